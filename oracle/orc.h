/*
 * orc.h -- CPU oracle for the bifurcation-analysis hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's per-trajectory algorithms (KiShiVi/CUDABifurcation3.0,
 * Library/cudaLibrary.cu and the driver arithmetic of Library/hostLibrary.cu).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline.  The product
 * (libbifhip.so) never links or calls it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no -march),
 * so every double operation is one IEEE-754 binary64 rounding in the
 * reference's source order, with no FMA contraction.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_step                 exact rational evaluation of the sequential
 *                            Euler step (a simultaneous-Euler step fails it)
 *   orc_loop                 record-before-step / divergence KATs
 *   orc_value_by_idx         closed form a_i = s + i*(f-s)/(n-1)
 *   orc_peak_finder          reference KAT {0,1,3,3,2,0,1,4,1,...}; brute-force
 *                            textbook comparison on plateau-free series
 *   orc_distance             glibc hypotf == (float)sqrt(double) check
 *   orc_dbscan               reference KATs ({1,2}->1, {1..8}->7, ...); Rossler
 *                            cluster counts at c=2.5,3.5,4.0,5.7,9.0
 *   orc_kde                  the values it compares (orc_kde_sums) against the
 *                            KDE definition; the count against a run-length
 *                            statement of the rule on exact-plateau / exact-
 *                            zero fuzz; hand cases for each rule; planted
 *                            mistakes caught (tests/test_oracle_pins.py,
 *                            tests/test_oracle_mutants.py)
 *   orc_gram_schmidt         orthonormality, |det| = prod(norms), numpy QR
 *   orc_lle_cell             linear model: LLE = log|1+h*lambda|/h
 *   orc_ls_cell              linear model spectrum; volume-contraction sum;
 *                            reference directions by one renormalisation of
 *                            the diagonal linear model against numpy QR
 *   orc_ls_var_cell          agreement with the finite-difference spectrum
 *   memristive orc_step      first-order convergence to its ODE (solve_ivp)
 *                            for explicit shares 1, 1/2, 0; float state
 *   orc_ostrovsky_cells      signs over the reference Rossler counts;
 *                            peak-free cells follow the last sample
 *   orc_denis                linear-model closed form per sample
 *   orc_bif_cell32 (FP32)    each step against the FP64 step from the same
 *                            float state (relative error at float rounding;
 *                            a dropped term or sign fails); linear model
 *                            x_n = x_0 * (float)(1+h*a)^n rounded per step;
 *                            period-1 Rossler peak counts equal the FP64 ones
 */
#ifndef BIFHIP_ORACLE_H
#define BIFHIP_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model ids shared with the product (same numbering). */
#define ORC_MODEL_ROSSLER 0   /* sequential explicit Euler, cudaLibrary.cu:15-17 */
#define ORC_MODEL_LINEAR  1   /* x_i += h*(a_i*x_i): closed-form test model     */
#define ORC_MODEL_MEMRISTIVE 2 /* five-variable system, cudaLibrary.cu:19-43    */

int  orc_model_dim(int model);
/* swept entries go into the initial conditions (1) or the parameters (0) */
void orc_set_sweep_ic(int on);
void orc_step(int model, double *x, const double *a, double h);

int  orc_loop(int model, double *x, const double *values, double h,
              int nIter, int preScaller, int writableVar, double maxValue,
              double *data, long start, int writeStep);

double orc_value_by_idx(int idx, int nPts, double startRange,
                        double finishRange, int valueNumber);

int  orc_peak_finder(double *data, long start, int amountOfPoints,
                     double *outPeaks, double *timeOfPeaks, double h);

int  orc_kde(const double *data, long start, int amountOfPoints,
             int maxAmountOfPeaks, int kdeSampling, double I1, double I2,
             double smoothH);

int  orc_kde_sums(const double *data, int amountOfPoints, int kdeSampling, double I1,
                  double I2, double smoothH, double *out, int cap);

double orc_distance(double x1, double y1, double x2, double y2);

int  orc_dbscan(const double *data, const double *intervals, double *helpful,
                long start, int amountOfPeaks, int sizeOfHelpful,
                int maxAmountOfPeaks, double eps);

void orc_gram_schmidt(const double *a, double *b, int n, double *denoms);

/* One cell of calculateDiscreteModelCUDA (cudaLibrary.cu:78-117).
 * Writes N recorded samples to series; returns 0 on divergence, 1 otherwise.
 * cellIdx is the global flattened grid index. */
int  orc_bif_cell(int model, int nPts, long cellIdx, int dimension,
                  const double *ranges, const int *mutIdx, const double *ic,
                  int nIC, const double *values, int nVal, double h,
                  int skip, int N, int preScaller, int writableVar,
                  double maxValue, double *series);

/* Perturbation-direction modes. */
#define ORC_PERT_REFERENCE 0  /* 0.5*sinf(...) of the chunk-local index */
#define ORC_PERT_AXIS_X    1  /* unit vector along x[0] */

double orc_lle_cell(int model, int nPts, long cellIdx, long localIdx,
                    int dimension, const double *ranges, const int *mutIdx,
                    const double *ic, int nIC, const double *values, int nVal,
                    double NT, double tMax, double h, double eps, int skip,
                    double maxValue, int pertMode);

void orc_ls_cell(int model, int nPts, long cellIdx, long localIdx,
                 int dimension, const double *ranges, const int *mutIdx,
                 const double *ic, int nIC, const double *values, int nVal,
                 double NT, double tMax, double h, double eps, int skip,
                 double maxValue, int pertMode, double *out);

/* Variational (tangent-map) spectrum of the Rossler step in FP64 with the
 * reference's Gram-Schmidt every nNT steps; out[3] = exponents.  Oracle of
 * the FP32 variational tile kernels.  Pinned by: the finite-difference LS
 * above (same cells, small eps) and the volume-contraction sum. */
void orc_ls_var_cell(double a, double b, double c, const double *ic, double h,
                     int skip, int nNT, int nBlocks, double *out);
void orc_ls_var_cells(const double *a, const double *b, const double *c,
                      const double *ic, double h, int skip, int nNT,
                      const int *nBlocks, long n, double *out);

/* ---- map-level drivers (OpenMP over cells; in-memory results) ---- */

/* bifurcation1D (hostLibrary.cu:7-158): counts[nPts], peaks[nPts*cap]. */
void orc_bifurcation1d(int model, double tMax, int nPts, double h, int nIC,
                       const double *ic, const double *ranges,
                       const int *mutIdx, int writableVar, double maxValue,
                       double transientTime, const double *values, int nVal,
                       int preScaller, int *counts, double *peaks, int cap);

/* Per-cell peak counts and peaks for a 2D (or 1D) grid slice [c0, c0+nc). */
void orc_peaks_cells(int model, int dimension, double tMax, int nPts, double h,
                     int nIC, const double *ic, const double *ranges,
                     const int *mutIdx, int writableVar, double maxValue,
                     double transientTime, const double *values, int nVal,
                     int preScaller, long c0, long nc, int *counts,
                     double *peaks, double *intervals, int cap);

/* bifurcation2DDBSCAN (hostLibrary.cu:571-744) over cells [c0, c0+nc). */
void orc_dbscan_cells(int model, double tMax, int nPts, double h, int nIC,
                      const double *ic, const double *ranges,
                      const int *mutIdx, int writableVar, double maxValue,
                      int maxAmountOfPeaks, double transientTime,
                      const double *values, int nVal, int preScaller,
                      double eps, long c0, long nc, int *out);

/* FP32 form of the peak maps (the library's FP32 precision; no reference
 * counterpart): float state with the fused FP32 step, samples widened to
 * double.  orc_bif_cell32 returns 1, 0 (diverged) or -1 (model or nIC not
 * supported: Rossler and linear, nIC = 3). */
int  orc_bif_cell32(int model, int nPts, long cellIdx, int dimension,
                    const double *ranges, const int *mutIdx, const double *ic,
                    int nIC, const double *values, int nVal, double h, int skip,
                    int N, int preScaller, int writableVar, double maxValue,
                    double *series);
void orc_peaks_cells32(int model, int dimension, double tMax, int nPts,
                       double h, int nIC, const double *ic,
                       const double *ranges, const int *mutIdx,
                       int writableVar, double maxValue, double transientTime,
                       const double *values, int nVal, int preScaller,
                       long c0, long nc, int *counts, double *peaks,
                       double *intervals, int cap);
void orc_dbscan_cells32(int model, double tMax, int nPts, double h, int nIC,
                        const double *ic, const double *ranges,
                        const int *mutIdx, int writableVar, double maxValue,
                        int maxAmountOfPeaks, double transientTime,
                        const double *values, int nVal, int preScaller,
                        double eps, long c0, long nc, int *out);

/* bifurcation2DKDE (hostLibrary.cu:397-567) over cells [c0, c0+nc). */
/* calculationOfPeriodicityByOstrovsky cells (initial-condition sweep) */
void orc_ostrovsky_cells(int model, double tMax, int nPts, double h, int nIC,
                         const double *ic, const double *ranges,
                         const int *mutIdx, int writableVar, double maxValue,
                         int maxAmountOfPeaks, double transientTime,
                         const double *values, int nVal, int preScaller,
                         double eps, double threshold, long c0, long nc,
                         double *out);
/* modelingOneSystemDenis series, out[nOut] (nOut = (int)(tMax/h)) */
void orc_denis(int model, double tMax, double h, double hS, int nIC,
               const double *ic, int writableVar, const double *values,
               int nVal, double *out, long nOut);
void orc_kde_cells(int model, double tMax, int nPts, double h, int nIC,
                   const double *ic, const double *ranges, const int *mutIdx,
                   int writableVar, double maxValue, int maxAmountOfPeaks,
                   double transientTime, const double *values, int nVal,
                   int preScaller, int kdeSampling, double I1, double I2,
                   double smoothH, long c0, long nc, int *out);

/* LLE1D / LLE2D (hostLibrary.cu:1270-1393, 1524-1663) over [c0, c0+nc).
 * dirChunk: chunk size whose local index seeds the reference direction. */
void orc_lle_cells(int model, int dimension, double tMax, double NT, int nPts,
                   double h, double eps, const double *ic, int nIC,
                   const double *ranges, const int *mutIdx, double maxValue,
                   double transientTime, const double *values, int nVal,
                   int pertMode, long dirChunk, long c0, long nc, double *out);

/* LS1D / LS2D (hostLibrary.cu:1802-2082) over [c0, c0+nc); out[nc*nIC]. */
void orc_ls_cells(int model, int dimension, double tMax, double NT, int nPts,
                  double h, double eps, const double *ic, int nIC,
                  const double *ranges, const int *mutIdx, double maxValue,
                  double transientTime, const double *values, int nVal,
                  int pertMode, long dirChunk, long c0, long nc, double *out);

/* Number of OpenMP threads the map drivers use. */
int  orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
