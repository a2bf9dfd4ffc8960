/*
 * bifhip.h -- C ABI of libbifhip.so, the MI355X (gfx950) bifurcation-analysis
 * library.  Plain C types only; no HIP or torch types cross this boundary.
 *
 * Every driver keeps the parameter order, types and meaning of the
 * reference's host function it replaces (KiShiVi/CUDABifurcation3.0,
 * Library/hostLibrary.cuh), returns a status instead of calling exit(),
 * and writes the same CSV files to the configured output path.
 *
 * Status: 0 success; > 0 a hipError_t value; < 0 one of BIFHIP_E_*.
 * bifhip_last_error() describes the last failure.
 *
 * Threading: one context per process, calls are synchronous on one HIP
 * stream and not re-entrant (as in the reference, which also shares one
 * output file).
 */
#ifndef BIFHIP_H
#define BIFHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BIFHIP_OK 0
#define BIFHIP_E_ARG (-1)      /* invalid argument                          */
#define BIFHIP_E_FILE (-2)     /* output file could not be opened/written   */
#define BIFHIP_E_NODEV (-3)    /* no usable gfx950 device                   */
#define BIFHIP_E_LIMIT (-4)    /* size beyond what the library supports     */

#define BIFHIP_MODEL_ROSSLER 0 /* Library/cudaLibrary.cu:15-17 (values {a,b,c}) */
#define BIFHIP_MODEL_LINEAR 1  /* x_i += h*a_i*x_i: closed-form test model      */
#define BIFHIP_MODEL_MEMRISTIVE 2 /* Library/cudaLibrary.cu:19-43: five variables,
                                     values a0..a7 (a7 = explicit share of the step);
                                     FP64 drivers only (no fused map, tiles, FP32) */

#define BIFHIP_PREC_FP64 0     /* reference operation order, no contraction     */
#define BIFHIP_PREC_FP32 1     /* FP32 bulk integration, FP64 renorm/log        */

#define BIFHIP_SWEEP_PARAMETERS 0         /* indicesOfMutVars index values[]        */
#define BIFHIP_SWEEP_INITIAL_CONDITIONS 1 /* ... index initialConditions[]         */

#define BIFHIP_PERT_REFERENCE 0 /* 0.5*sinf(...) of the chunk-local index      */
#define BIFHIP_PERT_AXIS_X 1    /* unit perturbation along x[0]                */

/* ------------------------------------------------------------ context */

/* Output file path.  Replaces the compile-time OUT_FILE_PATH macro
 * (Library/hostLibrary.cu:3).  Default "mat.csv". */
int bifhip_set_output_path(const char* path);
const char* bifhip_get_output_path(void);
/* Model plug-in.  Replaces editing calculateDiscreteModel
 * (Library/cudaLibrary.cu:2-44).  Default BIFHIP_MODEL_ROSSLER. */
int bifhip_set_model(int model);
/* Integration precision.  Default BIFHIP_PREC_FP64 (the reference's).
 * BIFHIP_PREC_FP32 applies to the LLE drivers (FP32 state and separation,
 * FP64 norms and log sums), to the 1D diagram and to the 2D DBSCAN / KDE
 * maps (FP32 state with the fused FP32 step, samples widened to double for
 * the same peak finder and clustering), for the Rossler and linear models.
 * No reference counterpart: FP32 results are judged by their agreement
 * with the FP64 ones.  The surrogate's re-solves ignore it. */
int bifhip_set_precision(int precision);
/* LLE/LS initial perturbation directions.  Default BIFHIP_PERT_REFERENCE
 * (Library/cudaLibrary.cu:720-732, :1025-1040). */
int bifhip_set_perturbation(int mode);
/* Chunk whose local index seeds the reference perturbation directions
 * (the reference's nPtsLimiter): -1 = the driver's own (200000 in LLE2D,
 * Library/hostLibrary.cu:1552; one chunk elsewhere), 0 = none. */
int bifhip_set_direction_chunk(long long cells);
/* Upper bound on cells per kernel launch (0 = automatic).  Results do not
 * depend on it; tests use it to force multi-launch execution. */
int bifhip_set_max_cells_per_launch(long long cells);
/* Peak slots per cell kept in HBM by the 2D peak-count drivers (>= 2;
 * default 256).  Cells with more peaks are recomputed exactly. */
int bifhip_set_peak_capacity(int cap);
/* Clamp applied to 2D DBSCAN cluster counts (0 = none, the reference). */
int bifhip_set_cluster_clamp(int clamp);
/* Progress messages on stdout like the reference's DEBUG build. */
int bifhip_set_verbose(int on);
/* HIP events around every kernel launch for bifhip_kernel_time(). */
int bifhip_set_timing(int on);
/* Accumulated event time (ms) and launch count of a kernel family
 * ("bif_peaks", "dbscan", "kde", "lle", "ls", ...) since the last reset;
 * kernel == NULL resets all counters. */
int bifhip_kernel_time(const char* kernel, double* total_ms, long long* launches);
/* Cells the timed launches of a kernel family covered since the last reset
 * (families that count them: "bif_peaks", "bif_peaks2" -- the two-cell FP64
 * peak kernel --, "lle_peaks", "lle", "lle32", "dbscan"). */
int bifhip_kernel_cells(const char* kernel, double* cells);
/* Release cached device workspaces. */
int bifhip_release(void);
const char* bifhip_last_error(void);
/* Number of visible HIP devices (0 when none). */
int bifhip_device_count(void);
/* Library version string. */
const char* bifhip_version(void);

/* Device memory on the library's HIP runtime, for callers that bring no
 * GPU framework of their own (no counterpart in the reference, whose
 * drivers only take host arrays). */
int bifhip_device_malloc(void** ptr, size_t bytes);
int bifhip_device_free(void* ptr);
int bifhip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int bifhip_memcpy_d2h(void* dst, const void* src, size_t bytes);
int bifhip_memcpy_d2d(void* dst, const void* src, size_t bytes);
int bifhip_device_memset(void* ptr, int value, size_t bytes);
int bifhip_synchronize(void);

/* ------------------------------------------------------ drivers (files) */

/* Library/hostLibrary.cuh:29-42 (bifurcation1D).  Writes "param, peak\n"
 * lines, 12 significant digits. */
int bifhip_bifurcation1D(double tMax, int nPts, double h, int amountOfInitialConditions,
                         const double* initialConditions, const double* ranges,
                         const int* indicesOfMutVars, int writableVar, double maxValue,
                         double transientTime, const double* values, int amountOfValues,
                         int preScaller);

/* Library/hostLibrary.cuh:129-147 (bifurcation2DKDE). */
int bifhip_bifurcation2DKDE(double tMax, int nPts, double h, int amountOfInitialConditions,
                            const double* initialConditions, const double* ranges,
                            const int* indicesOfMutVars, int writableVar, double maxValue,
                            int maxAmountOfPeaks, double transientTime, const double* values,
                            int amountOfValues, int preScaller, int kdeSampling,
                            double kdeSamplesInterval1, double kdeSamplesInterval2,
                            double kdeSamplesSmooth);

/* Library/hostLibrary.cuh:171-186 (bifurcation2DDBSCAN). */
int bifhip_bifurcation2DDBSCAN(double tMax, int nPts, double h, int amountOfInitialConditions,
                               const double* initialConditions, const double* ranges,
                               const int* indicesOfMutVars, int writableVar, double maxValue,
                               int maxAmountOfPeaks, double transientTime, const double* values,
                               int amountOfValues, int preScaller, double eps);

/* Library/hostLibrary.cuh:334-348 (LLE1D). */
int bifhip_LLE1D(double tMax, double NT, int nPts, double h, double eps,
                 const double* initialConditions, int amountOfInitialConditions,
                 const double* ranges, const int* indicesOfMutVars, int writableVar,
                 double maxValue, double transientTime, const double* values,
                 int amountOfValues);

/* Library/hostLibrary.cuh:408-422 (LLE2D). */
int bifhip_LLE2D(double tMax, double NT, int nPts, double h, double eps,
                 const double* initialConditions, int amountOfInitialConditions,
                 const double* ranges, const int* indicesOfMutVars, int writableVar,
                 double maxValue, double transientTime, const double* values,
                 int amountOfValues);

/* Library/hostLibrary.cuh:482-496 (LS1D). */
int bifhip_LS1D(double tMax, double NT, int nPts, double h, double eps,
                const double* initialConditions, int amountOfInitialConditions,
                const double* ranges, const int* indicesOfMutVars, int writableVar,
                double maxValue, double transientTime, const double* values,
                int amountOfValues);

/* Library/hostLibrary.cuh:520-534 (LS2D).  Writes <path>1.csv .. <path>n.csv. */
int bifhip_LS2D(double tMax, double NT, int nPts, double h, double eps,
                const double* initialConditions, int amountOfInitialConditions,
                const double* ranges, const int* indicesOfMutVars, int writableVar,
                double maxValue, double transientTime, const double* values,
                int amountOfValues);

/* Drivers sweeping initial conditions: the same computations with the
 * swept values written into the initial state (indicesOfMutVars index
 * initialConditions), as calculateDiscreteModelICCUDA and LLEKernelICCUDA
 * (Library/cudaLibrary.cu:165-207, 789-908); LLE directions use
 * LLEKernelICCUDA's formula of the global index. */

/* Library/hostLibrary.cuh:89-102 (bifurcation1DIC), bifurcation1D's format. */
int bifhip_bifurcation1DIC(double tMax, int nPts, double h, int amountOfInitialConditions,
                           const double* initialConditions, const double* ranges,
                           const int* indicesOfMutVars, int writableVar, double maxValue,
                           double transientTime, const double* values, int amountOfValues,
                           int preScaller);

/* Library/hostLibrary.cuh:254-272 (bifurcation2DKDEIC): no range header. */
int bifhip_bifurcation2DKDEIC(double tMax, int nPts, double h, int amountOfInitialConditions,
                              const double* initialConditions, const double* ranges,
                              const int* indicesOfMutVars, int writableVar, double maxValue,
                              int maxAmountOfPeaks, double transientTime, const double* values,
                              int amountOfValues, int preScaller, int kdeSampling,
                              double kdeSamplesInterval1, double kdeSamplesInterval2,
                              double kdeSamplesSmooth);

/* Library/hostLibrary.cuh:296-311 (bifurcation2DDBSCANIC). */
int bifhip_bifurcation2DDBSCANIC(double tMax, int nPts, double h, int amountOfInitialConditions,
                                 const double* initialConditions, const double* ranges,
                                 const int* indicesOfMutVars, int writableVar, double maxValue,
                                 int maxAmountOfPeaks, double transientTime,
                                 const double* values, int amountOfValues, int preScaller,
                                 double eps);

/* Library/hostLibrary.cuh:371-385 (LLE1DIC), LLE1D's format. */
int bifhip_LLE1DIC(double tMax, double NT, int nPts, double h, double eps,
                   const double* initialConditions, int amountOfInitialConditions,
                   const double* ranges, const int* indicesOfMutVars, int writableVar,
                   double maxValue, double transientTime, const double* values,
                   int amountOfValues);

/* Library/hostLibrary.cuh:445-459 (LLE2DIC): no range header. */
int bifhip_LLE2DIC(double tMax, double NT, int nPts, double h, double eps,
                   const double* initialConditions, int amountOfInitialConditions,
                   const double* ranges, const int* indicesOfMutVars, int writableVar,
                   double maxValue, double transientTime, const double* values,
                   int amountOfValues);

/* Library/hostLibrary.cuh:211-227 (calculationOfPeriodicityByOstrovsky): the
 * periodicity classifier over a sweep of initial conditions (the reference
 * uses calculateDiscreteModelICCUDA).  Per cell: the DBSCAN count of the
 * peaks, negated unless every kept peak after the first exceeds
 * ostrovskyThreshold; +-1 by the last sample's side of the threshold when
 * at most one peak is kept; 0 when the peaks lie on both sides; halved to
 * +-0.5 when no peak rises above its predecessor
 * (Library/cudaLibrary.cu:286-352, 628-662).  A diverged trajectory sets
 * only the few-peaks flag (the reference leaves the others unwritten).
 * Writes the 2D map layout with the range header. */
int bifhip_calculationOfPeriodicityByOstrovsky(double tMax, int nPts, double h,
                                               int amountOfInitialConditions,
                                               const double* initialConditions,
                                               const double* ranges,
                                               const int* indicesOfMutVars, int writableVar,
                                               double maxValue, int maxAmountOfPeaks,
                                               double transientTime, const double* values,
                                               int amountOfValues, int preScaller, double eps,
                                               double ostrovskyThreshold);
/* The same map into out[nPts * nPts] (host memory). */
int bifhip_ostrovsky_to_buffer(double tMax, int nPts, double h, int amountOfInitialConditions,
                               const double* initialConditions, const double* ranges,
                               const int* indicesOfMutVars, int writableVar, double maxValue,
                               int maxAmountOfPeaks, double transientTime, const double* values,
                               int amountOfValues, int preScaller, double eps,
                               double ostrovskyThreshold, double* out);

/* Library/hostLibrary.cuh:59-67 (modelingOneSystemDenis): one system from
 * initialConditions sampled every h over tMax by (int)(hSpecial/h)
 * trajectories of step hSpecial, trajectory k started k steps of h in
 * (calculateDiscreteModelDenisCUDA, Library/cudaLibrary.cu:121-158).
 * Writes "t, x[writableVar]" lines, 12 significant digits. */
int bifhip_modelingOneSystemDenis(double tMax, double h, double hSpecial,
                                  int amountOfInitialConditions,
                                  const double* initialConditions, int writableVar,
                                  const double* values, int amountOfValues);
/* The same series into out[nOut], nOut >= (int)(tMax/h); samples no
 * trajectory reaches are 0. */
int bifhip_denis_to_buffer(double tMax, double h, double hSpecial, int amountOfInitialConditions,
                           const double* initialConditions, int writableVar,
                           const double* values, int amountOfValues, double* out,
                           long long nOut);

/* Sweep target of the buffer and device entry points below
 * (BIFHIP_SWEEP_PARAMETERS by default); the named file drivers fix their
 * own target. */
int bifhip_set_sweep(int target);

/* ------------------------------------------- drivers (host buffers) */

/* Same computations, results in caller-owned host memory instead of a file.
 * counts[nPts]; peaks[nPts * peakStride] (row k holds counts[k] peaks;
 * peakStride >= (int)(tMax/h/preScaller)/2 + 1 always suffices). */
int bifhip_bifurcation1D_to_buffer(double tMax, int nPts, double h,
                                   int amountOfInitialConditions,
                                   const double* initialConditions, const double* ranges,
                                   const int* indicesOfMutVars, int writableVar,
                                   double maxValue, double transientTime, const double* values,
                                   int amountOfValues, int preScaller, int* counts,
                                   double* peaks, int peakStride);

/* out[nPts*nPts], row-major with the first swept parameter fastest. */
int bifhip_bifurcation2DKDE_to_buffer(double tMax, int nPts, double h,
                                      int amountOfInitialConditions,
                                      const double* initialConditions, const double* ranges,
                                      const int* indicesOfMutVars, int writableVar,
                                      double maxValue, int maxAmountOfPeaks,
                                      double transientTime, const double* values,
                                      int amountOfValues, int preScaller, int kdeSampling,
                                      double kdeSamplesInterval1, double kdeSamplesInterval2,
                                      double kdeSamplesSmooth, int* out);

int bifhip_bifurcation2DDBSCAN_to_buffer(double tMax, int nPts, double h,
                                         int amountOfInitialConditions,
                                         const double* initialConditions, const double* ranges,
                                         const int* indicesOfMutVars, int writableVar,
                                         double maxValue, int maxAmountOfPeaks,
                                         double transientTime, const double* values,
                                         int amountOfValues, int preScaller, double eps,
                                         int* out);

/* dimension 1 (LLE1D) or 2 (LLE2D); out[nPts^dimension]. */
int bifhip_LLE_to_buffer(int dimension, double tMax, double NT, int nPts, double h, double eps,
                         const double* initialConditions, int amountOfInitialConditions,
                         const double* ranges, const int* indicesOfMutVars, int writableVar,
                         double maxValue, double transientTime, const double* values,
                         int amountOfValues, double* out);

/* dimension 1 (LS1D) or 2 (LS2D); out[nPts^dimension * nIC], cell-major. */
int bifhip_LS_to_buffer(int dimension, double tMax, double NT, int nPts, double h, double eps,
                        const double* initialConditions, int amountOfInitialConditions,
                        const double* ranges, const int* indicesOfMutVars, int writableVar,
                        double maxValue, double transientTime, const double* values,
                        int amountOfValues, double* out);

/* ----------------------------------------- drivers (device buffers) */

/* Device-resident outputs for composition and benchmarking: d_out is a
 * device pointer the caller owns; nothing is copied back. */
int bifhip_bifurcation2DDBSCAN_device(double tMax, int nPts, double h,
                                      int amountOfInitialConditions,
                                      const double* initialConditions, const double* ranges,
                                      const int* indicesOfMutVars, int writableVar,
                                      double maxValue, int maxAmountOfPeaks,
                                      double transientTime, const double* values,
                                      int amountOfValues, int preScaller, double eps,
                                      int* d_out);
int bifhip_bifurcation2DKDE_device(double tMax, int nPts, double h,
                                   int amountOfInitialConditions,
                                   const double* initialConditions, const double* ranges,
                                   const int* indicesOfMutVars, int writableVar,
                                   double maxValue, int maxAmountOfPeaks, double transientTime,
                                   const double* values, int amountOfValues, int preScaller,
                                   int kdeSampling, double kdeSamplesInterval1,
                                   double kdeSamplesInterval2, double kdeSamplesSmooth,
                                   int* d_out);

int bifhip_LLE_device(int dimension, double tMax, double NT, int nPts, double h, double eps,
                      const double* initialConditions, int amountOfInitialConditions,
                      const double* ranges, const int* indicesOfMutVars, int writableVar,
                      double maxValue, double transientTime, const double* values,
                      int amountOfValues, double* d_out);

/* LLE map and DBSCAN peak-count map of one 2D grid in one pass.  No single
 * reference function: it produces exactly what LLE2D
 * (Library/hostLibrary.cuh:408-422) and bifurcation2DDBSCAN
 * (Library/hostLibrary.cuh:171-186) produce for the same grid, because the
 * LLE's reference trajectory passes through every state the peak window
 * records.  Requires preScaller 1, no divergence check, and the peak window
 * (peakTransient .. peakTransient + peakTMax) inside the LLE trajectory
 * (lleTransient + lleTMax).  d_peak[nPts^2] int, d_lle[nPts^2] double. */
int bifhip_lle_peak_map_device(int nPts, double h, int amountOfInitialConditions,
                               const double* initialConditions, const double* ranges,
                               const int* indicesOfMutVars, const double* values,
                               int amountOfValues, double peakTMax, double peakTransient,
                               int writableVar, int maxAmountOfPeaks, double dbscanEps,
                               double lleTMax, double NT, double lleTransient, double lleEps,
                               int* d_peak, double* d_lle);

/* ------------------------------------ reference kernels on device data */

/* peakFinderCUDA (Library/cudaLibrary.cu:270-284) with peakFinder's h
 * (Library/cudaLibrary.cuh:204, default 0.01): series d_data[nBlocks *
 * sizeOfBlock]; d_counts holds -1 for flagged blocks on entry and the
 * peak count on exit; peaks and intervals are written from each block's
 * start (d_outPeaks may equal d_data). */
int bifhip_peak_finder_device(const double* d_data, int sizeOfBlock, int amountOfBlocks,
                              int* d_counts, double* d_outPeaks, double* d_timeOfPeaks,
                              double h);

/* dbscanCUDA (Library/cudaLibrary.cu:609-624) over per-block peak rows of
 * stride sizeOfBlock: d_out[b] = dbscan(...) (-1 counts give 0). */
int bifhip_dbscan_device(const double* d_peaks, const double* d_intervals, int sizeOfBlock,
                         int amountOfBlocks, const int* d_counts, int maxAmountOfPeaks,
                         double eps, int* d_out);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): bifhip_dbscan_device with the same arguments and
 * results, whose scratch (overflow lists, counters, lane label rows) is a
 * fixed device array of BIFHIP_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; it calls no allocating routine.  Limits: amountOfBlocks <=
 * BIFHIP_NOALLOC_MAX_BLOCKS per call, and sizeOfBlock <=
 * BIFHIP_NOALLOC_MAX_ROW; beyond them it returns BIFHIP_E_LIMIT.  Calls
 * share the array of their GPU: calls that overlap on one GPU race. */
#define BIFHIP_NOALLOC_MAX_BLOCKS (1 << 20)
#define BIFHIP_NOALLOC_MAX_ROW (1 << 20)
#define BIFHIP_NOALLOC_SCRATCH_BYTES \
    (4LL * (64 + 2LL * BIFHIP_NOALLOC_MAX_BLOCKS + 2LL * BIFHIP_NOALLOC_MAX_ROW))
int bifhip_dbscan_device_noalloc(const double* d_peaks, const double* d_intervals,
                                 int sizeOfBlock, int amountOfBlocks, const int* d_counts,
                                 int maxAmountOfPeaks, double eps, int* d_out);

/* kdeCUDA (Library/cudaLibrary.cu:426-441). */
int bifhip_kde_device(const double* d_peaks, int sizeOfBlock, int amountOfBlocks,
                      const int* d_counts, int maxAmountOfPeaks, int kdeSampling,
                      double kdeSamplesInterval1, double kdeSamplesInterval2,
                      double kdeSamplesSmooth, int* d_out);

/* gramSchmidtProcess (Library/cudaLibrary.cu:931-957), batched: count
 * row-major n x n matrices, n <= 8; d_denominators may be NULL. */
int bifhip_gram_schmidt_device(const double* d_a, double* d_b, double* d_denominators, int n,
                               int count);

/* calculateDiscreteModelCUDA record pass (Library/cudaLibrary.cu:78-117):
 * the recorded series of every cell of a 1D/2D grid, d_data[cells * N]
 * with N = (int)(tMax/h/preScaller); d_flags[cell] = -1 on divergence. */
int bifhip_series_device(int dimension, double tMax, int nPts, double h,
                         int amountOfInitialConditions, const double* initialConditions,
                         const double* ranges, const int* indicesOfMutVars, int writableVar,
                         double maxValue, double transientTime, const double* values,
                         int amountOfValues, int preScaller, double* d_data, int* d_flags);

/* ------------------------------- Lyapunov spectrum on uneven tiles */

/* No counterpart in the reference (whose LS2D, Library/hostLibrary.cu:
 * 1933-2082, integrates one uniform length with finite-difference copies):
 * the full 3-exponent spectrum of tiles of tileW x tileH cells placed on an
 * nPts x nPts lattice over `ranges`, each tile with its own length, in one
 * persistent launch replayed from a captured HIP graph.  FP32 state and
 * tangent vectors (variational equation of the discrete step), FP64
 * Gram-Schmidt (the reference's, Library/cudaLibrary.cu:913-957) and log
 * accumulation every NT.  Tile tables are host arrays; d_out is device
 * memory [nTiles][tileW*tileH][3]. */

/* Seeded tile table: origins (half uniform, half clustered), lengths
 * uniform in [tMinTime, tMaxTime] as whole NT blocks. */
int bifhip_tiles_generate(int nTiles, int tileW, int tileH, int nPts,
                          unsigned long long seedOrigins, unsigned long long seedLengths,
                          double tMinTime, double tMaxTime, double NT, int* originI,
                          int* originJ, int* nBlocks);

int bifhip_ls_tiles_device(int nTiles, int tileW, int tileH, const int* originI,
                           const int* originJ, const int* nBlocks, int nPts, const double* ranges,
                           const int* indicesOfMutVars, const double* initialConditions,
                           int amountOfInitialConditions, const double* values,
                           int amountOfValues, double h, double NT, double transientTime,
                           double* d_out);

/* Same, after timing every (block size, cells per lane) candidate on a
 * sample of the queue and keeping the fastest as the tile configuration. */
int bifhip_ls_tiles_autotune(int nTiles, int tileW, int tileH, const int* originI,
                             const int* originJ, const int* nBlocks, int nPts,
                             const double* ranges, const int* indicesOfMutVars,
                             const double* initialConditions, int amountOfInitialConditions,
                             const double* values, int amountOfValues, double h, double NT,
                             double transientTime, double* d_out);

/* Tile configuration: block size 0 (autotune on next use), 64, 128, 256;
 * cells per lane 1 or 2; blocks per CU 0 = occupancy maximum. */
int bifhip_set_tile_config(int blockSize, int cellsPerLane, int blocksPerCU);
int bifhip_get_tile_config(int* blockSize, int* cellsPerLane, int* blocksPerCU);
/* Lanes per cell of the FP32 tile kernel: 1 (a lane carries the state and
 * the three tangent vectors) or 3 (a lane per tangent vector, each stepping
 * the state itself; vectors exchanged across lanes for Gram-Schmidt).
 * Results are bit-identical either way; the autotuner times both. */
int bifhip_set_tile_lanes(int lanesPerCell);
int bifhip_get_tile_lanes(int* lanesPerCell);
/* The last autotuner run: up to maxEntries rows of (block size, lanes per
 * cell, cells per lane, ms on the sample) to out[4*i..]; returns the number
 * of candidates timed. */
int bifhip_tile_autotune_log(double* out, int maxEntries);
/* State precision of the tile path: BIFHIP_PREC_FP32 (default, FP32 state
 * and tangent vectors, FP64 Gram-Schmidt) or BIFHIP_PREC_FP64 (FP64 state
 * and tangent vectors, the reference's Gram-Schmidt and log sum: the
 * reference against which the FP32 path's error is stated). */
int bifhip_set_tile_precision(int precision);
/* Replay the tile launches and the surrogate's training epochs from captured
 * HIP graphs (default 1); 0 launches the same kernels directly. */
int bifhip_set_graph(int on);
/* Surrogate-refined map: cells whose LLE regime is predicted with
 * confidence and that need only their peak-count class re-solve the peak
 * count alone (default 1); 0 re-solves every flagged cell in full. */
int bifhip_set_partial_resolve(int on);
/* Surrogate-refined map: cells whose chaos-logit confidence is below this
 * threshold (in [0.5, 1], default 0.975) have their LLE re-solved. */
int bifhip_set_chaos_confidence(double threshold);
/* Surrogate-refined map: cells re-solved for the regime only take their LLE
 * from the FP32 bulk kernel (FP32 state and separation, FP64 norms and log
 * sums, as bifhip_set_precision(BIFHIP_PREC_FP32) computes the LLE map);
 * those whose FP32 LLE lies within `margin` of chaosThreshold, or is NaN,
 * are re-solved again in FP64.  Default 0.02; a negative margin keeps
 * every regime re-solve in FP64.  The memristive model always uses FP64. */
int bifhip_set_lle32_margin(double margin);
/* Surrogate-refined map: what a cell's regime is when the map decides
 * which cells to re-solve.  No reference counterpart.
 *   BIFHIP_READING_CLASS_AND_CHAOS (default): the peak-count class
 *     clamp(count, 1, 16) and the chaos bit (LLE > chaosThreshold); cells
 *     unsure of either are re-solved for it;
 *   BIFHIP_READING_CHAOS: the chaos bit only; peak counts stay predicted
 *     except where a re-solved LLE changes the predicted regime;
 *   BIFHIP_READING_CLASS: the class only; LLE values stay predicted. */
/* Surrogate-refined map: a fine cell inside a quad of the coarse lattice
 * whose four corners (exact, from the coarse map) disagree on the chaos
 * bit, or whose corners the network predicts wrongly, has its LLE re-solved
 * unless its chaos confidence is >= threshold (in [0, 2]; 0 turns the rule
 * off, above 1 applies it to every such cell).  Default 0.  No reference
 * counterpart. */
int bifhip_set_quad_confidence(double threshold);
#define BIFHIP_READING_CLASS_AND_CHAOS 0
#define BIFHIP_READING_CHAOS 1
#define BIFHIP_READING_CLASS 2
int bifhip_set_regime_reading(int reading);

/* ------------------------------------------ surrogate-refined map */

/* No counterpart in the reference.  A small network (64 Fourier features of
 * (a, c) -> 3 x 128 ReLU -> LLE, 16 peak-count classes and a chaos logit),
 * trained on a coarse map the plain FP64 method produces and run with BF16
 * matrix-core products and FP32 accumulation, predicts the fine map: the
 * class from the class logits, the LLE regime (chaotic when LLE >
 * chaosThreshold) from the chaos logit, the LLE value from the regression
 * head, kept on the predicted side of chaosThreshold.  A cell needs its
 * class re-solved when the class softmax confidence is below confThreshold
 * or a 4-neighbour has another predicted class, and its regime when the
 * chaos confidence is below bifhip_set_chaos_confidence's threshold (0.975
 * by default) or a 4-neighbour has another predicted regime.  Cells that
 * need their class and regime are re-solved by the plain FP64 fused kernel
 * (bifhip_lle_peak_map_device semantics).  With partial re-solves on
 * (bifhip_set_partial_resolve, default), cells that need only their class
 * get the plain FP64 DBSCAN peak count (bifurcation2DDBSCAN semantics) and
 * keep the predicted LLE, and cells that need only their regime get the
 * plain LLE (LLE2D semantics; FP32 bulk away from the threshold, FP64
 * near it, bifhip_set_lle32_margin) and keep the predicted class unless
 * the re-solved regime differs from the predicted one, in which case they
 * get the FP64 peak count too; with them off, every flagged cell is
 * re-solved in full.
 * The re-solves score the surrogate; a validation sample -- each remaining
 * cell drawn with probability validationFraction by a hash of its index and
 * the seed -- is re-solved in full to score it there.  Outputs: d_peak[nFine^2] (re-solved cells: DBSCAN count;
 * others: predicted class + 1 in 1..16), d_lle[nFine^2],
 * d_resolved[nFine^2] (1 re-solved in full, 2 peak count only, 3 validation
 * sample re-solved in full, 4 LLE only, 5 LLE then peak count, 6 and 7 as
 * 4 and 5 with the LLE from the FP32 bulk kernel, 0 predicted; may be
 * NULL),
 * stats[BIFHIP_SSTAT_COUNT]. */
#define BIFHIP_SSTAT_MS_TOTAL 0
#define BIFHIP_SSTAT_MS_COARSE 1
#define BIFHIP_SSTAT_MS_TRAIN 2
#define BIFHIP_SSTAT_MS_INFER 3
#define BIFHIP_SSTAT_MS_RESOLVE 4
#define BIFHIP_SSTAT_RESOLVED 5            /* cells re-solved in full            */
#define BIFHIP_SSTAT_RESOLVED_FRAC 6
#define BIFHIP_SSTAT_VALIDATION 7          /* unflagged cells re-solved to score */
#define BIFHIP_SSTAT_FLAGGED_CLASS_ACC 8   /* surrogate vs FP64 on flagged cells */
#define BIFHIP_SSTAT_FLAGGED_REGIME_ACC 9
#define BIFHIP_SSTAT_FLAGGED_LLE_MAE 10
#define BIFHIP_SSTAT_VALID_CLASS_ACC 11    /* surrogate vs FP64 on the sample    */
#define BIFHIP_SSTAT_VALID_REGIME_ACC 12
#define BIFHIP_SSTAT_VALID_LLE_MAE 13
#define BIFHIP_SSTAT_LOSS_CE 14            /* last-epoch training losses         */
#define BIFHIP_SSTAT_LOSS_MSE 15
#define BIFHIP_SSTAT_LLE_MEAN 16
#define BIFHIP_SSTAT_LLE_SD 17
#define BIFHIP_SSTAT_PEAK_ONLY 18          /* cells re-solved for the class only */
#define BIFHIP_SSTAT_PEAK_ONLY_CLASS_ACC 19
#define BIFHIP_SSTAT_MS_FLAG 20            /* parts of MS_RESOLVE: flagging and   */
#define BIFHIP_SSTAT_MS_SOLVE 21           /* sampling, the FP64 re-solves, and   */
#define BIFHIP_SSTAT_MS_SCORE 22           /* scoring + writing the results       */
#define BIFHIP_SSTAT_LLE_ONLY 23           /* cells re-solved for the regime only */
#define BIFHIP_SSTAT_LLE_ONLY_REGIME_ACC 24 /* surrogate chaos regime right there */
#define BIFHIP_SSTAT_LLE_ONLY_RECOUNTED 25 /* of those, regime changed: peak count */
#define BIFHIP_SSTAT_LLE_ONLY_FP32 26      /* of those, LLE kept from FP32       */
#define BIFHIP_SSTAT_COUNT 27

int bifhip_surrogate_map_device(int nFine, int nCoarse, double h, int amountOfInitialConditions,
                                const double* initialConditions, const double* ranges,
                                const int* indicesOfMutVars, const double* values,
                                int amountOfValues, double peakTMax, double peakTransient,
                                int writableVar, int maxAmountOfPeaks, double dbscanEps,
                                double lleTMax, double NT, double lleTransient, double lleEps,
                                int epochs, int batch, double lr, unsigned long long seed,
                                double confThreshold, double chaosThreshold,
                                double validationFraction, int* d_peak, double* d_lle,
                                unsigned char* d_resolved, double* stats);

/* The network's pieces, for tests and custom pipelines (device buffers). */
int bifhip_mlp_param_count(void);
/* Training on a labelled nLat x nLat lattice: d_cls class 0..15, d_lle
 * standardised LLE (NaN = no target), chaosStd the chaos threshold in the
 * same units; loss_out[3] = last-epoch mean CE, MSE, chaos BCE. */
int bifhip_mlp_train_device(int nLat, const int* d_cls, const float* d_lle, int epochs,
                            int batch, double lr, double wLle, unsigned long long seed,
                            double fourierSigma, double chaosStd, int useGraph, float* loss_out);
/* Inference of an nLat x nLat lattice: class, chaos regime, confidence, LLE.
 * The regime comes from the chaos logit; the LLE value from the regression
 * head, kept on that side of chaosThr.  With d_confChaos NULL, d_conf =
 * min(class confidence, chaos confidence); else d_conf = class confidence
 * and d_confChaos = chaos confidence. */
int bifhip_mlp_infer_device(int nLat, double lleMean, double lleSd, double chaosThr,
                            unsigned char* d_cls, unsigned char* d_chaos, float* d_conf,
                            float* d_confChaos, float* d_lle);
int bifhip_mlp_get_params(float* params, int nParams, float* fourier, int nFourier);

/* No counterpart in the reference: the optimiser kernel of the surrogate's
 * training on caller-given state, for checks against a CPU Adam.  `steps`
 * updates (step t = 1..steps) of n FP32 parameters, params/m/v host arrays
 * updated in place; the gradient of step k is `splits` partials
 * grads[(k*splits + z)*n + i], added in order z; keep[i] == 0 holds
 * parameter i at zero; wbf (may be NULL) receives the BF16 bit patterns of
 * the parameters after the last step. */
int bifhip_mlp_adam(int n, float* params, float* m, float* v, const float* grads, int splits,
                    int steps, const unsigned char* keep, double lr, double b1, double b2,
                    double eps, unsigned short* wbf);
/* One forward + backward pass (host arrays): cells are flattened indices on
 * an nLat x nLat lattice, cls/lle their targets; grads in parameter
 * layout W1[128x64] b1 W2[128x128] b2 W3[128x128] b3 Wh[32x128] bh[32];
 * losses[3]; outputs[rows x 32] head values. */
int bifhip_mlp_gradient(int nLat, const float* params, const float* fourier, const int* cells,
                        const int* cls, const float* lle, int rows, double wLle, double chaosStd,
                        float* grads, float* losses, float* outputs);

/* ------------------------------------------------ CSV writers (host) */

/* The reference's file formats, usable without a GPU.
 * 2D maps: header "r0 r1\nr2 r3\n" (when header != 0), values joined by
 * ", " with a line break after every nPts values, the ", " written before
 * the break (Library/hostLibrary.cu:651-655, :708-725); precision = the
 * stream precision (6 default, 12 for bifurcation2DKDE). */
int bifhip_write_map_int(const char* path, int nPts, const double* ranges, int header,
                         int precision, const int* values, long long count);
int bifhip_write_map_double(const char* path, int nPts, const double* ranges, int header,
                            int precision, const double* values, long long count,
                            int stride, int offset);
/* bifurcation1D lines "param, peak\n" (Library/hostLibrary.cu:125-133). */
int bifhip_write_bif1d(const char* path, int nPts, const double* ranges, const int* counts,
                       const double* peaks, int peakStride);
/* LLE1D / LS1D lines "param, v1[, v2 ...]\n" (Library/hostLibrary.cu:1368-1373,
 * :1900-1908). */
int bifhip_write_1d_values(const char* path, int nPts, const double* ranges,
                           const double* values, int perCell);
/* getValueByIdx (Library/cudaLibrary.cu:210-216) on the host. */
double bifhip_value_by_idx(int idx, int nPts, double startRange, double finishRange,
                           int valueNumber);

#ifdef __cplusplus
}
#endif
#endif
