/*
 * orc.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see orc.h).
 *
 * Every function restates one reference routine of KiShiVi/CUDABifurcation3.0
 * in plain C and cites the file:line it follows.  Compiled with
 * -ffp-contract=off and without -march, so double arithmetic is SSE2
 * binary64 with one rounding per operation in source order.
 */
#include "orc.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXD 8   /* largest state dimension handled */
#define ORC_MAXV 16  /* largest parameter vector handled */

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_model_dim(int model)
{
    return model == ORC_MODEL_ROSSLER ? 3 : (model == ORC_MODEL_MEMRISTIVE ? 5 : ORC_MAXD);
}

/* calculateDiscreteModel, Library/cudaLibrary.cu:15-17 (the Rossler body):
 * sequential explicit Euler -- x1 uses the new x0, x2 uses the new x0. */
static inline void rossler_step(double *x, const double *a, double h)
{
    x[0] = x[0] + h * (-x[1] - x[2]);
    x[1] = x[1] + h * (x[0] + a[0] * x[1]);
    x[2] = x[2] + h * (a[1] + x[2] * (x[0] - a[2]));
}

/* Linear diagonal test model x_i <- x_i + h*(a_i*x_i).  Not in the
 * reference; its Lyapunov spectrum is log|1+h*a_i|/h in closed form. */
static inline void linear_step(double *x, const double *a, double h, int n)
{
    for (int i = 0; i < n; ++i)
        x[i] = x[i] + h * (a[i] * x[i]);
}

/* calculateDiscreteModel, Library/cudaLibrary.cu:19-43 (the active body):
 * five-variable memristive system, composition step -- an explicit
 * sequential pass of length k = h*a7 with k held in a float (:19-27), the
 * state rounded to float (:29-33), then a semi-implicit pass of length
 * h*(1-a7) with the linear terms solved in closed form (:36-43).  abs() of
 * the double x4 is taken as the double absolute value. */
static inline void memristive_step(double *x, const double *a, double h)
{
    float y[5], k;
    k = (float)(h * a[7]);
    x[4] = x[4] + k * (x[0]);
    x[3] = x[3] + k * (-a[2] * x[1] - a[3] * x[3]);
    x[2] = x[2] + k * (a[1] * (x[1] - x[0] - x[2]));
    x[1] = x[1] + k * (-x[2] + x[3]);
    x[0] = x[0] + k * (a[0] * (x[2] - (a[4] + a[5] + a[6] * fabs(x[4])) * x[0]));
    for (int i = 0; i < 5; ++i)
        y[i] = (float)x[i];
    double h1 = h * (1 - a[7]);
    x[0] = (y[0] + h1 * (a[0] * x[2])) / (1 + h1 * a[0] * (a[4] + a[5] + a[6] * fabs(x[4])));
    x[1] = y[1] + h1 * (-x[2] + x[3]);
    x[2] = (y[2] + h1 * (a[1] * (x[1] - x[0]))) / (1 + h1 * a[1]);
    x[3] = (y[3] + h1 * (-a[2] * x[1])) / (1 + h1 * a[3]);
    x[4] = y[4] + h1 * (x[0]);
}

static _Thread_local int g_linear_dim = 3;

void orc_step(int model, double *x, const double *a, double h)
{
    if (model == ORC_MODEL_ROSSLER)
        rossler_step(x, a, h);
    else if (model == ORC_MODEL_MEMRISTIVE)
        memristive_step(x, a, h);
    else
        linear_step(x, a, h, g_linear_dim);
}

/* loopCalculateDiscreteModel, Library/cudaLibrary.cu:55-74:
 * record x[wv] BEFORE each iteration, take preScaller steps
 * (preScaller-1 extra, then one), then check |(float)x[wv]| > maxValue
 * when maxValue != 0 and stop early on divergence. */
int orc_loop(int model, double *x, const double *values, double h,
             int nIter, int preScaller, int writableVar, double maxValue,
             double *data, long start, int writeStep)
{
    for (int i = 0; i < nIter; ++i) {
        if (data != NULL)
            data[start + (long)i * writeStep] = x[writableVar];
        for (int j = 0; j < preScaller - 1; ++j)
            orc_step(model, x, values, h);
        orc_step(model, x, values, h);
        if (maxValue != 0)
            if (fabsf((float)x[writableVar]) > maxValue)
                return 0;
    }
    return 1;
}

/* getValueByIdx, Library/cudaLibrary.cu:210-216: float powf and float
 * division for the axis index, then one double product and one double sum. */
double orc_value_by_idx(int idx, int nPts, double startRange,
                        double finishRange, int valueNumber)
{
    float p = powf((float)nPts, (float)valueNumber);
    float q = (float)idx / p;
    int t = (int)q;
    return startRange + ((double)(t % nPts) *
                         ((double)(finishRange - startRange) / (double)(nPts - 1)));
}

/* peakFinder, Library/cudaLibrary.cu:220-266.  Indices are relative to
 * start (the reference stores absolute ones; only differences are used). */
int orc_peak_finder(double *data, long start, int amountOfPoints,
                    double *outPeaks, double *timeOfPeaks, double h)
{
    const double *d = data + start;
    double *op = outPeaks ? outPeaks + start : NULL;
    double *tp = timeOfPeaks ? timeOfPeaks + start : NULL;
    int count = 0;
    for (int i = 1; i < amountOfPoints - 1; ++i) {
        if (d[i] > d[i - 1] && d[i] >= d[i + 1]) {
            for (int j = i; j < amountOfPoints - 1; ++j) {
                if (d[j] < d[j + 1]) {
                    i = j + 1;
                    break;
                }
                if (d[j] > d[j + 1]) {
                    if (op)
                        op[count] = d[j];
                    if (tp)
                        tp[count] = (double)j;
                    ++count;
                    i = j + 1;
                    break;
                }
            }
        }
    }
    /* drop the first peak; intervals are index differences times h */
    if (count > 1) {
        for (int i = 0; i < count - 1; ++i) {
            if (op)
                op[i] = op[i + 1];
            if (tp)
                tp[i] = (double)((tp[i + 1] - tp[i]) * h);
        }
        count = count - 1;
    } else {
        count = 0;
    }
    return count;
}

/* kde, Library/cudaLibrary.cu:356-422: count the local maxima of a
 * Gaussian KDE of the peak amplitudes sampled on k1-1 points. */
int orc_kde(const double *data, long start, int amountOfPoints,
            int maxAmountOfPeaks, int kdeSampling, double I1, double I2,
            double smoothH)
{
    if (amountOfPoints == 0)
        return 0;
    if (amountOfPoints == 1 || amountOfPoints == 2)
        return 1;
    if (amountOfPoints > maxAmountOfPeaks)
        return maxAmountOfPeaks;

    double k1 = kdeSampling * amountOfPoints;
    double k2 = (I2 - I1) / (k1 - 1);
    double delt = 0, pp = 0, p = 0, cur = 0;
    int strange = 0, result = 0;
    for (int w = 0; w < k1 - 1; ++w) {
        delt = w * k2 + I1;
        pp = p;
        p = cur;
        cur = 0;
        for (int m = 0; m < amountOfPoints; ++m) {
            double t = (data[start + m] - delt) / smoothH;
            cur += expf((float)(-((t * t) / 2)));
        }
        if (w < 2)
            continue;
        if (strange) {
            if (p == cur)
                continue;
            else if (p < cur) {
                strange = 0;
                continue;
            } else if (p > cur) {
                strange = 0;
                ++result;
                continue;
            }
        } else if (p > pp && p > cur) {
            ++result;
            continue;
        } else if (p > pp && p == cur) {
            strange = 1;
            continue;
        }
    }
    if (p < cur)
        ++result;
    return result;
}

/* The KDE values kde() compares, Library/cudaLibrary.cu:368-386: sum w of
 * the k1-1 sample points (w < k1-1, k1 = kdeSampling*P, step (I2-I1)/(k1-1),
 * point w*step + I1) of sum_m expf(-(((p_m - point)/H)^2)/2), in the same
 * order and precision as orc_kde.  Writes min(k1-1, cap) values to out and
 * returns k1-1 (0 when the loop would not run). */
int orc_kde_sums(const double *data, int amountOfPoints, int kdeSampling, double I1,
                 double I2, double smoothH, double *out, int cap)
{
    double k1 = kdeSampling * amountOfPoints;
    double k2 = (I2 - I1) / (k1 - 1);
    int n = 0;
    for (int w = 0; w < k1 - 1; ++w, ++n) {
        double delt = w * k2 + I1, cur = 0;
        for (int m = 0; m < amountOfPoints; ++m) {
            double t = (data[m] - delt) / smoothH;
            cur += expf((float)(-((t * t) / 2)));
        }
        if (n < cap)
            out[n] = cur;
    }
    return n;
}

/* distance, Library/cudaLibrary.cu:445-453: float hypot of the double
 * differences, 0 for identical points. */
double orc_distance(double x1, double y1, double x2, double y2)
{
    if (x1 == x2 && y1 == y2)
        return 0;
    double dx = x2 - x1;
    double dy = y2 - y1;
    return hypotf((float)dx, (float)dy);
}

/* dbscan, Library/cudaLibrary.cu:457-529.  helpful[start .. start+size)
 * holds labels in [0,P) and the neighbour slots in [P, 2P).  Kept exactly:
 * the slot written is k, the slot popped is NumNeibor-1, the popped value
 * becomes the loop variable (then ++), scans stop at k < P-1, the result
 * is cluster-1. */
int orc_dbscan(const double *data, const double *intervals, double *helpful,
               long start, int amountOfPeaks, int sizeOfHelpful,
               int maxAmountOfPeaks, double eps)
{
    const int P = amountOfPeaks;
    if (P <= 0)
        return 0;
    if (P == 1)
        return 1;
    if (P > maxAmountOfPeaks)
        return 0;
    int cluster = 0, nn = 0;
    double *H = helpful + start;
    const double *A = data + start;
    const double *I = intervals + start;
    for (int i = 0; i < sizeOfHelpful; ++i)
        H[i] = 0;
    for (int i = 0; i < P; i++) {
        if (nn >= 1) {
            i = (int)H[P + nn - 1];
            H[P + nn - 1] = 0;
            nn = nn - 1;
            for (int k = 0; k < P - 1; k++) {
                if (i != k && H[k] == 0) {
                    if (orc_distance(A[i], I[i], A[k], I[k]) < eps) {
                        H[k] = cluster;
                        H[P + k] = k;
                        ++nn;
                    }
                }
            }
        } else if (H[i] == 0) {
            nn = 0;
            ++cluster;
            H[i] = cluster;
            for (int k = 0; k < P - 1; k++) {
                if (i != k && H[k] == 0) {
                    if (orc_distance(A[i], I[i], A[k], I[k]) < eps) {
                        H[k] = cluster;
                        H[P + k] = k;
                        ++nn;
                    }
                }
            }
        }
    }
    return cluster - 1;
}

/* projectionOperator + gramSchmidtProcess, Library/cudaLibrary.cu:913-957:
 * classical GS (project the original a_i on the unnormalised b_j), then
 * normalise each b_i; zero norm gives a zero vector.  Norms to denoms. */
void orc_gram_schmidt(const double *a, double *b, int n, double *denoms)
{
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j)
            b[j + i * n] = a[j + i * n];
        for (int j = 0; j < i; ++j) {
            const double *ai = a + i * n, *bj = b + j * n;
            double *bi = b + i * n;
            double num = 0, den = 0;
            for (int m = 0; m < n; ++m) {
                num += ai[m] * bj[m];
                den += bj[m] * bj[m];
            }
            double f = den == 0 ? 0 : num / den;
            for (int m = 0; m < n; ++m)
                bi[m] -= f * bj[m];
        }
    }
    for (int i = 0; i < n; ++i) {
        double den = 0;
        for (int j = 0; j < n; ++j)
            den += b[i * n + j] * b[i * n + j];
        den = sqrt(den);
        for (int j = 0; j < n; ++j)
            b[i * n + j] = den == 0 ? 0 : b[i * n + j] / den;
        if (denoms)
            denoms[i] = den;
    }
}

/* Sweep target: parameters (0, calculateDiscreteModelCUDA,
 * Library/cudaLibrary.cu:96-105) or initial conditions (1,
 * calculateDiscreteModelICCUDA and LLEKernelICCUDA, Library/cudaLibrary.cu:
 * 190-192, 837-839).  Set before a batch; read-only inside it. */
static int g_sweep_ic = 0;

void orc_set_sweep_ic(int on) { g_sweep_ic = on != 0; }

static void setup_cell(int model, int nPts, long cellIdx, int dimension,
                       const double *ranges, const int *mutIdx,
                       const double *ic, int nIC, const double *values,
                       int nVal, double *x, double *v)
{
    (void)model;
    for (int i = 0; i < nIC; ++i)
        x[i] = ic[i];
    for (int i = 0; i < nVal; ++i)
        v[i] = values[i];
    for (int i = 0; i < dimension; ++i) {
        double val = orc_value_by_idx((int)cellIdx, nPts, ranges[2 * i],
                                      ranges[2 * i + 1], i);
        if (g_sweep_ic)
            x[mutIdx[i]] = val;
        else
            v[mutIdx[i]] = val;
    }
}

/* calculateDiscreteModelCUDA, Library/cudaLibrary.cu:78-117: set the swept
 * parameters, transient without check, then record N samples. */
int orc_bif_cell(int model, int nPts, long cellIdx, int dimension,
                 const double *ranges, const int *mutIdx, const double *ic,
                 int nIC, const double *values, int nVal, double h, int skip,
                 int N, int preScaller, int writableVar, double maxValue,
                 double *series)
{
    double x[ORC_MAXD], v[ORC_MAXV];
    if (model == ORC_MODEL_LINEAR)
        g_linear_dim = nIC;
    setup_cell(model, nPts, cellIdx, dimension, ranges, mutIdx, ic, nIC,
               values, nVal, x, v);
    orc_loop(model, x, v, h, skip, 1, 0, 0, NULL, 0, 1);
    return orc_loop(model, x, v, h, N, preScaller, writableVar, maxValue,
                    series, 0, 1);
}

/* LLE direction, Library/cudaLibrary.cu:720-732: 0.5*sinf of the wrapped
 * 32-bit integer idx*(i*idx+1)+1, normalised in double; for a sweep over
 * initial conditions the LLEKernelICCUDA form (Library/cudaLibrary.cu:846). */
static void lle_direction(long localIdx, int nIC, int pertMode, double *z)
{
    if (pertMode == ORC_PERT_AXIS_X) {
        for (int i = 0; i < nIC; ++i)
            z[i] = i == 0 ? 1.0 : 0.0;
        return;
    }
    double zp = 0;
    uint32_t id = (uint32_t)localIdx;
    for (int i = 0; i < nIC; ++i) {
        if (g_sweep_ic) {
            /* LLEKernelICCUDA, Library/cudaLibrary.cu:846 */
            int idx = (int)localIdx;
            double arg = 0.2171828 * (i + 1) + idx + (0.2171828 + i * idx);
            z[i] = sinf((float)arg) * 0.5;
        } else {
            int32_t arg = (int32_t)(id * ((uint32_t)i * id + 1u) + 1u);
            z[i] = 0.5 * (double)sinf((float)arg);
        }
        zp += z[i] * z[i];
    }
    zp = sqrt(zp);
    for (int i = 0; i < nIC; ++i)
        z[i] /= zp;
}

/* LLEKernelCUDA, Library/cudaLibrary.cu:666-785 (Benettin two-trajectory). */
double orc_lle_cell(int model, int nPts, long cellIdx, long localIdx,
                    int dimension, const double *ranges, const int *mutIdx,
                    const double *ic, int nIC, const double *values, int nVal,
                    double NT, double tMax, double h, double eps, int skip,
                    double maxValue, int pertMode)
{
    double x[ORC_MAXD], y[ORC_MAXD], z[ORC_MAXD], v[ORC_MAXV];
    if (model == ORC_MODEL_LINEAR)
        g_linear_dim = nIC;
    setup_cell(model, nPts, cellIdx, dimension, ranges, mutIdx, ic, nIC,
               values, nVal, x, v);
    size_t nNT = (size_t)(NT / h);
    int sizeOfBlock = (int)(tMax / NT);
    lle_direction(localIdx, nIC, pertMode, z);
    orc_loop(model, x, v, h, skip, 1, 0, maxValue, NULL, 0, 1);
    for (int i = 0; i < nIC; ++i)
        y[i] = z[i] * eps + x[i];
    double result = 0;
    for (int i = 0; i < sizeOfBlock; ++i) {
        orc_loop(model, x, v, h, (int)nNT, 1, 0, maxValue, NULL, 0, 1);
        orc_loop(model, y, v, h, (int)nNT, 1, 0, maxValue, NULL, 0, 1);
        double t = 0;
        for (int l = 0; l < nIC; ++l)
            t += (x[l] - y[l]) * (x[l] - y[l]);
        t = sqrt(t) / eps;
        result += log(t);
        if (t != 0)
            t = (1 / t);
        for (int j = 0; j < nIC; ++j)
            y[j] = (double)(x[j] - ((x[j] - y[j]) * t));
    }
    return result / tMax;
}

/* LS directions, Library/cudaLibrary.cu:1025-1040. */
static void ls_directions(long localIdx, int n, int pertMode, double *z)
{
    for (int j = 0; j < n; ++j) {
        if (pertMode == ORC_PERT_AXIS_X) {
            for (int i = 0; i < n; ++i)
                z[j * n + i] = i == j ? 1.0 : 0.0;
            continue;
        }
        double zp = 0;
        int idx = (int)localIdx;
        for (int i = 0; i < n; ++i) {
            double arg = 0.2171828 * (i + 1) * (j + 1) + idx +
                         (0.2171828 + i * j * idx);
            z[j * n + i] = sinf((float)arg) * 0.5;
            zp += z[j * n + i] * z[j * n + i];
        }
        zp = sqrt(zp);
        for (int i = 0; i < n; ++i)
            z[j * n + i] /= zp;
    }
}

/* LSKernelCUDA, Library/cudaLibrary.cu:961-1100: n finite-difference
 * copies, GS every NT, sum of log(norm/eps). */
void orc_ls_cell(int model, int nPts, long cellIdx, long localIdx,
                 int dimension, const double *ranges, const int *mutIdx,
                 const double *ic, int nIC, const double *values, int nVal,
                 double NT, double tMax, double h, double eps, int skip,
                 double maxValue, int pertMode, double *out)
{
    const int n = nIC;
    double x[ORC_MAXD], v[ORC_MAXV];
    double y[ORC_MAXD * ORC_MAXD], z[ORC_MAXD * ORC_MAXD];
    double res[ORC_MAXD], den[ORC_MAXD];
    if (model == ORC_MODEL_LINEAR)
        g_linear_dim = nIC;
    setup_cell(model, nPts, cellIdx, dimension, ranges, mutIdx, ic, nIC,
               values, nVal, x, v);
    for (int i = 0; i < n; ++i)
        res[i] = den[i] = 0;
    size_t nNT = (size_t)(NT / h);
    int sizeOfBlock = (int)(tMax / NT);
    ls_directions(localIdx, n, pertMode, z);
    orc_loop(model, x, v, h, skip, 1, 0, maxValue, NULL, 0, 1);
    orc_gram_schmidt(z, y, n, NULL);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            y[j * n + i] = y[j * n + i] * eps + x[i];
    for (int it = 0; it < sizeOfBlock; ++it) {
        orc_loop(model, x, v, h, (int)nNT, 1, 0, maxValue, NULL, 0, 1);
        for (int j = 0; j < n; ++j)
            orc_loop(model, y + j * n, v, h, (int)nNT, 1, 0, maxValue, NULL, 0, 1);
        for (int k = 0; k < n; ++k)
            for (int l = 0; l < n; ++l)
                y[k * n + l] = y[k * n + l] - x[l];
        orc_gram_schmidt(y, z, n, den);
        for (int k = 0; k < n; ++k) {
            res[k] += log(den[k] / eps);
            for (int j = 0; j < n; ++j)
                y[k * n + j] = (double)(x[j] + z[k * n + j] * eps);
        }
    }
    for (int i = 0; i < n; ++i)
        out[i] = res[i] / tMax;
}

/* Variational (tangent-map) Lyapunov spectrum of the Rossler step, FP64.
 * Not in the reference, which uses finite-difference copies
 * (Library/cudaLibrary.cu:961-1100); this is the oracle of the variational
 * tile path.  The tangent map of the sequential Euler step
 *   x0' = x0 - h(x1+x2); x1' = x1 + h(x0' + a x1); x2' = x2 + h(b + x2(x0'-c))
 * is  v0' = v0 - h(v1+v2); v1' = (1+ha)v1 + h v0';
 *     v2' = (1 + h(x0'-c)) v2 + h x2 v0'    (x2 the pre-step value).
 * Every nNT steps the three tangent vectors are orthonormalised with the
 * reference's classical Gram-Schmidt (orc_gram_schmidt) and log(norm)
 * accumulates; out = sum / (nBlocks * nNT * h). */
void orc_ls_var_cell(double a, double b, double c, const double *ic, double h,
                     int skip, int nNT, int nBlocks, double *out)
{
    double x0 = ic[0], x1 = ic[1], x2 = ic[2];
    double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, w[9], den[3];
    double res[3] = {0, 0, 0};
    for (int s = 0; s < skip; ++s) {
        x0 = x0 + h * (-x1 - x2);
        x1 = x1 + h * (x0 + a * x1);
        x2 = x2 + h * (b + x2 * (x0 - c));
    }
    for (int blk = 0; blk < nBlocks; ++blk) {
        for (int s = 0; s < nNT; ++s) {
            const double x2o = x2;
            x0 = x0 + h * (-x1 - x2);
            x1 = x1 + h * (x0 + a * x1);
            x2 = x2 + h * (b + x2 * (x0 - c));
            for (int k = 0; k < 3; ++k) {
                double *q = v + 3 * k;
                q[0] = q[0] - h * (q[1] + q[2]);
                q[1] = (1 + h * a) * q[1] + h * q[0];
                q[2] = (1 + h * (x0 - c)) * q[2] + h * x2o * q[0];
            }
        }
        orc_gram_schmidt(v, w, 3, den);
        for (int k = 0; k < 3; ++k)
            res[k] += log(den[k]);
        memcpy(v, w, sizeof(v));
    }
    const double T = (double)nBlocks * nNT * h;
    for (int k = 0; k < 3; ++k)
        out[k] = res[k] / T;
}

void orc_ls_var_cells(const double *a, const double *b, const double *c,
                      const double *ic, double h, int skip, int nNT,
                      const int *nBlocks, long n, double *out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < n; ++i)
        orc_ls_var_cell(a[i], b[i], c[i], ic, h, skip, nNT, nBlocks[i], out + 3 * i);
}

/* ---------------- map-level drivers ---------------- */

void orc_peaks_cells(int model, int dimension, double tMax, int nPts, double h,
                     int nIC, const double *ic, const double *ranges,
                     const int *mutIdx, int writableVar, double maxValue,
                     double transientTime, const double *values, int nVal,
                     int preScaller, long c0, long nc, int *counts,
                     double *peaks, double *intervals, int cap)
{
    /* hostLibrary.cu:22-23 / :588-589 derived counts */
    int N = (int)(tMax / h / preScaller);
    int skip = (int)(transientTime / h);
#pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * (size_t)N);
        double *iv = (double *)malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 64)
        for (long c = 0; c < nc; ++c) {
            int ok = orc_bif_cell(model, nPts, c0 + c, dimension, ranges,
                                  mutIdx, ic, nIC, values, nVal, h, skip, N,
                                  preScaller, writableVar, maxValue, s);
            int P = 0;
            if (ok) /* peakFinderCUDA, cudaLibrary.cu:277-282 */
                P = orc_peak_finder(s, 0, N, s, iv, 0.01);
            counts[c] = ok ? P : -1;
            int m = P < cap ? P : cap;
            if (peaks)
                for (int k = 0; k < m; ++k)
                    peaks[c * (long)cap + k] = s[k];
            if (intervals)
                for (int k = 0; k < m; ++k)
                    intervals[c * (long)cap + k] = iv[k];
        }
        free(s);
        free(iv);
    }
}

void orc_bifurcation1d(int model, double tMax, int nPts, double h, int nIC,
                       const double *ic, const double *ranges,
                       const int *mutIdx, int writableVar, double maxValue,
                       double transientTime, const double *values, int nVal,
                       int preScaller, int *counts, double *peaks, int cap)
{
    orc_peaks_cells(model, 1, tMax, nPts, h, nIC, ic, ranges, mutIdx,
                    writableVar, maxValue, transientTime, values, nVal,
                    preScaller, 0, nPts, counts, peaks, NULL, cap);
    for (int c = 0; c < nPts; ++c)
        if (counts[c] < 0)
            counts[c] = 0;
}

void orc_dbscan_cells(int model, double tMax, int nPts, double h, int nIC,
                      const double *ic, const double *ranges,
                      const int *mutIdx, int writableVar, double maxValue,
                      int maxAmountOfPeaks, double transientTime,
                      const double *values, int nVal, int preScaller,
                      double eps, long c0, long nc, int *out)
{
    int N = (int)(tMax / h / preScaller);
    int skip = (int)(transientTime / h);
#pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * (size_t)N);
        double *iv = (double *)malloc(sizeof(double) * (size_t)N);
        double *hp = (double *)malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 64)
        for (long c = 0; c < nc; ++c) {
            int ok = orc_bif_cell(model, nPts, c0 + c, 2, ranges, mutIdx, ic,
                                  nIC, values, nVal, h, skip, N, preScaller,
                                  writableVar, maxValue, s);
            if (!ok) { /* dbscanCUDA, cudaLibrary.cu:617-621 */
                out[c] = 0;
                continue;
            }
            int P = orc_peak_finder(s, 0, N, s, iv, 0.01);
            out[c] = orc_dbscan(s, iv, hp, 0, P, N, maxAmountOfPeaks, eps);
        }
        free(s);
        free(iv);
        free(hp);
    }
}

/* FP32 form of the peak maps (the library's bifhip_set_precision(FP32) for
 * the 1D diagram and the 2D cluster maps; no reference counterpart).  The
 * state is float, stepped with the fused FP32 map below (fmaf is correctly
 * rounded in glibc, as v_fma_f32 is), each sample widened to double for the
 * reference's peak finder and DBSCAN.  Rossler:
 *   x0 <- fma(-h, x1 + x2, x0)
 *   x1 <- fma(x1, (float)(1 + h*a), h*x0)
 *   x2 <- fma(x2, fma(h, x0 - c, 1), (float)(h*b))
 * linear: x_i <- x_i * (float)(1 + h*a_i). */
static void step32_params(int model, const double *v, double h, float *q)
{
    if (model == ORC_MODEL_ROSSLER) {
        q[0] = (float)h;
        q[1] = (float)(1.0 + h * v[0]);
        q[2] = (float)v[2];
        q[3] = (float)(h * v[1]);
    } else {
        for (int i = 0; i < 3; ++i)
            q[i] = (float)(1.0 + h * v[i]);
        q[3] = 0;
    }
}

static inline void step32(int model, float *x, const float *q)
{
    if (model == ORC_MODEL_ROSSLER) {
        x[0] = fmaf(-q[0], x[1] + x[2], x[0]);
        x[1] = fmaf(x[1], q[1], q[0] * x[0]);
        x[2] = fmaf(x[2], fmaf(q[0], x[0] - q[2], 1.0f), q[3]);
    } else {
        x[0] *= q[0];
        x[1] *= q[1];
        x[2] *= q[2];
    }
}

int orc_bif_cell32(int model, int nPts, long cellIdx, int dimension,
                   const double *ranges, const int *mutIdx, const double *ic,
                   int nIC, const double *values, int nVal, double h, int skip,
                   int N, int preScaller, int writableVar, double maxValue,
                   double *series)
{
    double x[ORC_MAXD], v[ORC_MAXV];
    float q[4], y[3];
    if (model == ORC_MODEL_MEMRISTIVE || nIC != 3)
        return -1;
    setup_cell(model, nPts, cellIdx, dimension, ranges, mutIdx, ic, nIC,
               values, nVal, x, v);
    step32_params(model, v, h, q);
    for (int i = 0; i < 3; ++i)
        y[i] = (float)x[i];
    for (int i = 0; i < skip; ++i)
        step32(model, y, q);
    for (int i = 0; i < N; ++i) {
        series[i] = (double)y[writableVar];
        for (int j = 0; j < preScaller; ++j)
            step32(model, y, q);
        if (maxValue != 0 && fabsf(y[writableVar]) > maxValue)
            return 0;
    }
    return 1;
}

void orc_peaks_cells32(int model, int dimension, double tMax, int nPts,
                       double h, int nIC, const double *ic,
                       const double *ranges, const int *mutIdx,
                       int writableVar, double maxValue, double transientTime,
                       const double *values, int nVal, int preScaller,
                       long c0, long nc, int *counts, double *peaks,
                       double *intervals, int cap)
{
    int N = (int)(tMax / h / preScaller);
    int skip = (int)(transientTime / h);
#pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        double *iv = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
#pragma omp for schedule(dynamic, 64)
        for (long c = 0; c < nc; ++c) {
            int ok = orc_bif_cell32(model, nPts, c0 + c, dimension, ranges,
                                    mutIdx, ic, nIC, values, nVal, h, skip,
                                    N, preScaller, writableVar, maxValue, s);
            int P = 0;
            if (ok == 1)
                P = orc_peak_finder(s, 0, N, s, iv, 0.01);
            counts[c] = ok == 1 ? P : -1;
            int m = P < cap ? P : cap;
            if (peaks)
                for (int k = 0; k < m; ++k)
                    peaks[c * (long)cap + k] = s[k];
            if (intervals)
                for (int k = 0; k < m; ++k)
                    intervals[c * (long)cap + k] = iv[k];
        }
        free(s);
        free(iv);
    }
}

void orc_dbscan_cells32(int model, double tMax, int nPts, double h, int nIC,
                        const double *ic, const double *ranges,
                        const int *mutIdx, int writableVar, double maxValue,
                        int maxAmountOfPeaks, double transientTime,
                        const double *values, int nVal, int preScaller,
                        double eps, long c0, long nc, int *out)
{
    int N = (int)(tMax / h / preScaller);
    int skip = (int)(transientTime / h);
#pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        double *iv = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        double *hp = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
#pragma omp for schedule(dynamic, 64)
        for (long c = 0; c < nc; ++c) {
            int ok = orc_bif_cell32(model, nPts, c0 + c, 2, ranges, mutIdx,
                                    ic, nIC, values, nVal, h, skip, N,
                                    preScaller, writableVar, maxValue, s);
            if (ok != 1) {
                out[c] = 0;
                continue;
            }
            int P = orc_peak_finder(s, 0, N, s, iv, 0.01);
            out[c] = orc_dbscan(s, iv, hp, 0, P, N, maxAmountOfPeaks, eps);
        }
        free(s);
        free(iv);
        free(hp);
    }
}

/* calculationOfPeriodicityByOstrovsky, Library/hostLibrary.cu:748-925:
 * initial-condition sweep (calculateDiscreteModelICCUDA), peaks with five
 * flags (peakFinderCUDAForCalculationOfPeriodicityByOstrovsky,
 * Library/cudaLibrary.cu:286-352), DBSCAN (dbscanDouble, the dbscan
 * algorithm, :533-605) and the flag rules of
 * dbscanCUDAForCalculationOfPeriodicityByOstrovsky (:628-662).  A diverged
 * cell sets only flag 3 in the reference (the others are unwritten device
 * memory); here they are false. */
void orc_ostrovsky_cells(int model, double tMax, int nPts, double h, int nIC,
                         const double *ic, const double *ranges,
                         const int *mutIdx, int writableVar, double maxValue,
                         int maxAmountOfPeaks, double transientTime,
                         const double *values, int nVal, int preScaller,
                         double eps, double threshold, long c0, long nc,
                         double *out)
{
    int N = (int)(tMax / h / preScaller);
    int skip = (int)(transientTime / h);
    int saved = g_sweep_ic;
    g_sweep_ic = 1;
#pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * (size_t)N);
        double *iv = (double *)malloc(sizeof(double) * (size_t)N);
        double *hp = (double *)malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 64)
        for (long c = 0; c < nc; ++c) {
            int ok = orc_bif_cell(model, nPts, c0 + c, 2, ranges, mutIdx, ic,
                                  nIC, values, nVal, h, skip, N, preScaller,
                                  writableVar, maxValue, s);
            int f[5] = {0, 0, 0, 0, 0};
            double o;
            if (!ok) {
                f[3] = 1;
                o = 0;
            } else {
                double last = s[N - 1];
                int P = orc_peak_finder(s, 0, N, s, iv, 0.01);
                f[0] = 1;
                for (int i = 1; i < P; ++i)
                    if (s[i] - s[i - 1] > 0) {
                        f[0] = 0;
                        break;
                    }
                int one = 0, zero = 0;
                for (int i = 1; i < P; ++i) {
                    if (s[i] > threshold)
                        one = 1;
                    else
                        zero = 1;
                    if (one && zero)
                        break;
                }
                f[1] = one && zero;
                f[2] = !(one && !zero);
                f[3] = P == 0 || P == 1;
                f[4] = last > threshold;
                o = orc_dbscan(s, iv, hp, 0, P, N, maxAmountOfPeaks, eps);
            }
            if (f[2])
                o = -1 * o;
            if (f[3])
                o = f[4] ? 1 : -1;
            if (f[1])
                o = 0;
            if (f[0])
                o = o > 0 ? 0.5 : (o == 0 ? 0 : -0.5);
            out[c] = o;
        }
        free(s);
        free(iv);
        free(hp);
    }
    g_sweep_ic = saved;
}

/* modelingOneSystemDenis + calculateDiscreteModelDenisCUDA,
 * Library/hostLibrary.cu:162-239, Library/cudaLibrary.cu:121-158:
 * trajectory idx < (int)(hS/h) takes idx steps of h, then (int)(tMax/hS)
 * steps of hS recording x[wv] before each into out[idx + i*T]; out has
 * (int)(tMax/h) points, the unwritten ones stay 0. */
void orc_denis(int model, double tMax, double h, double hS, int nIC,
               const double *ic, int writableVar, const double *values,
               int nVal, double *out, long nOut)
{
    int T = (int)(hS / h);
    int iters = (int)(tMax / hS);
    for (long j = 0; j < nOut; ++j)
        out[j] = 0;
    if (model == ORC_MODEL_LINEAR)
        g_linear_dim = nIC;
    for (int idx = 0; idx < T; ++idx) {
        double x[ORC_MAXD], v[ORC_MAXV];
        for (int i = 0; i < nIC; ++i)
            x[i] = ic[i];
        for (int i = 0; i < nVal; ++i)
            v[i] = values[i];
        for (int s = 0; s < idx; ++s)
            orc_step(model, x, v, h);
        for (int i = 0; i < iters; ++i) {
            long j = idx + (long)i * T;
            if (j < nOut)
                out[j] = x[writableVar];
            orc_step(model, x, v, hS);
        }
    }
}

void orc_kde_cells(int model, double tMax, int nPts, double h, int nIC,
                   const double *ic, const double *ranges, const int *mutIdx,
                   int writableVar, double maxValue, int maxAmountOfPeaks,
                   double transientTime, const double *values, int nVal,
                   int preScaller, int kdeSampling, double I1, double I2,
                   double smoothH, long c0, long nc, int *out)
{
    int N = (int)(tMax / h / preScaller);
    int skip = (int)(transientTime / h);
#pragma omp parallel
    {
        double *s = (double *)malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 64)
        for (long c = 0; c < nc; ++c) {
            int ok = orc_bif_cell(model, nPts, c0 + c, 2, ranges, mutIdx, ic,
                                  nIC, values, nVal, h, skip, N, preScaller,
                                  writableVar, maxValue, s);
            if (!ok) { /* kdeCUDA, cudaLibrary.cu:434-438 */
                out[c] = 0;
                continue;
            }
            int P = orc_peak_finder(s, 0, N, s, NULL, 0.01);
            out[c] = orc_kde(s, 0, P, maxAmountOfPeaks, kdeSampling, I1, I2,
                             smoothH);
        }
        free(s);
    }
}

void orc_lle_cells(int model, int dimension, double tMax, double NT, int nPts,
                   double h, double eps, const double *ic, int nIC,
                   const double *ranges, const int *mutIdx, double maxValue,
                   double transientTime, const double *values, int nVal,
                   int pertMode, long dirChunk, long c0, long nc, double *out)
{
    int skip = (int)(transientTime / h);
#pragma omp parallel for schedule(dynamic, 64)
    for (long c = 0; c < nc; ++c) {
        long g = c0 + c;
        long local = dirChunk > 0 ? g % dirChunk : g;
        out[c] = orc_lle_cell(model, nPts, g, local, dimension, ranges, mutIdx,
                              ic, nIC, values, nVal, NT, tMax, h, eps, skip,
                              maxValue, pertMode);
    }
}

void orc_ls_cells(int model, int dimension, double tMax, double NT, int nPts,
                  double h, double eps, const double *ic, int nIC,
                  const double *ranges, const int *mutIdx, double maxValue,
                  double transientTime, const double *values, int nVal,
                  int pertMode, long dirChunk, long c0, long nc, double *out)
{
    int skip = (int)(transientTime / h);
#pragma omp parallel for schedule(dynamic, 64)
    for (long c = 0; c < nc; ++c) {
        long g = c0 + c;
        long local = dirChunk > 0 ? g % dirChunk : g;
        orc_ls_cell(model, nPts, g, local, dimension, ranges, mutIdx, ic, nIC,
                    values, nVal, NT, tMax, h, eps, skip, maxValue, pertMode,
                    out + c * nIC);
    }
}
