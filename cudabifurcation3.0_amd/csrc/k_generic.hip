// k_generic.hip -- kernels for models with a state of D variables held as a
// register array (the five-variable memristive system of the reference's
// calculateDiscreteModel, Library/cudaLibrary.cu:19-43).  Same algorithms
// and operation order as the three-variable kernels in k_bif.hip and
// k_lyap.hip, one trajectory per lane:
//
//  k_gen_bif_peaks  transient, recorded window into the streaming peak
//                   finder (calculateDiscreteModelCUDA + peakFinder,
//                   Library/cudaLibrary.cu:78-117, 220-266);
//  k_gen_series     the raw recorded series (parity tests);
//  k_gen_lle        Benettin two-trajectory LLE (LLEKernelCUDA,
//                   Library/cudaLibrary.cu:666-785);
//  k_gen_ls         finite-difference spectrum with classical Gram-Schmidt
//                   (LSKernelCUDA / gramSchmidtProcess,
//                   Library/cudaLibrary.cu:913-1100);
//  k_denis          one system sampled from offset coarse-step trajectories
//                   (calculateDiscreteModelDenisCUDA, :121-158);
//  k_osc_flags,     Ostrovsky's periodicity classifier on top of the peak
//  k_osc_combine    rows and DBSCAN counts (Library/cudaLibrary.cu:286-352,
//                   628-662).
// k_gen_bif_peaks also serves the three-variable models when the last
// recorded sample is wanted (the classifier's fifth flag).
#include "bifhip_dev.h"
#include "peak_stream.h"

namespace bifhip {

template <int D>
__device__ __forceinline__ double component(const double (&x)[D], int i)
{
    double r = x[0];
#pragma unroll
    for (int k = 1; k < D; ++k)
        if (k == i) r = x[k];
    return r;
}

// nIter steps with the reference's early stop on x[wv] (writableVar 0 in
// the Lyapunov kernels).
template <class M, bool CHECK>
__device__ __forceinline__ void run_n(double (&x)[M::D], const typename M::P64& p, double h,
                                      int nIter, int wv, double maxValue)
{
    for (int i = 0; i < nIter; ++i) {
        M::stepn(x, p, h);
        if (CHECK && diverged(component<M::D>(x, wv), maxValue)) return;
    }
}

// last[t] (when given): the last recorded sample of cell t.
template <class M, bool CHECK>
__global__ __launch_bounds__(256) void k_gen_bif_peaks(const BifArgs A, int wv,
                                                       int* __restrict__ counts,
                                                       double* __restrict__ amps,
                                                       double* __restrict__ ivls,
                                                       double* __restrict__ last)
{
    constexpr int D = M::D;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal], x[D];
    cell_setup_n<D>(A.g, idx, v, x);
    const typename M::P64 p = M::params(v);
    const double h = A.h;
    // transient, no divergence check (Library/cudaLibrary.cu:107-108)
    for (int s = 0; s < A.skip; ++s) M::stepn(x, p, h);

    PeakSink sink;
    sink.amps = amps ? amps + (long long)t * A.rowStride : nullptr;
    sink.ivls = ivls ? ivls + (long long)t * A.rowStride : nullptr;
    sink.cap = A.cap;
    sink.hpf = A.hpf;
    PeakStream ps;
    ps.init();
    bool bad = false;
    double rec = 0;
    for (int i = 0; i < A.N && !bad; ++i) {
        const double d = component<D>(x, wv);
        rec = d;
        if (i >= 2) ps.push(d, i, sink);
        else { ps.p2 = ps.p1; ps.p1 = d; }
        for (int j = 0; j < A.pre - 1; ++j) M::stepn(x, p, h);
        M::stepn(x, p, h);
        if (CHECK && diverged(component<D>(x, wv), A.maxValue)) bad = true;
    }
    counts[t] = bad ? -1 : ps.count();
    if (last) last[t] = rec;
}

// Ostrovsky's five flags of a cell from its full peak row
// (peakFinderCUDAForCalculationOfPeriodicityByOstrovsky,
// Library/cudaLibrary.cu:293-351), bit k = flag k:
//   0 no kept peak rises above its predecessor
//   1 kept peaks after the first lie on both sides of the threshold
//   2 not (all of them above it)
//   3 at most one kept peak (or a diverged trajectory)
//   4 the last recorded sample is above the threshold
// A diverged cell (count -1) sets only flag 3.
__global__ __launch_bounds__(256) void k_osc_flags(int n, const int* __restrict__ counts,
                                                   const double* __restrict__ amps,
                                                   long long rowStride,
                                                   const double* __restrict__ last, double thr,
                                                   unsigned char* __restrict__ flags)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int P = counts[c];
    if (P == -1) {
        flags[c] = 1u << 3;
        return;
    }
    const double* pk = amps + (long long)c * rowStride;
    bool f0 = true, one = false, zero = false;
    for (int i = 1; i < P; ++i) {
        const double a = pk[i];
        if (a - pk[i - 1] > 0) f0 = false;
        if (a > thr) one = true;
        else zero = true;
    }
    const bool f1 = one && zero;
    const bool f2 = !(one && !zero);
    const bool f3 = P == 0 || P == 1;
    const bool f4 = last[c] > thr;
    flags[c] = (unsigned char)((f0 ? 1 : 0) | (f1 ? 2 : 0) | (f2 ? 4 : 0) | (f3 ? 8 : 0) |
                               (f4 ? 16 : 0));
}

// The classifier's value from the DBSCAN count and the flags
// (dbscanCUDAForCalculationOfPeriodicityByOstrovsky,
// Library/cudaLibrary.cu:641-661), rules applied in the reference's order.
__global__ __launch_bounds__(256) void k_osc_combine(int n, const int* __restrict__ clusters,
                                                     const unsigned char* __restrict__ flags,
                                                     double* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const unsigned f = flags[c];
    double o = (double)clusters[c];
    if (f & 4) o = -1 * o;
    if (f & 8) o = (f & 16) ? 1 : -1;
    if (f & 2) o = 0;
    if (f & 1) o = o > 0 ? 0.5 : (o == 0 ? 0 : -0.5);
    out[c] = o;
}

template <class M>
__global__ __launch_bounds__(256) void k_gen_series(const BifArgs A, int wv,
                                                    double* __restrict__ data,
                                                    int* __restrict__ flags)
{
    constexpr int D = M::D;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal], x[D];
    cell_setup_n<D>(A.g, idx, v, x);
    const typename M::P64 p = M::params(v);
    const double h = A.h;
    for (int s = 0; s < A.skip; ++s) M::stepn(x, p, h);
    double* row = data + (long long)t * A.rowStride;
    int flag = 0;
    for (int i = 0; i < A.N; ++i) {
        row[i] = component<D>(x, wv);
        for (int j = 0; j < A.pre - 1; ++j) M::stepn(x, p, h);
        M::stepn(x, p, h);
        if (A.maxValue != 0 && diverged(component<D>(x, wv), A.maxValue)) {
            flag = -1;
            break;
        }
    }
    flags[t] = flag;
}

template <class M, bool CHECK>
__global__ __launch_bounds__(256) void k_gen_lle(const LyapArgs A, double* __restrict__ out)
{
    constexpr int D = M::D;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal], x[D], y[D], z[D];
    cell_setup_n<D>(A.g, idx, v, x);
    const typename M::P64 p = M::params(v);
    const double h = A.h, eps = A.eps;
#pragma unroll
    for (int i = 0; i < D; ++i) z[i] = A.dirs ? A.dirs[(long long)t * D + i] : (i == 0 ? 1.0 : 0.0);
    run_n<M, CHECK>(x, p, h, A.skip, 0, A.maxValue);
#pragma unroll
    for (int i = 0; i < D; ++i) y[i] = z[i] * eps + x[i];
    double result = 0;
    for (int b = 0; b < A.nBlocks; ++b) {
        run_n<M, CHECK>(x, p, h, A.nNT, 0, A.maxValue);
        run_n<M, CHECK>(y, p, h, A.nNT, 0, A.maxValue);
        double s = 0;
#pragma unroll
        for (int l = 0; l < D; ++l) s += (x[l] - y[l]) * (x[l] - y[l]);
        s = sqrt(s) / eps;
        result += log(s);
        if (s != 0) s = (1 / s);
#pragma unroll
        for (int j = 0; j < D; ++j) y[j] = x[j] - ((x[j] - y[j]) * s);
    }
    out[t] = result / A.tMax;
}

// Classical Gram-Schmidt of D row vectors a -> b with norms.
template <int D>
__device__ __forceinline__ void gs_n(const double (&a)[D * D], double (&b)[D * D],
                                     double (&den)[D])
{
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int j = 0; j < D; ++j) b[i * D + j] = a[i * D + j];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double num = 0, dd = 0;
#pragma unroll
            for (int m = 0; m < D; ++m) {
                num += a[i * D + m] * b[j * D + m];
                dd += b[j * D + m] * b[j * D + m];
            }
            const double f = dd == 0 ? 0 : num / dd;
#pragma unroll
            for (int m = 0; m < D; ++m) b[i * D + m] -= f * b[j * D + m];
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double d = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) d += b[i * D + j] * b[i * D + j];
        d = sqrt(d);
#pragma unroll
        for (int j = 0; j < D; ++j) b[i * D + j] = d == 0 ? 0 : b[i * D + j] / d;
        den[i] = d;
    }
}

template <class M, bool CHECK>
__global__ __launch_bounds__(256) void k_gen_ls(const LyapArgs A, double* __restrict__ out)
{
    constexpr int D = M::D;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal], x[D];
    cell_setup_n<D>(A.g, idx, v, x);
    const typename M::P64 p = M::params(v);
    const double h = A.h, eps = A.eps;
    double z[D * D], y[D * D], den[D], res[D];
#pragma unroll
    for (int i = 0; i < D * D; ++i)
        z[i] = A.dirs ? A.dirs[(long long)t * D * D + i] : ((i % (D + 1)) == 0 ? 1.0 : 0.0);
#pragma unroll
    for (int i = 0; i < D; ++i) res[i] = 0;
    run_n<M, CHECK>(x, p, h, A.skip, 0, A.maxValue);
    gs_n<D>(z, y, den);
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
        for (int i = 0; i < D; ++i) y[j * D + i] = y[j * D + i] * eps + x[i];
    for (int b = 0; b < A.nBlocks; ++b) {
        run_n<M, CHECK>(x, p, h, A.nNT, 0, A.maxValue);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double yk[D];
#pragma unroll
            for (int i = 0; i < D; ++i) yk[i] = y[k * D + i];
            run_n<M, CHECK>(yk, p, h, A.nNT, 0, A.maxValue);
#pragma unroll
            for (int i = 0; i < D; ++i) y[k * D + i] = yk[i] - x[i];
        }
        gs_n<D>(y, z, den);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            res[k] += log(den[k] / eps);
#pragma unroll
            for (int j = 0; j < D; ++j) y[k * D + j] = (double)(x[j] + z[k * D + j] * eps);
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) out[(long long)t * D + i] = res[i] / A.tMax;
}

// One system sampled finely from many coarse trajectories
// (calculateDiscreteModelDenisCUDA, Library/cudaLibrary.cu:121-158): lane
// idx < T takes idx steps of h, then `iters` steps of hS, recording x[wv]
// before each into data[idx + i*T] (only indices below nPoints exist).
template <class M>
__global__ __launch_bounds__(256) void k_denis(const Grid g, int T, double h, double hS,
                                               int iters, int wv, long long nPoints,
                                               double* __restrict__ data)
{
    constexpr int D = M::D;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T) return;
    double v[kMaxVal], x[D];
    cell_setup_n<D>(g, 0, v, x);
    const typename M::P64 p = M::params(v);
    for (int s = 0; s < idx; ++s) M::stepn(x, p, h);
    for (int i = 0; i < iters; ++i) {
        const long long j = idx + (long long)i * T;
        if (j < nPoints) data[j] = component<D>(x, wv);
        M::stepn(x, p, hS);
    }
}

// ------------------------------------------------------------ launchers

template <class M>
static void gen_bif_m(const BifArgs& A, int wv, bool check, int* counts, double* amps,
                      double* ivls, double* last, hipStream_t st)
{
    const dim3 blk(256), grd((A.nCells + 255) / 256);
    if (check)
        hipLaunchKernelGGL((k_gen_bif_peaks<M, true>), grd, blk, 0, st, A, wv, counts, amps, ivls,
                           last);
    else
        hipLaunchKernelGGL((k_gen_bif_peaks<M, false>), grd, blk, 0, st, A, wv, counts, amps,
                           ivls, last);
}

hipError_t launch_gen_bif(int model, const BifArgs& A, int wv, bool check, int* counts,
                          double* amps, double* ivls, double* last, hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model == kMemristive) gen_bif_m<Memristive>(A, wv, check, counts, amps, ivls, last, st);
    else if (model == kLinear) gen_bif_m<Linear>(A, wv, check, counts, amps, ivls, last, st);
    else gen_bif_m<Rossler>(A, wv, check, counts, amps, ivls, last, st);
    return hipGetLastError();
}

hipError_t launch_osc_flags(int n, const int* counts, const double* amps, long long rowStride,
                            const double* last, double thr, unsigned char* flags, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_osc_flags, dim3((n + 255) / 256), dim3(256), 0, st, n, counts, amps,
                       rowStride, last, thr, flags);
    return hipGetLastError();
}

hipError_t launch_osc_combine(int n, const int* clusters, const unsigned char* flags,
                              double* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_osc_combine, dim3((n + 255) / 256), dim3(256), 0, st, n, clusters,
                       flags, out);
    return hipGetLastError();
}

hipError_t launch_gen_series(int model, const BifArgs& A, int wv, double* data, int* flags,
                             hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model != kMemristive) return hipErrorInvalidValue;
    const dim3 blk(256), grd((A.nCells + 255) / 256);
    hipLaunchKernelGGL((k_gen_series<Memristive>), grd, blk, 0, st, A, wv, data, flags);
    return hipGetLastError();
}

hipError_t launch_gen_lle(int model, const LyapArgs& A, bool check, double* out,
                          hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model != kMemristive) return hipErrorInvalidValue;
    const dim3 blk(256), grd((A.nCells + 255) / 256);
    if (check)
        hipLaunchKernelGGL((k_gen_lle<Memristive, true>), grd, blk, 0, st, A, out);
    else
        hipLaunchKernelGGL((k_gen_lle<Memristive, false>), grd, blk, 0, st, A, out);
    return hipGetLastError();
}

hipError_t launch_gen_ls(int model, const LyapArgs& A, bool check, double* out, hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model != kMemristive) return hipErrorInvalidValue;
    const dim3 blk(256), grd((A.nCells + 255) / 256);
    if (check)
        hipLaunchKernelGGL((k_gen_ls<Memristive, true>), grd, blk, 0, st, A, out);
    else
        hipLaunchKernelGGL((k_gen_ls<Memristive, false>), grd, blk, 0, st, A, out);
    return hipGetLastError();
}

}  // namespace bifhip

namespace bifhip {
hipError_t launch_denis(int model, const Grid& g, int T, double h, double hS, int iters, int wv,
                        long long nPoints, double* data, hipStream_t st)
{
    if (T <= 0) return hipSuccess;
    const dim3 blk(256), grd((T + 255) / 256);
    if (model == kMemristive)
        hipLaunchKernelGGL((k_denis<Memristive>), grd, blk, 0, st, g, T, h, hS, iters, wv,
                           nPoints, data);
    else if (model == kLinear)
        hipLaunchKernelGGL((k_denis<Linear>), grd, blk, 0, st, g, T, h, hS, iters, wv, nPoints,
                           data);
    else
        hipLaunchKernelGGL((k_denis<Rossler>), grd, blk, 0, st, g, T, h, hS, iters, wv, nPoints,
                           data);
    return hipGetLastError();
}
}  // namespace bifhip
