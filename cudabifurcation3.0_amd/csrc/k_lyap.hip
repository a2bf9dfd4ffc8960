// k_lyap.hip -- Lyapunov kernels for gfx950.
//
//  k_lle64   Benettin two-trajectory LLE (LLEKernelCUDA,
//            Library/cudaLibrary.cu:666-785), FP64, reference operation
//            order; x and y stepped interleaved (independent chains give the
//            scheduler two dependency streams per lane).
//  k_lle32   FP32 bulk: x in FP32 with fused multiply-adds, the separation
//            carried as a scaled deviation u (y = x + eps*u, exact
//            difference of the two maps, so eps = 1e-8 survives FP32);
//            norm, renormalisation (rsqrt) and the log-norm sum in FP64,
//            the sum kept as a product of squared norms and logged once.
//  k_ls64    finite-difference Lyapunov spectrum (LSKernelCUDA,
//            Library/cudaLibrary.cu:961-1100) with classical Gram-Schmidt
//            (Library/cudaLibrary.cu:913-957), FP64, reference order.
//  k_gs      batched gramSchmidtProcess on n x n matrices (parity tests).
#include "bifhip_dev.h"

namespace bifhip {


// nIter steps with the reference's early stop (loopCalculateDiscreteModel,
// Library/cudaLibrary.cu:59-72; writableVar 0 in the LLE/LS kernels).
template <class M, bool CHECK>
__device__ __forceinline__ void run(double& x0, double& x1, double& x2,
                                    const typename M::P64& p, double h, int nIter,
                                    double maxValue)
{
    for (int i = 0; i < nIter; ++i) {
        M::step(x0, x1, x2, p, h);
        if (CHECK && diverged(x0, maxValue)) return;
    }
}

// Two independent trajectories advanced together; each keeps its own early
// stop, which leaves its arithmetic identical to running it alone.
template <class M, bool CHECK>
__device__ __forceinline__ void run2(double& x0, double& x1, double& x2, double& y0, double& y1,
                                     double& y2, const typename M::P64& p, double h, int nIter,
                                     double maxValue)
{
    if (!CHECK) {
        int i = 0;
        for (; i + 2 <= nIter; i += 2) {
            M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
            M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
        }
        for (; i < nIter; ++i) M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
        return;
    }
    bool gx = true, gy = true;
    for (int i = 0; i < nIter && (gx || gy); ++i) {
        if (gx) {
            M::step(x0, x1, x2, p, h);
            if (diverged(x0, maxValue)) gx = false;
        }
        if (gy) {
            M::step(y0, y1, y2, p, h);
            if (diverged(y0, maxValue)) gy = false;
        }
    }
}

template <class M, bool CHECK>
__global__ __launch_bounds__(256) void k_lle64(const LyapArgs A, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal];
    double x0, x1, x2;
    cell_setup(A.g, idx, v, x0, x1, x2);
    const typename M::P64 p = M::params(v);
    const double h = A.h, eps = A.eps;
    double z0 = 1, z1 = 0, z2 = 0;
    if (A.dirs) {
        const double* d = A.dirs + (long long)t * 3;
        z0 = d[0]; z1 = d[1]; z2 = d[2];
    }
    run<M, CHECK>(x0, x1, x2, p, h, A.skip, A.maxValue);
    double y0 = z0 * eps + x0, y1 = z1 * eps + x1, y2 = z2 * eps + x2;
    double result = 0;
    for (int b = 0; b < A.nBlocks; ++b) {
        run2<M, CHECK>(x0, x1, x2, y0, y1, y2, p, h, A.nNT, A.maxValue);
        double s = 0;
        s += (x0 - y0) * (x0 - y0);
        s += (x1 - y1) * (x1 - y1);
        s += (x2 - y2) * (x2 - y2);
        s = sqrt(s) / eps;
        result += log(s);
        if (s != 0) s = (1 / s);
        y0 = x0 - ((x0 - y0) * s);
        y1 = x1 - ((x1 - y1) * s);
        y2 = x2 - ((x2 - y2) * s);
    }
    out[t] = result / A.tMax;
}

template <class M, bool CHECK>
__global__ __launch_bounds__(256) void k_lle32(const LyapArgs A, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal];
    double s0, s1, s2;
    cell_setup(A.g, idx, v, s0, s1, s2);
    const typename M::P32 q = M::params32(v, A.h);
    const float mv = (float)A.maxValue;
    float x0 = (float)s0, x1 = (float)s1, x2 = (float)s2;
    float u0 = 1, u1 = 0, u2 = 0;
    if (A.dirs) {
        const double* d = A.dirs + (long long)t * 3;
        u0 = (float)d[0]; u1 = (float)d[1]; u2 = (float)d[2];
    }
    const float s = (float)A.eps;
    for (int i = 0; i < A.skip; ++i) {
        M::step32(x0, x1, x2, q);
        if (CHECK && fabsf(x0) > mv) break;
    }
    LogNormSum result;
    result.init();
    bool live = true;
    for (int b = 0; b < A.nBlocks; ++b) {
        if (live) {
#pragma unroll 2
            for (int i = 0; i < A.nNT; ++i) {
                const float x2o = x2;
                M::step32(x0, x1, x2, q);
                M::dev32(u0, u1, u2, x0, x2o, s, q);
                if (CHECK && fabsf(x0) > mv) { live = false; break; }
            }
        }
        const double d0 = u0, d1 = u1, d2 = u2;
        const double n2 = fma(d2, d2, fma(d1, d1, d0 * d0));
        result.add(n2);
        const double inv = n2 > 0.0 ? rsqrt(n2) : 1.0;
        u0 = (float)(d0 * inv);
        u1 = (float)(d1 * inv);
        u2 = (float)(d2 * inv);
    }
    out[t] = result.value() / A.tMax;
}

template <class M, bool CHECK>
__global__ __launch_bounds__(256) void k_ls64(const LyapArgs A, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal];
    double x0, x1, x2;
    cell_setup(A.g, idx, v, x0, x1, x2);
    const typename M::P64 p = M::params(v);
    const double h = A.h, eps = A.eps;
    double z[9], y[9], den[3], res[3] = {0, 0, 0};
    if (A.dirs) {
        const double* d = A.dirs + (long long)t * 9;
#pragma unroll
        for (int i = 0; i < 9; ++i) z[i] = d[i];
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) z[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    run<M, CHECK>(x0, x1, x2, p, h, A.skip, A.maxValue);
    gs3(z, y, den);
    double xs[3] = {x0, x1, x2};
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) y[j * 3 + i] = y[j * 3 + i] * eps + xs[i];

    for (int b = 0; b < A.nBlocks; ++b) {
        if (!CHECK) {
            for (int i = 0; i < A.nNT; ++i) {
                M::step(x0, x1, x2, p, h);
                M::step(y[0], y[1], y[2], p, h);
                M::step(y[3], y[4], y[5], p, h);
                M::step(y[6], y[7], y[8], p, h);
            }
        } else {
            run<M, true>(x0, x1, x2, p, h, A.nNT, A.maxValue);
            run<M, true>(y[0], y[1], y[2], p, h, A.nNT, A.maxValue);
            run<M, true>(y[3], y[4], y[5], p, h, A.nNT, A.maxValue);
            run<M, true>(y[6], y[7], y[8], p, h, A.nNT, A.maxValue);
        }
        const double xv[3] = {x0, x1, x2};
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = 0; l < 3; ++l) y[k * 3 + l] = y[k * 3 + l] - xv[l];
        gs3(y, z, den);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            res[k] += log(den[k] / eps);
#pragma unroll
            for (int j = 0; j < 3; ++j) y[k * 3 + j] = (double)(xv[j] + z[k * 3 + j] * eps);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) out[(long long)t * 3 + i] = res[i] / A.tMax;
}

// Batched gramSchmidtProcess for n <= 8 (one matrix per lane).
__global__ __launch_bounds__(64) void k_gs(const double* __restrict__ a, double* __restrict__ b,
                                           double* __restrict__ den, int n, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const double* A = a + (long long)t * n * n;
    double* B = b + (long long)t * n * n;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) B[j + i * n] = A[j + i * n];
        for (int j = 0; j < i; ++j) {
            double num = 0, dd = 0;
            for (int m = 0; m < n; ++m) {
                num += A[i * n + m] * B[j * n + m];
                dd += B[j * n + m] * B[j * n + m];
            }
            const double f = dd == 0 ? 0 : num / dd;
            for (int m = 0; m < n; ++m) B[i * n + m] -= f * B[j * n + m];
        }
    }
    for (int i = 0; i < n; ++i) {
        double d = 0;
        for (int j = 0; j < n; ++j) d += B[i * n + j] * B[i * n + j];
        d = sqrt(d);
        for (int j = 0; j < n; ++j) B[i * n + j] = d == 0 ? 0 : B[i * n + j] / d;
        if (den) den[(long long)t * n + i] = d;
    }
}

// --------------------------------------------------------- launchers

hipError_t launch_gen_lle(int model, const LyapArgs& A, bool check, double* out,
                          hipStream_t st);
hipError_t launch_gen_ls(int model, const LyapArgs& A, bool check, double* out, hipStream_t st);

hipError_t launch_lle(int model, int precision, const LyapArgs& A, bool check, double* out,
                      hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model == kMemristive)
        return precision == 1 ? hipErrorInvalidValue : launch_gen_lle(model, A, check, out, st);
    const dim3 blk(256), grd((A.nCells + 255) / 256);
#define LLE_CASE(KER, MOD, C) hipLaunchKernelGGL((KER<MOD, C>), grd, blk, 0, st, A, out)
    if (precision == 1) {
        if (model == kLinear) { if (check) LLE_CASE(k_lle32, Linear, true); else LLE_CASE(k_lle32, Linear, false); }
        else { if (check) LLE_CASE(k_lle32, Rossler, true); else LLE_CASE(k_lle32, Rossler, false); }
    } else {
        if (model == kLinear) { if (check) LLE_CASE(k_lle64, Linear, true); else LLE_CASE(k_lle64, Linear, false); }
        else { if (check) LLE_CASE(k_lle64, Rossler, true); else LLE_CASE(k_lle64, Rossler, false); }
    }
#undef LLE_CASE
    return hipGetLastError();
}

hipError_t launch_ls(int model, const LyapArgs& A, bool check, double* out, hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model == kMemristive) return launch_gen_ls(model, A, check, out, st);
    const dim3 blk(256), grd((A.nCells + 255) / 256);
#define LS_CASE(MOD, C) hipLaunchKernelGGL((k_ls64<MOD, C>), grd, blk, 0, st, A, out)
    if (model == kLinear) { if (check) LS_CASE(Linear, true); else LS_CASE(Linear, false); }
    else { if (check) LS_CASE(Rossler, true); else LS_CASE(Rossler, false); }
#undef LS_CASE
    return hipGetLastError();
}

hipError_t launch_gs(const double* a, double* b, double* den, int n, int count, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gs, dim3((count + 63) / 64), dim3(64), 0, st, a, b, den, n, count);
    return hipGetLastError();
}

}  // namespace bifhip
