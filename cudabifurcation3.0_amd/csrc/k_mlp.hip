// k_mlp.hip -- surrogate network kernels for gfx950 (BF16 MFMA, FP32 accumulate).
//
// The surrogate maps a parameter point (a, c) to the LLE and the peak-count
// class of the cell.  Network: 64 Fourier features -> 3 x 128 ReLU -> one
// 32-wide head (column 0: standardised LLE, columns 1..16: class logits,
// the rest zero).  Every product -- forward, input gradient and weight
// gradient -- runs through k_gemm, a v_mfma_f32_16x16x32_bf16 kernel whose
// operands are addressed with (row, k) strides, so transposed operands need
// no copies.  Weights are kept in FP32 (Adam state) with BF16 copies for
// the matrix cores.
#include "bifhip_dev.h"

namespace bifhip {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFeat = 64;      // Fourier features
constexpr int kHid = 128;      // hidden width
constexpr int kHead = 32;      // head width (1 LLE + 16 classes + chaos + padding)
constexpr int kClasses = 16;
constexpr int kChaosCol = 1 + kClasses;   // chaos logit column

enum Epi : int {
    kEpiReluBf16 = 0,   // out(bf16) = relu(acc + bias)
    kEpiBiasF32 = 1,    // out(f32)  = acc + bias
    kEpiMaskBf16 = 2,   // out(bf16) = acc * (mask > 0)
    kEpiF32 = 3,        // out(f32)  = acc
};

// C[M x N] = sum_k A(m, k) B(k, n) with A(m,k) = A[m*sAm + k*sAk],
// B(k,n) = B[k*sBk + n*sBn].  Block 256 = 4 waves, block tile 64 x 64,
// each wave a 32 x 32 quadrant = 2 x 2 tiles of 16 x 16, K in steps of 32.
// MFMA 16x16x32 operand layout (one lane = 8 consecutive k):
//   A: row (lane & 15), k = 8 * (lane >> 4) + e
//   B: col (lane & 15), k = 8 * (lane >> 4) + e
//   D: col (lane & 15), row = 4 * (lane >> 4) + e   (e < 4)
// Split-K: block z sums k in [z*kChunk, min(K, (z+1)*kChunk)) and writes its
// partial to C + z * splitStride (kEpiF32 only); k_sum_splits adds them.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16* __restrict__ A,
                                              long long sAm, long long sAk,
                                              const bf16* __restrict__ B, long long sBk,
                                              long long sBn, const float* __restrict__ bias,
                                              const bf16* __restrict__ mask, long long sMm,
                                              void* __restrict__ Cbase, long long sCm,
                                              int kChunk, long long splitStride)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int li = lane & 15;
    const int kq = (lane >> 4) * 8;
    const int kBegin = blockIdx.z * kChunk;
    const int kEnd = min(K, kBegin + kChunk);
    void* C = EPI == kEpiF32 ? (void*)(static_cast<float*>(Cbase) + blockIdx.z * splitStride)
                             : Cbase;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

    const bool aVec = (sAk == 1) && (K % 8 == 0) && (sAm % 8 == 0);
    const bool bVec = (sBk == 1) && (K % 8 == 0) && (sBn % 8 == 0);
    for (int k0 = kBegin; k0 < kEnd; k0 += 32) {
        bf16x8 af[2], bfr[2];
        const int kb = k0 + kq;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int r = m0 + 16 * t + li;
            if (aVec && r < M && kb + 8 <= kEnd) {
                af[t] = *reinterpret_cast<const bf16x8*>(A + (long long)r * sAm + kb);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = kb + e;
                    af[t][e] = (r < M && k < kEnd) ? A[(long long)r * sAm + (long long)k * sAk]
                                                   : (bf16)0.0f;
                }
            }
            const int c = n0 + 16 * t + li;
            if (bVec && c < N && kb + 8 <= kEnd) {
                bfr[t] = *reinterpret_cast<const bf16x8*>(B + (long long)c * sBn + kb);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = kb + e;
                    bfr[t][e] = (c < N && k < kEnd) ? B[(long long)k * sBk + (long long)c * sBn]
                                                    : (bf16)0.0f;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = n0 + 16 * j + li;
            if (c >= N) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = m0 + 16 * i + (lane >> 4) * 4 + e;
                if (r >= M) continue;
                float v = acc[i][j][e];
                if (EPI == kEpiReluBf16) {
                    v += bias ? bias[c] : 0.0f;
                    static_cast<bf16*>(C)[(long long)r * sCm + c] = (bf16)(v > 0 ? v : 0.0f);
                } else if (EPI == kEpiBiasF32) {
                    v += bias ? bias[c] : 0.0f;
                    static_cast<float*>(C)[(long long)r * sCm + c] = v;
                } else if (EPI == kEpiMaskBf16) {
                    const float mk = (float)mask[(long long)r * sMm + c];
                    static_cast<bf16*>(C)[(long long)r * sCm + c] = (bf16)(mk > 0 ? v : 0.0f);
                } else {
                    static_cast<float*>(C)[(long long)r * sCm + c] = v;
                }
            }
        }
    }
}

// Fourier features of lattice cells: cell = list ? list[i] : base + i on an
// n x n lattice; p = (ix/(n-1), iy/(n-1)); x = [sin(2 pi B p), cos(2 pi B p)].
__global__ void k_features(int rows, int n, const int* __restrict__ list, long long base,
                           const float* __restrict__ Bff, bf16* __restrict__ X)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = t / (kFeat / 2);
    const int f = t % (kFeat / 2);
    if (r >= rows) return;
    const long long cell = list ? (long long)list[r] : base + r;
    const float inv = n > 1 ? 1.0f / (float)(n - 1) : 0.0f;
    const float u = (float)(cell % n) * inv;
    const float v = (float)(cell / n) * inv;
    const float ph = 6.283185307179586f * (Bff[2 * f] * u + Bff[2 * f + 1] * v);
    float s, c;
    sincosf(ph, &s, &c);
    X[(long long)r * kFeat + f] = (bf16)s;
    X[(long long)r * kFeat + kFeat / 2 + f] = (bf16)c;
}

// Loss gradient of one batch: softmax cross-entropy over columns 1..16 and
// w_lle * squared error on column 0 (skipped where the target is not
// finite).  dOut = dLoss/dOut (mean over the batch) in BF16; loss summed
// into loss[0] (CE) and loss[1] (MSE).
// Column 17 is the chaos logit: binary cross-entropy against
// (standardised LLE > chaosStd), a non-finite LLE counting as not chaotic;
// its loss goes to loss[2].
__global__ void k_loss_grad(int rows, const int* __restrict__ list, const float* __restrict__ out,
                            const int* __restrict__ clsAll, const float* __restrict__ lleAll,
                            float wLle, float chaosStd, bf16* __restrict__ dOut,
                            float* __restrict__ loss)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int src = list[r];
    const int y = clsAll[src];
    const float target = lleAll[src];
    const float* o = out + (long long)r * kHead;
    float mx = o[1];
    for (int k = 2; k <= kClasses; ++k) mx = fmaxf(mx, o[k]);
    float se = 0;
    for (int k = 1; k <= kClasses; ++k) se += __expf(o[k] - mx);
    const float inv = 1.0f / (float)rows;
    bf16* d = dOut + (long long)r * kHead;
    for (int k = 1; k <= kClasses; ++k) {
        const float p = __expf(o[k] - mx) / se;
        d[k] = (bf16)((p - (k - 1 == y ? 1.0f : 0.0f)) * inv);
    }
    float ce = -(o[1 + y] - mx - __logf(se));
    float mse = 0, g0 = 0;
    if (isfinite(target)) {
        const float e = o[0] - target;
        mse = e * e;
        g0 = 2.0f * wLle * e * inv;
    }
    d[0] = (bf16)g0;
    const float z = o[kChaosCol];
    const float tc = (target > chaosStd) ? 1.0f : 0.0f;     // NaN -> 0
    const float pc = 1.0f / (1.0f + __expf(-z));
    d[kChaosCol] = (bf16)((pc - tc) * inv);
    const float bce = fmaxf(z, 0.0f) - z * tc + __logf(1.0f + __expf(-fabsf(z)));
    for (int k = kChaosCol + 1; k < kHead; ++k) d[k] = (bf16)0.0f;
    atomicAdd(&loss[0], ce * inv);
    atomicAdd(&loss[1], mse * inv);
    atomicAdd(&loss[2], bce * inv);
}

// Sum of split partials in a fixed order: out[i] = sum_z part[z][i].
__global__ void k_sum_splits(int n, int splits, const float* __restrict__ part,
                             float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0;
    for (int z = 0; z < splits; ++z) s += part[(long long)z * n + i];
    out[i] = s;
}

// Split partials of the four bias gradients (column sums of the layers'
// output gradients) in one launch: blockIdx.y picks the layer, blockIdx.x a
// group of 64 columns, blockIdx.z the row chunk; the partial of split z goes
// to gpart[z * stride + offset + column], next to the weight-gradient
// partials, and k_adam adds them in split order.  Block (0,0,0) also
// advances the optimiser's step counter (read by the following k_adam).
struct BiasSrc {
    const bf16* g[4];
    int cols[4];
    int off[4];
};

__global__ __launch_bounds__(256) void k_bias_partials(int rows, int rChunk, BiasSrc src,
                                                       float* __restrict__ gpart,
                                                       long long stride, int* __restrict__ step)
{
    __shared__ float s_part[4][64];
    const int layer = blockIdx.y;
    const int N = src.cols[layer];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int rs = threadIdx.x >> 6;
    const int r0 = blockIdx.z * rChunk, r1 = min(rows, r0 + rChunk);
    float s = 0;
    if (c < N) {
        const bf16* g = src.g[layer];
        for (int r = r0 + rs; r < r1; r += 4) s += (float)g[(long long)r * N + c];
    }
    s_part[rs][threadIdx.x & 63] = s;
    __syncthreads();
    if (rs == 0 && c < N)
        gpart[(long long)blockIdx.z * stride + src.off[layer] + c] =
            ((s_part[0][c & 63] + s_part[1][c & 63]) + s_part[2][c & 63]) + s_part[3][c & 63];
    if (step && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0)
        *step += 1;
}

// Adam step on the FP32 master parameters with the gradient summed from its
// split partials (gpart[z * stride + i], z < splits, fixed order); refreshes
// the BF16 copies of the weight matrices and keeps padded head rows at zero.
__global__ void k_adam(int n, float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ gpart, int splits, long long stride,
                       const unsigned char* __restrict__ keep, const int* __restrict__ bfIndex,
                       bf16* __restrict__ wbf, float lr, float b1, float b2, float eps,
                       const int* __restrict__ step)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // bias corrections of step t (1-based) read from device memory, so a
    // captured graph replays correctly step after step
    const float t = (float)(*step);
    const float c1 = 1.0f / (1.0f - powf(b1, t));
    const float c2 = 1.0f / (1.0f - powf(b2, t));
    if (!keep[i]) {
        p[i] = 0;
    } else {
        float gi = 0;
        for (int z = 0; z < splits; ++z) gi += gpart[(long long)z * stride + i];
        const float mi = b1 * m[i] + (1 - b1) * gi;
        const float vi = b2 * v[i] + (1 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
    const int bi = bfIndex[i];
    if (bi >= 0) wbf[bi] = (bf16)p[i];
}

__global__ void k_fill_f64(int n, double* p, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(int n, float* p, float v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Inference post-processing of head outputs: class = argmax of the logits,
// chaos regime from the chaos logit, LLE = column 0 de-standardised and
// kept on the side of chaosThr the chaos logit predicts (the regime of a
// predicted cell is the classifier's; the regression supplies the value).
// Class confidence = the class's softmax probability; chaos confidence =
// the chaos classifier's probability of its answer.  With confChaos null,
// conf = min of the two; else conf = class confidence and confChaos =
// chaos confidence.
__global__ void k_infer_post(int rows, const float* __restrict__ out, float lleMean,
                             float lleStd, float chaosThr, unsigned char* __restrict__ cls,
                             unsigned char* __restrict__ chaos, float* __restrict__ conf,
                             float* __restrict__ confChaos, float* __restrict__ lle)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float* o = out + (long long)r * kHead;
    int best = 0;
    float mx = o[1];
    for (int k = 2; k <= kClasses; ++k)
        if (o[k] > mx) { mx = o[k]; best = k - 1; }
    float se = 0;
    for (int k = 1; k <= kClasses; ++k) se += __expf(o[k] - mx);
    const float pc = 1.0f / (1.0f + __expf(-o[kChaosCol]));
    const bool ch = pc > 0.5f;
    float l = o[0] * lleStd + lleMean;
    const float margin = fmaxf(1e-6f, fabsf(chaosThr) * 1e-4f);
    if (ch && !(l > chaosThr)) l = chaosThr + margin;
    if (!ch && l > chaosThr) l = chaosThr - margin;
    const float cc = ch ? pc : 1.0f - pc;
    const float ck = 1.0f / se;
    cls[r] = (unsigned char)best;
    chaos[r] = ch ? 1 : 0;
    if (confChaos) {
        conf[r] = ck;
        confChaos[r] = cc;
    } else {
        conf[r] = fminf(ck, cc);
    }
    lle[r] = l;
}

// Wave-aggregated append of the lanes with `take` to list[]: one atomic per
// wavefront, lanes in lane order after the wave's base.
__device__ __forceinline__ void wave_append(bool take, int value, int* __restrict__ list,
                                            int* __restrict__ count)
{
    const uint64_t m = __ballot(take);
    if (!m) return;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __builtin_ctzll(m);
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Cells to re-solve.  A cell needs its peak-count class when the class
// confidence is below confThr or a 4-neighbour has another class, and its
// LLE regime when the chaos confidence is below chaosConfThr or a
// 4-neighbour has another regime.  With confChaos null (conf = min of both,
// compared with confThr) every such cell is re-solved in full.  Otherwise
// cells that need their regime are re-solved in full (flag 1, listFull)
// and cells that only need their class get the peak count alone (flag 2,
// listPeak); with `partial` false both go to listFull.
__global__ void k_flag(int n, const unsigned char* __restrict__ cls,
                       const unsigned char* __restrict__ chaos, const float* __restrict__ conf,
                       const float* __restrict__ confChaos, float confThr, float chaosConfThr,
                       bool partial, int* __restrict__ listFull, int* __restrict__ countFull,
                       int* __restrict__ listPeak, int* __restrict__ countPeak,
                       unsigned char* __restrict__ flag)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = idx < (long long)n * n;   // no early return: the appends are wave-wide
    bool fc = false, fh = false;
    if (live) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        const int c0 = cls[idx];
        const int ch0 = chaos[idx];
        fc = conf[idx] < confThr;
        fh = confChaos ? confChaos[idx] < chaosConfThr : fc;
        const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ii = i + di[q], jj = j + dj[q];
            if (ii < 0 || jj < 0 || ii >= n || jj >= n) continue;
            const long long o = (long long)jj * n + ii;
            fc |= cls[o] != c0;
            fh |= chaos[o] != ch0;
        }
        if (!confChaos || !partial) fh = fh || fc;
        flag[idx] = fh ? 1 : (fc ? 2 : 0);
    }
    wave_append(live && fh, (int)idx, listFull, countFull);
    wave_append(live && !fh && fc, (int)idx, listPeak, countPeak);
}

// Write re-solved cells into the output maps (LLE only when lleIn is given).
__global__ void k_scatter(int n, const int* __restrict__ list, const int* __restrict__ peak,
                          const double* __restrict__ lleIn, int* __restrict__ peakMap,
                          double* __restrict__ lleMap)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = list[t];
    peakMap[c] = peak[t];
    if (lleIn) lleMap[c] = lleIn[t];
}

// Surrogate prediction into the output maps (class k -> peak count k+1).
__global__ void k_fill_pred(long long n, const unsigned char* __restrict__ cls,
                            const float* __restrict__ lle, int* __restrict__ peakMap,
                            double* __restrict__ lleMap)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    peakMap[i] = (int)cls[i] + 1;
    lleMap[i] = (double)lle[i];
}

// Scores of the surrogate on re-solved cells, accumulated per group:
// 0 = flagged and re-solved in full (entries [0, nFull) of list),
// 1 = validation sample (entries [nFull, nTot)), 2 = peak count only
// (plist).  acc[g*5 + 0..4] = cells, class right, regime right, finite
// LLE cells, sum |LLE error| (as doubles).
__global__ __launch_bounds__(256) void k_score(int nFull, int nTot, const int* __restrict__ list,
                                               const int* __restrict__ rpeak,
                                               const double* __restrict__ rlle, int nPeak,
                                               const int* __restrict__ plist,
                                               const int* __restrict__ ppeak,
                                               const unsigned char* __restrict__ pcls,
                                               const unsigned char* __restrict__ pch,
                                               const float* __restrict__ plle, double chaosThr,
                                               double* __restrict__ acc)
{
    __shared__ double s_acc[15];
    if (threadIdx.x < 15) s_acc[threadIdx.x] = 0;
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nTot) {
        const int g = t < nFull ? 0 : 1;
        const int cell = list[t];
        const int tc = min(kClasses, max(1, rpeak[t])) - 1;
        const double tl = rlle[t];
        const bool tch = tl > chaosThr;
        const bool okc = tc == (int)pcls[cell];
        const bool okr = okc && (tch == (pch[cell] != 0));
        atomicAdd(&s_acc[g * 5 + 0], 1.0);
        if (okc) atomicAdd(&s_acc[g * 5 + 1], 1.0);
        if (okr) atomicAdd(&s_acc[g * 5 + 2], 1.0);
        if (isfinite(tl)) {
            atomicAdd(&s_acc[g * 5 + 3], 1.0);
            atomicAdd(&s_acc[g * 5 + 4], fabs(tl - (double)plle[cell]));
        }
    }
    if (t < nPeak) {
        const int cell = plist[t];
        const int tc = min(kClasses, max(1, ppeak[t])) - 1;
        atomicAdd(&s_acc[10], 1.0);
        if (tc == (int)pcls[cell]) atomicAdd(&s_acc[11], 1.0);
    }
    __syncthreads();
    if (threadIdx.x < 15 && s_acc[threadIdx.x] != 0) atomicAdd(&acc[threadIdx.x], s_acc[threadIdx.x]);
}

// ------------------------------------------------------------ launchers

// splits > 1 (kEpiF32 only): K is cut into `splits` chunks written to
// C + z*M*sCm, to be added by launch_sum_splits.
hipError_t launch_gemm(int epi, int M, int N, int K, const bf16* A, long long sAm, long long sAk,
                       const bf16* B, long long sBk, long long sBn, const float* bias,
                       const bf16* mask, long long sMm, void* C, long long sCm, hipStream_t st,
                       int splits = 1, long long splitStride = -1)
{
    if (epi != kEpiF32) splits = 1;
    // chunks are whole 32-wide k steps; trailing splits may be empty and
    // then write zero partials, so the caller always sums `splits` of them
    int kChunk = (K + splits - 1) / splits;
    kChunk = (kChunk + 31) / 32 * 32;
    const dim3 grd((N + 63) / 64, (M + 63) / 64, splits), blk(256);
    if (splitStride < 0) splitStride = (long long)M * sCm;
#define GEMM_CASE(E)                                                                          \
    hipLaunchKernelGGL((k_gemm<E>), grd, blk, 0, st, M, N, K, A, sAm, sAk, B, sBk, sBn, bias, \
                       mask, sMm, C, sCm, kChunk, splitStride)
    switch (epi) {
        case kEpiReluBf16: GEMM_CASE(kEpiReluBf16); break;
        case kEpiBiasF32: GEMM_CASE(kEpiBiasF32); break;
        case kEpiMaskBf16: GEMM_CASE(kEpiMaskBf16); break;
        default: GEMM_CASE(kEpiF32); break;
    }
#undef GEMM_CASE
    return hipGetLastError();
}

hipError_t launch_features(int rows, int n, const int* list, long long base, const float* Bff,
                           bf16* X, hipStream_t st)
{
    const long long th = (long long)rows * (kFeat / 2);
    hipLaunchKernelGGL(k_features, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, st, rows, n,
                       list, base, Bff, X);
    return hipGetLastError();
}

hipError_t launch_loss_grad(int rows, const int* list, const float* out, const int* cls,
                            const float* lle, float wLle, float chaosStd, bf16* dOut, float* loss,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_loss_grad, dim3((rows + 255) / 256), dim3(256), 0, st, rows, list, out,
                       cls, lle, wLle, chaosStd, dOut, loss);
    return hipGetLastError();
}

hipError_t launch_sum_splits(int n, int splits, const float* part, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_splits, dim3((n + 255) / 256), dim3(256), 0, st, n, splits, part,
                       out);
    return hipGetLastError();
}

hipError_t launch_fill_f64(int n, double* p, double v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_f64, dim3((n + 255) / 256), dim3(256), 0, st, n, p, v);
    return hipGetLastError();
}

hipError_t launch_fill_f32(int n, float* p, float v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_f32, dim3((n + 255) / 256), dim3(256), 0, st, n, p, v);
    return hipGetLastError();
}

// split partials of the four bias gradients (g[l] is [rows x cols[l]]) into
// gpart[z * stride + off[l] + c]; advances *step when step is not null
hipError_t launch_bias_partials(int rows, int splits, const bf16* const g[4], const int cols[4],
                                const int off[4], float* gpart, long long stride, int* step,
                                hipStream_t st)
{
    BiasSrc src;
    int maxCols = 0;
    for (int l = 0; l < 4; ++l) {
        src.g[l] = g[l];
        src.cols[l] = cols[l];
        src.off[l] = off[l];
        maxCols = max(maxCols, cols[l]);
    }
    const int rChunk = (rows + splits - 1) / splits;
    hipLaunchKernelGGL(k_bias_partials, dim3((maxCols + 63) / 64, 4, splits), dim3(256), 0, st,
                       rows, rChunk, src, gpart, stride, step);
    return hipGetLastError();
}

hipError_t launch_adam(int n, float* p, float* m, float* v, const float* gpart, int splits,
                       long long stride, const unsigned char* keep, const int* bfIndex, bf16* wbf,
                       float lr, float b1, float b2, float eps, const int* step, hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, n, p, m, v, gpart, splits,
                       stride, keep, bfIndex, wbf, lr, b1, b2, eps, step);
    return hipGetLastError();
}

hipError_t launch_infer_post(int rows, const float* out, float mean, float sd, float chaosThr,
                             unsigned char* cls, unsigned char* chaos, float* conf,
                             float* confChaos, float* lle, hipStream_t st)
{
    hipLaunchKernelGGL(k_infer_post, dim3((rows + 255) / 256), dim3(256), 0, st, rows, out, mean,
                       sd, chaosThr, cls, chaos, conf, confChaos, lle);
    return hipGetLastError();
}

hipError_t launch_flag(int n, const unsigned char* cls, const unsigned char* chaos,
                       const float* conf, const float* confChaos, float confThr,
                       float chaosConfThr, bool partial, int* listFull, int* countFull,
                       int* listPeak, int* countPeak, unsigned char* flag, hipStream_t st)
{
    const long long cells = (long long)n * n;
    hipLaunchKernelGGL(k_flag, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, n, cls,
                       chaos, conf, confChaos, confThr, chaosConfThr, partial, listFull,
                       countFull, listPeak, countPeak, flag);
    return hipGetLastError();
}

hipError_t launch_scatter(int n, const int* list, const int* peak, const double* lle,
                          int* peakMap, double* lleMap, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, st, n, list, peak, lle,
                       peakMap, lleMap);
    return hipGetLastError();
}

hipError_t launch_fill_pred(long long n, const unsigned char* cls, const float* lle,
                            int* peakMap, double* lleMap, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_pred, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, cls,
                       lle, peakMap, lleMap);
    return hipGetLastError();
}

}  // namespace bifhip

namespace bifhip {
hipError_t launch_score(int nFull, int nTot, const int* list, const int* rpeak, const double* rlle,
                        int nPeak, const int* plist, const int* ppeak, const unsigned char* pcls,
                        const unsigned char* pch, const float* plle, double chaosThr,
                        double* acc, hipStream_t st)
{
    const int n = max(nTot, nPeak);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_score, dim3((n + 255) / 256), dim3(256), 0, st, nFull, nTot, list, rpeak,
                       rlle, nPeak, plist, ppeak, pcls, pch, plle, chaosThr, acc);
    return hipGetLastError();
}
}  // namespace bifhip
