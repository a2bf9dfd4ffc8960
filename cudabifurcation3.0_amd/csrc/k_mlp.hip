// k_mlp.hip -- surrogate network kernels for gfx950 (BF16 MFMA, FP32 accumulate).
//
// The surrogate maps a parameter point (a, c) to the LLE and the peak-count
// class of the cell.  Network: 64 Fourier features -> 3 x 128 ReLU -> one
// 32-wide head (column 0: standardised LLE, columns 1..16: class logits,
// column 17: chaos logit, the rest zero).  Every product is a
// v_mfma_f32_16x16x32_bf16 tile with FP32 accumulation.  Inference is one
// k_infer_fused launch (features, layers and head of 64 rows per workgroup
// with the activations in LDS); a training step is one k_train_fused launch
// (forward, loss gradient, backward, gradient partials per workgroup) and
// one k_adam launch.
// Weights are kept in FP32 (Adam state) with BF16 copies, and transposed
// BF16 copies for the backward products, for the matrix cores.
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

// Loss gradient of one row of a batch of 1/inv rows: softmax cross-entropy
// over columns 1..16 and w_lle * squared error on column 0 (skipped where
// the target is not finite); column 17 is the chaos logit, binary
// cross-entropy against (standardised LLE > chaosStd), a non-finite LLE
// counting as not chaotic.  d = dLoss/dOut (mean over the batch) in BF16;
// the row's CE, squared error and BCE in ce, mse, bce.
__device__ __forceinline__ void loss_row(const float* o, int y, float target, float wLle,
                                         float chaosStd, float inv, bf16 (&d)[kHead], float& ce,
                                         float& mse, float& bce)
{
    float mx = o[1];
    for (int k = 2; k <= kClasses; ++k) mx = fmaxf(mx, o[k]);
    float se = 0;
    for (int k = 1; k <= kClasses; ++k) se += __expf(o[k] - mx);
    for (int k = 1; k <= kClasses; ++k) {
        const float p = __expf(o[k] - mx) / se;
        d[k] = (bf16)((p - (k - 1 == y ? 1.0f : 0.0f)) * inv);
    }
    ce = -(o[1 + y] - mx - __logf(se));
    mse = 0;
    float g0 = 0;
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
    bce = fmaxf(z, 0.0f) - z * tc + __logf(1.0f + __expf(-fabsf(z)));
    for (int k = kChaosCol + 1; k < kHead; ++k) d[k] = (bf16)0.0f;
}

// sum_z part[z * stride + i] over z < splits, added in z order; the loads
// go out sixteen (then eight) at a time so their latencies overlap (one
// load in flight at a time made the sum wait for every partial in turn;
// sixteen against eight: 20 epochs of 2048-row steps 15.75 against 15.98 ms)
__device__ __forceinline__ float sum_partials(const float* __restrict__ part, int splits,
                                              long long stride, int i)
{
    float s = 0;
    int z = 0;
    for (; z + 16 <= splits; z += 16) {
        float t[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) t[u] = part[(long long)(z + u) * stride + i];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += t[u];
    }
    for (; z + 8 <= splits; z += 8) {
        float t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = part[(long long)(z + u) * stride + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += t[u];
    }
    for (; z < splits; ++z) s += part[(long long)z * stride + i];
    return s;
}

// Sum of split partials in a fixed order: out[i] = sum_z part[z][i].
__global__ void k_sum_splits(int n, int splits, const float* __restrict__ part,
                             float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sum_partials(part, splits, n, i);
}

// Adam step on the FP32 master parameters with the gradient summed from its
// split partials (gpart[z * stride + i], z < splits, fixed order); refreshes
// the BF16 copies of the weight matrices and keeps padded head rows at zero.
__global__ void k_adam(int n, float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ gpart, int splits, long long stride,
                       const unsigned char* __restrict__ keep, const int* __restrict__ bfIndex,
                       bf16* __restrict__ wbf, const int* __restrict__ bfIndexT,
                       bf16* __restrict__ wbfT, float lr, float b1, float b2, float eps,
                       const int* __restrict__ step)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // every load is issued before the first wait: the entry's flags and
    // state, then the partials (summed for held entries too, and unused)
    const bool kept = keep[i] != 0;
    const int bi = bfIndex[i];
    const int ti = bfIndexT ? bfIndexT[i] : -1;
    const float m0 = m[i], v0 = v[i], p0 = p[i];
    const float gi = sum_partials(gpart, splits, stride, i);
    // bias corrections of step t (1-based) read from device memory, so a
    // captured graph replays correctly step after step
    const float t = (float)(*step);
    const float c1 = 1.0f / (1.0f - powf(b1, t));
    const float c2 = 1.0f / (1.0f - powf(b2, t));
    float pn = 0;
    if (kept) {
        const float mi = b1 * m0 + (1 - b1) * gi;
        const float vi = b2 * v0 + (1 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        pn = p0 - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
    p[i] = pn;
    if (bi >= 0) wbf[bi] = (bf16)pn;
    if (ti >= 0) wbfT[ti] = (bf16)pn;
}

// ------------------------------------------------------ fused training step
//
// One workgroup (4 waves) takes kRows rows of a batch through a whole
// training step: Fourier features, the three ReLU layers and the head, the
// loss gradient (loss_row), the backward pass, and the weight- and
// bias-gradient sums over its rows, written as partial kb of the gradient
// to gpart[kb * stride + parameter]; k_adam adds the partials in block
// order.  Activations stay in LDS in two layouts, row-major [row][unit] and
// transposed [unit][row] (rows padded by 8 against bank conflicts), so every
// MFMA operand is one 16-byte read:
//   forward   H^T  = W   X^T     A = W (BF16, global)    B = X   row-major
//   backward  dX^T = W^T dY^T    A = W^T (BF16 copy)     B = dY  row-major
//   weights   dW   = dY^T X      A = dY^T transposed     B = X^T transposed
// Each D tile (4 consecutive units of one row per lane) is stored to both
// layouts.  MFMA 16x16x32 operand layout (one lane = 8 consecutive k):
//   A: row (lane & 15), k = 8 * (lane >> 4) + e
//   B: col (lane & 15), k = 8 * (lane >> 4) + e
//   D: col (lane & 15), row = 4 * (lane >> 4) + e   (e < 4)  The forward products run the same MFMA sequence as k_gemm
// (K ascending in steps of 32), so the activations equal the unfused path's.
// Rows past `rows` enter with zero features and a zero loss gradient and
// add nothing.
constexpr int kRows = 64;
// k_train_fused: eight waves per 64-row workgroup, one 16-unit tile each
constexpr int kTrainThreads = 512;
constexpr int kTr = kRows + 8;          // transposed row length
constexpr int kLdX = kFeat + 8;
constexpr int kLdH = kHid + 8;
constexpr int kLdO = kHead + 8;
// FP32 parameter offsets (row-major [out][in] matrices, then the bias)
constexpr int pW1 = 0, pB1 = pW1 + kHid * kFeat, pW2 = pB1 + kHid, pB2 = pW2 + kHid * kHid;
constexpr int pW3 = pB2 + kHid, pB3 = pW3 + kHid * kHid, pWh = pB3 + kHid, pBh = pWh + kHead * kHid;
// BF16 copies: wbf = W1 W2 W3 Wh; wbfT = W2^T W3^T Wh^T
constexpr int qW1 = 0, qW2 = qW1 + kHid * kFeat, qW3 = qW2 + kHid * kHid, qWh = qW3 + kHid * kHid;
constexpr int tW2 = 0, tW3 = tW2 + kHid * kHid, tWh = tW3 + kHid * kHid;

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// D tile (units u0..u0+3 of row r, one lane) to the row-major layout and,
// when tr is given, the transposed one
__device__ __forceinline__ void store_tile(bf16* rm, int ld, bf16* tr, int u0, int r,
                                           const float (&v)[4])
{
    bf16x4 q;
#pragma unroll
    for (int e = 0; e < 4; ++e) q[e] = (bf16)v[e];
    *reinterpret_cast<bf16x4*>(rm + r * ld + u0) = q;
    if (tr) {
#pragma unroll
        for (int e = 0; e < 4; ++e) tr[(u0 + e) * kTr + r] = q[e];
    }
}

// Fourier features of kRows rows (k_features' arithmetic) into the
// row-major layout and, when xt is given, the transposed one; cell of row r
// = list ? list[row0 + r] : base + row0 + r; rows from nrows on are zero.
// Each of the workgroup's NT threads takes one row (its cell read and split
// into lattice coordinates once) and every (NT / kRows)-th feature of it.
template <int NT>
__device__ __forceinline__ void features_rows(const int* __restrict__ list, long long base,
                                              long long row0, int nrows, int nLat,
                                              const float* __restrict__ Bff, bf16* x, bf16* xt)
{
    constexpr int kFeatThreadsPerRow = NT / kRows;
    static_assert(kFeatThreadsPerRow * kRows == NT, "features: NT threads per workgroup");
    const float inv = nLat > 1 ? 1.0f / (float)(nLat - 1) : 0.0f;
    const int r = threadIdx.x / kFeatThreadsPerRow;
    const bool live = r < nrows;
    float u = 0.0f, v = 0.0f;
    if (live) {
        const long long cell = list ? (long long)list[row0 + r] : base + row0 + r;
        if (cell <= 0x7fffffffLL) {
            const int c = (int)cell;
            u = (float)(c % nLat) * inv;
            v = (float)(c / nLat) * inv;
        } else {
            u = (float)(cell % nLat) * inv;
            v = (float)(cell / nLat) * inv;
        }
    }
    for (int f = threadIdx.x % kFeatThreadsPerRow; f < kFeat / 2; f += kFeatThreadsPerRow) {
        float sv = 0.0f, cv = 0.0f;
        if (live) {
            const float ph = 6.283185307179586f * (Bff[2 * f] * u + Bff[2 * f + 1] * v);
            sincosf(ph, &sv, &cv);
        }
        const bf16 bs = (bf16)sv, bc = (bf16)cv;
        x[r * kLdX + f] = bs;
        x[r * kLdX + kFeat / 2 + f] = bc;
        if (xt) {
            xt[f * kTr + r] = bs;
            xt[(kFeat / 2 + f) * kTr + r] = bc;
        }
    }
}

// head outputs (FP32, + bias) of the kRows rows held row-major in h3, one
// 16 x 16 tile per wave of eight: row tile (wave & 3), unit tile (wave >> 2)
__device__ __forceinline__ void head_tile(const bf16* __restrict__ Wh, const float* __restrict__ bh,
                                          const bf16* h3, float* out, int ldOut, int wave, int li,
                                          int kq, int rq)
{
    const int rt = wave & 3, ut = wave >> 2;
    f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int k0 = 0; k0 < kHid; k0 += 32) {
        const bf16x8 x = ld8(h3 + (rt * 16 + li) * kLdH + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(Wh + (ut * 16 + li) * kHid + k0 + kq), x,
                                                      acc, 0, 0, 0);
    }
    const int r = rt * 16 + li;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int u = ut * 16 + rq + e;
        out[r * ldOut + u] = acc[e] + bh[u];
    }
}

// hidden layer forward: out = relu(W in^T + b), W [kHid][K] (global BF16);
// this wave's unit tiles are NI * wave .. NI * wave + NI - 1, all row tiles
template <int K, int NI>
__device__ __forceinline__ void fwd_hidden(const bf16* __restrict__ W, const float* __restrict__ b,
                                           const bf16* in, int ldIn, bf16* outRm, bf16* outTr,
                                           int wave, int li, int kq, int rq)
{
    f32x4 acc[NI][4];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a[NI], x[4];
#pragma unroll
        for (int i = 0; i < NI; ++i) a[i] = ld8(W + ((NI * wave + i) * 16 + li) * K + k0 + kq);
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = ld8(in + (j * 16 + li) * ldIn + k0 + kq);
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], x[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int u0 = (NI * wave + i) * 16 + rq;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float t = acc[i][j][e] + b[u0 + e];
                v[e] = t > 0 ? t : 0.0f;
            }
            store_tile(outRm, kLdH, outTr, u0, j * 16 + li, v);
        }
    }
}

// input gradient of a hidden layer, masked by the layer input's ReLU:
// acc(k, r) = sum_m WT[k][m] dY[r][m] for this wave's NI x 4 tiles
template <int M, int NI>
__device__ __forceinline__ void bwd_input(const bf16* __restrict__ WT, const bf16* dY, int ldY,
                                          f32x4 (&acc)[NI][4], int wave, int li, int kq)
{
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int m0 = 0; m0 < M; m0 += 32) {
        bf16x8 a[NI], y[4];
#pragma unroll
        for (int i = 0; i < NI; ++i) a[i] = ld8(WT + ((NI * wave + i) * 16 + li) * M + m0 + kq);
#pragma unroll
        for (int j = 0; j < 4; ++j) y[j] = ld8(dY + (j * 16 + li) * ldY + m0 + kq);
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], y[j], acc[i][j], 0, 0, 0);
    }
}

// mask by (h > 0) from the row-major activation, then store both layouts
template <int NI>
__device__ __forceinline__ void store_masked(const f32x4 (&acc)[NI][4], const bf16* hRm,
                                             bf16* outRm, bf16* outTr, int wave, int li, int rq)
{
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int u0 = (NI * wave + i) * 16 + rq;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = j * 16 + li;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                v[e] = (float)hRm[r * kLdH + u0 + e] > 0 ? acc[i][j][e] : 0.0f;
            store_tile(outRm, kLdH, outTr, u0, r, v);
        }
    }
}

// weight gradient dW[M x N] (+)= sum over the kRows rows of dY^T X from the
// transposed layouts: tiles (i0 + i, j0 + j), i < TI, j < TJ
template <int TI, int TJ>
__device__ __forceinline__ void wgrad(const bf16* dYT, const bf16* XT, int i0, int j0,
                                      float* __restrict__ g, int ldG, int li, int kq, int rq)
{
    f32x4 acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int r0 = 0; r0 < kRows; r0 += 32) {
        bf16x8 a[TI], x[TJ];
#pragma unroll
        for (int i = 0; i < TI; ++i) a[i] = ld8(dYT + ((i0 + i) * 16 + li) * kTr + r0 + kq);
#pragma unroll
        for (int j = 0; j < TJ; ++j) x[j] = ld8(XT + ((j0 + j) * 16 + li) * kTr + r0 + kq);
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], x[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                g[((i0 + i) * 16 + rq + e) * ldG + (j0 + j) * 16 + li] = acc[i][j][e];
}

// list[row]: lattice cell of each batch row (features); targets are read at
// clsAll/lleAll[tlist ? tlist[row] : list[row]].
__global__ __launch_bounds__(kTrainThreads) void k_train_fused(
    int rows, int nLat, const int* __restrict__ list, const int* __restrict__ tlist,
    const float* __restrict__ Bff,
    const bf16* __restrict__ wbf, const bf16* __restrict__ wbfT, const float* __restrict__ p,
    const int* __restrict__ clsAll, const float* __restrict__ lleAll, float wLle, float chaosStd,
    float* __restrict__ gpart, long long stride, float* __restrict__ loss, int* __restrict__ step)
{
    __shared__ __attribute__((aligned(16))) bf16 sX[kRows * kLdX];
    __shared__ __attribute__((aligned(16))) bf16 sXT[kFeat * kTr];
    __shared__ __attribute__((aligned(16))) bf16 sH[3][kRows * kLdH];
    __shared__ __attribute__((aligned(16))) bf16 sHT[3][kHid * kTr];
    __shared__ __attribute__((aligned(16))) bf16 sDO[kRows * kLdO];
    __shared__ __attribute__((aligned(16))) bf16 sDOT[kHead * kTr];
    __shared__ float sOut[kRows * (kHead + 1)];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, kq = (lane >> 4) * 8, rq = (lane >> 4) * 4;
    const int row0 = blockIdx.x * kRows;
    const int nrows = min(kRows, rows - row0);
    float* gp = gpart + (long long)blockIdx.x * stride;
    if (step && blockIdx.x == 0 && tid == 0) *step += 1;

    // Fourier features; rows past the batch are zero
    features_rows<kTrainThreads>(list, 0, row0, nrows, nLat, Bff, sX, sXT);
    __syncthreads();

    // forward: each of the eight waves takes one 16-unit tile of every row
    fwd_hidden<kFeat, 1>(wbf + qW1, p + pB1, sX, kLdX, sH[0], sHT[0], wave, li, kq, rq);
    __syncthreads();
    fwd_hidden<kHid, 1>(wbf + qW2, p + pB2, sH[0], kLdH, sH[1], sHT[1], wave, li, kq, rq);
    __syncthreads();
    fwd_hidden<kHid, 1>(wbf + qW3, p + pB3, sH[1], kLdH, sH[2], sHT[2], wave, li, kq, rq);
    __syncthreads();
    head_tile(wbf + qWh, p + pBh, sH[2], sOut, kHead + 1, wave, li, kq, rq);
    __syncthreads();

    // loss gradient, one row per lane of wave 0
    if (wave == 0) {
        const int r = lane;
        bf16 d[kHead];
        float ce = 0, mse = 0, bce = 0;
        if (r < nrows) {
            const int src = tlist ? tlist[row0 + r] : list[row0 + r];
            loss_row(sOut + r * (kHead + 1), clsAll[src], lleAll[src], wLle, chaosStd,
                     1.0f / (float)rows, d, ce, mse, bce);
        } else {
#pragma unroll
            for (int k = 0; k < kHead; ++k) d[k] = (bf16)0.0f;
        }
#pragma unroll
        for (int k = 0; k < kHead; ++k) {
            sDO[r * kLdO + k] = d[k];
            sDOT[k * kTr + r] = d[k];
        }
        const float w = 1.0f / (float)rows;
        ce *= w;
        mse *= w;
        bce *= w;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            ce += __shfl_xor(ce, o);
            mse += __shfl_xor(mse, o);
            bce += __shfl_xor(bce, o);
        }
        if (lane == 0) {
            atomicAdd(&loss[0], ce);
            atomicAdd(&loss[1], mse);
            atomicAdd(&loss[2], bce);
        }
    }
    __syncthreads();

    // head: dWh = dOut^T H3 (this wave: both head tiles of H3 unit tile
    // `wave`); dH3 = (dOut Wh) masked by H3
    wgrad<2, 1>(sDOT, sHT[2], 0, wave, gp + pWh, kHid, li, kq, rq);
    f32x4 acc[1][4];
    bwd_input<kHead, 1>(wbfT + tWh, sDO, kLdO, acc, wave, li, kq);
    __syncthreads();   // H3 (both layouts) is read for the last time above
    store_masked<1>(acc, sH[2], sH[2], sHT[2], wave, li, rq);
    __syncthreads();
    // layer 3: dW3 = dH3^T H2; dH2 = (dH3 W3) masked by H2
    wgrad<1, 8>(sHT[2], sHT[1], wave, 0, gp + pW3, kHid, li, kq, rq);
    bwd_input<kHid, 1>(wbfT + tW3, sH[2], kLdH, acc, wave, li, kq);
    __syncthreads();
    store_masked<1>(acc, sH[1], sH[1], sHT[1], wave, li, rq);
    __syncthreads();
    // layer 2: dW2 = dH2^T H1; dH1 = (dH2 W2) masked by H1
    wgrad<1, 8>(sHT[1], sHT[0], wave, 0, gp + pW2, kHid, li, kq, rq);
    bwd_input<kHid, 1>(wbfT + tW2, sH[1], kLdH, acc, wave, li, kq);
    __syncthreads();
    store_masked<1>(acc, sH[0], sH[0], sHT[0], wave, li, rq);
    __syncthreads();
    // layer 1: dW1 = dH1^T X
    wgrad<1, 4>(sHT[0], sXT, wave, 0, gp + pW1, kFeat, li, kq, rq);

    // bias gradients: column sums over the rows of dOut, dH3, dH2, dH1
    for (int c = tid; c < kHead + 3 * kHid; c += kTrainThreads) {
        const bf16* src;
        int off;
        if (c < kHead) { src = sDOT + c * kTr; off = pBh + c; }
        else if (c < kHead + kHid) { src = sHT[2] + (c - kHead) * kTr; off = pB3 + c - kHead; }
        else if (c < kHead + 2 * kHid) { src = sHT[1] + (c - kHead - kHid) * kTr; off = pB2 + c - kHead - kHid; }
        else { src = sHT[0] + (c - kHead - 2 * kHid) * kTr; off = pB1 + c - kHead - 2 * kHid; }
        float sum = 0;
        for (int r = 0; r < kRows; ++r) sum += (float)src[r];
        gp[off] = sum;
    }
}

// Inference post-processing of one row's head outputs: class = argmax of
// the logits, chaos regime from the chaos logit, LLE = column 0
// de-standardised and kept on the side of chaosThr the chaos logit predicts
// (the regime of a predicted cell is the classifier's; the regression
// supplies the value).  Class confidence = the class's softmax probability;
// chaos confidence = the chaos classifier's probability of its answer.
// With confChaos null, conf = min of the two; else conf = class confidence
// and confChaos = chaos confidence.
__device__ __forceinline__ void infer_post_row(const float* o, float lleMean, float lleStd,
                                               float chaosThr, long long r,
                                               unsigned char* __restrict__ cls,
                                               unsigned char* __restrict__ chaos,
                                               float* __restrict__ conf,
                                               float* __restrict__ confChaos,
                                               float* __restrict__ lle)
{
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

// Hidden layer forward with this wave's weight tiles already in registers
// (w[i][kk]: unit tile 2*wave+i, k = 32*kk + kq..+7): the MFMA sequence,
// bias, ReLU and BF16 rounding of fwd_hidden, so the activations are
// bit-identical to it.
template <int K>
__device__ __forceinline__ void fwd_hidden_reg(const bf16x8 (&w)[2][K / 32],
                                               const float* __restrict__ b, const bf16* in,
                                               int ldIn, bf16* outRm, int wave, int li, int kq,
                                               int rq)
{
    f32x4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < K / 32; ++kk) {
        bf16x8 x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = ld8(in + (j * 16 + li) * ldIn + kk * 32 + kq);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[i][kk], x[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int u0 = (2 * wave + i) * 16 + rq;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float t = acc[i][j][e] + b[u0 + e];
                v[e] = t > 0 ? t : 0.0f;
            }
            store_tile(outRm, kLdH, nullptr, u0, j * 16 + li, v);
        }
    }
}

// Fused inference, persistent: each workgroup loads its waves' weight tiles
// into registers once (W1, W2 and W3: two 16-unit tiles per wave; Wh: both
// tiles in every wave, 112 VGPRs in all) and then takes kRows-row tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... (cell = list ? list[row] : base +
// row) through the features, the three layers and the head with the
// activations in LDS.  The products, in the order of k_train_fused's forward
// pass, give the same outputs as one workgroup per tile; only the weights'
// global reads are gone.  Writes the per-row predictions (cls != null)
// and/or the raw head outputs (rawOut != null, [rows][kHead]).
__global__ __launch_bounds__(256) void k_infer_fused(
    long long rows, long long base, const int* __restrict__ list, int nLat,
    const float* __restrict__ Bff, const bf16* __restrict__ wbf, const float* __restrict__ p,
    float lleMean, float lleStd, float chaosThr, unsigned char* __restrict__ cls,
    unsigned char* __restrict__ chaos, float* __restrict__ conf, float* __restrict__ confChaos,
    float* __restrict__ lle, float* __restrict__ rawOut)
{
    __shared__ __attribute__((aligned(16))) bf16 sX[kRows * kLdX];
    __shared__ __attribute__((aligned(16))) bf16 sH[2][kRows * kLdH];
    __shared__ float sOut[kRows * (kHead + 1)];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, kq = (lane >> 4) * 8, rq = (lane >> 4) * 4;

    bf16x8 w1[2][kFeat / 32], w2[2][kHid / 32], w3[2][kHid / 32], wh[2][kHid / 32];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int u = (2 * wave + i) * 16 + li;
#pragma unroll
        for (int kk = 0; kk < kFeat / 32; ++kk) w1[i][kk] = ld8(wbf + qW1 + u * kFeat + kk * 32 + kq);
#pragma unroll
        for (int kk = 0; kk < kHid / 32; ++kk) {
            w2[i][kk] = ld8(wbf + qW2 + u * kHid + kk * 32 + kq);
            w3[i][kk] = ld8(wbf + qW3 + u * kHid + kk * 32 + kq);
            wh[i][kk] = ld8(wbf + qWh + (i * 16 + li) * kHid + kk * 32 + kq);
        }
    }
    const long long tiles = (rows + kRows - 1) / kRows;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long long row0 = tile * kRows;
        const int nrows = (int)min((long long)kRows, rows - row0);
        features_rows<256>(list, base, row0, nrows, nLat, Bff, sX, nullptr);
        __syncthreads();
        fwd_hidden_reg<kFeat>(w1, p + pB1, sX, kLdX, sH[0], wave, li, kq, rq);
        __syncthreads();
        fwd_hidden_reg<kHid>(w2, p + pB2, sH[0], kLdH, sH[1], wave, li, kq, rq);
        __syncthreads();
        fwd_hidden_reg<kHid>(w3, p + pB3, sH[1], kLdH, sH[0], wave, li, kq, rq);
        __syncthreads();
        // head: this wave's row tile, both unit tiles (head_rows' sequence)
        {
            f32x4 acc[2] = {f32x4{0, 0, 0, 0}, f32x4{0, 0, 0, 0}};
#pragma unroll
            for (int kk = 0; kk < kHid / 32; ++kk) {
                const bf16x8 x = ld8(sH[0] + (wave * 16 + li) * kLdH + kk * 32 + kq);
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[i][kk], x, acc[i], 0, 0, 0);
            }
            const int r = wave * 16 + li;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int u = i * 16 + rq + e;
                    sOut[r * (kHead + 1) + u] = acc[i][e] + p[pBh + u];
                }
        }
        __syncthreads();
        if (rawOut)
            for (int t = tid; t < nrows * kHead; t += 256)
                rawOut[(row0 + t / kHead) * kHead + t % kHead] =
                    sOut[(t / kHead) * (kHead + 1) + t % kHead];
        if (cls && tid < nrows)
            infer_post_row(sOut + tid * (kHead + 1), lleMean, lleStd, chaosThr, row0 + tid, cls,
                           chaos, conf, confChaos, lle);
        // the next tile writes sX and sH[0] only after its first barrier, by
        // which every wave has finished reading sOut and sH[0] of this one
    }
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
// Validation draw of an unflagged cell: a seeded 64-bit mix of the cell
// index (splitmix64 finaliser), kept when its top 32 bits fall below vthr
// (vthr = fraction * 2^32; 0 draws nothing).
__device__ __forceinline__ bool validation_pick(long long idx, unsigned long long seed,
                                                unsigned int vthr)
{
    unsigned long long z = (unsigned long long)idx + seed * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (unsigned int)(z >> 32) < vthr;
}

__global__ void k_flag(int n, const unsigned char* __restrict__ cls,
                       const unsigned char* __restrict__ chaos, const float* __restrict__ conf,
                       const float* __restrict__ confChaos, float confThr, float chaosConfThr,
                       bool partial, int reading, const unsigned char* __restrict__ coarseChaos,
                       int nCoarse, float quadConf, int* __restrict__ listFull,
                       int* __restrict__ countFull,
                       int* __restrict__ listPeak, int* __restrict__ countPeak,
                       int* __restrict__ listLle, int* __restrict__ countLle,
                       unsigned long long vseed, unsigned int vthr, int* __restrict__ listVal,
                       int* __restrict__ countVal, unsigned char* __restrict__ flag)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = idx < (long long)n * n;   // no early return: the appends are wave-wide
    bool fc = false, fh = false, fv = false, fl = false;
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
        // the exact coarse map: a cell inside a coarse quad whose four
        // corners disagree on the chaos bit sits where a regime boundary
        // crosses the coarse lattice, and one whose corners the network
        // itself gets wrong lies where the network is unreliable; unless the
        // chaos head is sure of the cell (confidence >= quadConf), its LLE
        // is re-solved.  Coarse corner (ci, cj) is fine cell
        // (rint(ci (n-1)/(nc-1)), rint(cj (n-1)/(nc-1))).
        if (coarseChaos && confChaos[idx] < quadConf) {
            const int qi = min((int)((long long)i * (nCoarse - 1) / (n - 1)), nCoarse - 2);
            const int qj = min((int)((long long)j * (nCoarse - 1) / (n - 1)), nCoarse - 2);
            int s4 = 0;
            bool wrong = false;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int ci = qi + (c & 1), cj = qj + (c >> 1);
                const int b = coarseChaos[(long long)cj * nCoarse + ci];
                const long long fi = (2LL * ci * (n - 1) + (nCoarse - 1)) / (2LL * (nCoarse - 1));
                const long long fj = (2LL * cj * (n - 1) + (nCoarse - 1)) / (2LL * (nCoarse - 1));
                s4 += b;
                wrong |= (chaos[fj * n + fi] != 0) != (b != 0);
            }
            fh |= (s4 > 0 && s4 < 4) || wrong;
        }
        // the regime reading decides what a cell can be unsure of
        if (reading == 1) fc = false;        // chaos bit only
        else if (reading == 2) fh = false;   // peak-count class only
        if (!confChaos || !partial) {
            fh = fh || fc;
        } else if (fh && !fc) {
            fl = true;          // regime only: the LLE alone, class kept
            fh = false;
        }
        fv = !fh && !fc && !fl && vthr != 0 && validation_pick(idx, vseed, vthr);
        flag[idx] = fh ? 1 : (fc ? 2 : (fl ? 4 : (fv ? 3 : 0)));
    }
    wave_append(live && fh, (int)idx, listFull, countFull);
    wave_append(live && !fh && fc, (int)idx, listPeak, countPeak);
    wave_append(live && fl, (int)idx, listLle, countLle);
    if (vthr != 0) wave_append(live && fv, (int)idx, listVal, countVal);
}

// Cells re-solved for the regime only whose exact regime (lle > chaosThr)
// differs from the predicted one: mostly periodic windows inside the chaotic
// region, where the predicted class is suspect too; appended to out[] and
// marked 5 in flag[].
__global__ void k_regime_changed(int n, const int* __restrict__ list,
                                 const double* __restrict__ lle,
                                 const unsigned char* __restrict__ pch, double chaosThr,
                                 int* __restrict__ out, int* __restrict__ count,
                                 unsigned char* __restrict__ flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < n;   // no early return: the append is wave-wide
    const int cell = live ? list[t] : 0;
    const bool take = live && ((lle[t] > chaosThr) != (pch[cell] != 0));
    if (take) flag[cell] = flag[cell] == 6 ? 7 : 5;
    wave_append(take, cell, out, count);
}

// FP32/FP64 split of the regime-only re-solves.  A listed cell whose FP32
// bulk LLE lies at least `margin` from the chaos threshold keeps it (flag
// 6); the others, NaN included, are appended for the FP64 re-solve: the
// cell index to cells[], its position in the list to pos[] (same slot).
__global__ void k_lle32_split(int n, const int* __restrict__ list,
                              const double* __restrict__ lle32, double chaosThr, double margin,
                              int* __restrict__ cells, int* __restrict__ pos,
                              int* __restrict__ count, unsigned char* __restrict__ flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < n;   // no early return: the append is wave-wide
    const int cell = live ? list[t] : 0;
    const bool near = live && !(fabs(lle32[t] - chaosThr) >= margin);
    if (live && !near) flag[cell] = 6;
    const uint64_t m = __ballot(near);
    if (!m) return;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __builtin_ctzll(m);
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader);
    if (near) {
        const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
        cells[slot] = cell;
        pos[slot] = t;
    }
}

// Write re-solved cells into the output maps (the peak count when peak is
// given, the LLE when lleIn is given).
__global__ void k_scatter(int n, const int* __restrict__ list, const int* __restrict__ peak,
                          const double* __restrict__ lleIn, int* __restrict__ peakMap,
                          double* __restrict__ lleMap)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = list[t];
    if (peak) peakMap[c] = peak[t];
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
// (plist), 3 = LLE only (llist).  acc[g*5 + 0..4] = cells, class right,
// regime right, finite LLE cells, sum |LLE error| (as doubles); group 2
// fills entries 0..1, group 3 scores the chaos regime alone in entry 2.
constexpr int kScoreSlots = 20;
__global__ __launch_bounds__(256) void k_score(int nFull, int nTot, const int* __restrict__ list,
                                               const int* __restrict__ rpeak,
                                               const double* __restrict__ rlle, int nPeak,
                                               const int* __restrict__ plist,
                                               const int* __restrict__ ppeak, int nLle,
                                               const int* __restrict__ llist,
                                               const double* __restrict__ llle,
                                               const unsigned char* __restrict__ pcls,
                                               const unsigned char* __restrict__ pch,
                                               const float* __restrict__ plle, double chaosThr,
                                               double* __restrict__ acc)
{
    __shared__ double s_acc[kScoreSlots];
    if (threadIdx.x < kScoreSlots) s_acc[threadIdx.x] = 0;
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
    if (t < nLle) {
        const int cell = llist[t];
        const double tl = llle[t];
        atomicAdd(&s_acc[15], 1.0);
        if ((tl > chaosThr) == (pch[cell] != 0)) atomicAdd(&s_acc[17], 1.0);
        if (isfinite(tl)) {
            atomicAdd(&s_acc[18], 1.0);
            atomicAdd(&s_acc[19], fabs(tl - (double)plle[cell]));
        }
    }
    __syncthreads();
    if (threadIdx.x < kScoreSlots && s_acc[threadIdx.x] != 0)
        atomicAdd(&acc[threadIdx.x], s_acc[threadIdx.x]);
}

// ------------------------------------------------------------ launchers

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

hipError_t launch_adam(int n, float* p, float* m, float* v, const float* gpart, int splits,
                       long long stride, const unsigned char* keep, const int* bfIndex, bf16* wbf,
                       const int* bfIndexT, bf16* wbfT, float lr, float b1, float b2, float eps,
                       const int* step, hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, n, p, m, v, gpart, splits,
                       stride, keep, bfIndex, wbf, bfIndexT, wbfT, lr, b1, b2, eps, step);
    return hipGetLastError();
}

int train_fused_blocks(int rows) { return (rows + kRows - 1) / kRows; }

hipError_t launch_infer_fused(long long rows, long long base, const int* list, int nLat,
                              const float* Bff, const bf16* wbf, const float* p, float lleMean,
                              float lleStd, float chaosThr, unsigned char* cls,
                              unsigned char* chaos, float* conf, float* confChaos, float* lle,
                              float* rawOut, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    const long long tiles = (rows + kRows - 1) / kRows;
    // persistent: as many workgroups as the device holds at once
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, perCu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_infer_fused, 256, 0);
        if (e != hipSuccess) return e;
        resident = std::max(1, cus * std::max(1, perCu));
    }
    const long long blocks = std::min<long long>(tiles, resident);
    hipLaunchKernelGGL(k_infer_fused, dim3((unsigned)blocks), dim3(256), 0, st, rows, base, list,
                       nLat, Bff, wbf, p, lleMean, lleStd, chaosThr, cls, chaos, conf, confChaos,
                       lle, rawOut);
    return hipGetLastError();
}

hipError_t launch_train_fused(int rows, int nLat, const int* list, const int* tlist,
                              const float* Bff, const bf16* wbf, const bf16* wbfT, const float* p,
                              const int* cls, const float* lle, float wLle, float chaosStd,
                              float* gpart, long long stride, float* loss, int* step,
                              hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_train_fused, dim3(train_fused_blocks(rows)), dim3(kTrainThreads), 0, st,
                       rows, nLat, list, tlist, Bff, wbf, wbfT, p, cls, lle, wLle, chaosStd, gpart,
                       stride, loss, step);
    return hipGetLastError();
}



hipError_t launch_flag(int n, const unsigned char* cls, const unsigned char* chaos,
                       const float* conf, const float* confChaos, float confThr,
                       float chaosConfThr, bool partial, int reading,
                       const unsigned char* coarseChaos, int nCoarse, float quadConf,
                       int* listFull, int* countFull,
                       int* listPeak, int* countPeak, int* listLle, int* countLle,
                       unsigned long long vseed, unsigned int vthr, int* listVal, int* countVal,
                       unsigned char* flag, hipStream_t st)
{
    const long long cells = (long long)n * n;
    hipLaunchKernelGGL(k_flag, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, n, cls,
                       chaos, conf, confChaos, confThr, chaosConfThr, partial, reading, coarseChaos,
                       nCoarse, quadConf, listFull,
                       countFull, listPeak, countPeak, listLle, countLle, vseed, vthr, listVal,
                       countVal, flag);
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
hipError_t launch_regime_changed(int n, const int* list, const double* lle,
                                 const unsigned char* pch, double chaosThr, int* out, int* count,
                                 unsigned char* flag, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_regime_changed, dim3((n + 255) / 256), dim3(256), 0, st, n, list, lle,
                       pch, chaosThr, out, count, flag);
    return hipGetLastError();
}

hipError_t launch_lle32_split(int n, const int* list, const double* lle32, double chaosThr,
                              double margin, int* cells, int* pos, int* count,
                              unsigned char* flag, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lle32_split, dim3((n + 255) / 256), dim3(256), 0, st, n, list, lle32,
                       chaosThr, margin, cells, pos, count, flag);
    return hipGetLastError();
}

hipError_t launch_score(int nFull, int nTot, const int* list, const int* rpeak, const double* rlle,
                        int nPeak, const int* plist, const int* ppeak, int nLle,
                        const int* llist, const double* llle, const unsigned char* pcls,
                        const unsigned char* pch, const float* plle, double chaosThr,
                        double* acc, hipStream_t st)
{
    const int n = max(max(nTot, nPeak), nLle);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_score, dim3((n + 255) / 256), dim3(256), 0, st, nFull, nTot, list, rpeak,
                       rlle, nPeak, plist, ppeak, nLle, llist, llle, pcls, pch, plle, chaosThr,
                       acc);
    return hipGetLastError();
}
}  // namespace bifhip
