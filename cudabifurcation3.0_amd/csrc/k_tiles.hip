// k_tiles.hip -- Lyapunov spectrum over uneven refinement tiles, gfx950.
//
// One persistent launch covers every tile.  Work items are slices of a tile
// (blockDim * cellsPerLane cells); blocks pull items from an atomic counter
// in longest-tile-first order, so tiles of 200..2000 time units balance
// across the 256 CUs without a host round trip.  Per lane: the state and
// three tangent vectors in FP32 registers, advanced with the variational
// equation of the discrete step; every nNT steps the vectors are promoted to
// FP64 and orthonormalised with classical Gram-Schmidt (the reference's
// scheme, Library/cudaLibrary.cu:913-957, with normalised projections), and
// the norms accumulate in FP64 as a product, logged once at the end.
#include "bifhip_dev.h"

namespace bifhip {

// Gram-Schmidt of the three FP32 tangent vectors in FP64: the classical
// scheme (each vector minus its projections on the previous ones, as
// Library/cudaLibrary.cu:913-957), written with the previous vectors already
// normalised, so a projection is a dot product and no division; each norm
// from a reciprocal square root.  Zero vectors stay zero (the reference's
// guards), their log norm is -inf.
__device__ __forceinline__ void gs3_f(float (&v)[9], LogNormSum (&acc)[3])
{
    double u[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) u[i] = v[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double a[3] = {u[i * 3], u[i * 3 + 1], u[i * 3 + 2]};
        double b[3] = {a[0], a[1], a[2]};
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double p = fma(a[2], u[j * 3 + 2], fma(a[1], u[j * 3 + 1], a[0] * u[j * 3]));
#pragma unroll
            for (int m = 0; m < 3; ++m) b[m] = fma(-p, u[j * 3 + m], b[m]);
        }
        const double d2 = fma(b[2], b[2], fma(b[1], b[1], b[0] * b[0]));
        const double r = d2 > 0.0 ? rsqrt(d2) : 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            u[i * 3 + m] = b[m] * r;
            v[i * 3 + m] = (float)u[i * 3 + m];
        }
        acc[i].add(d2);
    }
}

template <class M, int CPL>
__global__ void k_ls_tiles(const TileArgs A)
{
    __shared__ int s_item;
    const int totalItems = A.nTiles * A.itemsPerTile;
    const int tileCells = A.tileW * A.tileH;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(A.counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();   // every lane has read s_item before lane 0 rewrites it
        if (item >= totalItems) return;
        const int tile = A.order[item / A.itemsPerTile];
        const int part = item % A.itemsPerTile;
        const int nb = A.nBlocks[tile];

        typename M::P32 q[CPL];
        float x0[CPL], x1[CPL], x2[CPL], v[CPL][9];
        LogNormSum acc[CPL][3];
        bool live[CPL];
        int cellOf[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int cell = part * A.itemCells + c * (int)blockDim.x + (int)threadIdx.x;
            cellOf[c] = cell;
            live[c] = cell < tileCells;
            const int ci = live[c] ? cell % A.tileW : 0;
            const int cj = live[c] ? cell / A.tileW : 0;
            const int idx = (A.originJ[tile] + cj) * A.g.nPts + (A.originI[tile] + ci);
            double vals[kMaxVal], s0, s1, s2;
            cell_setup(A.g, idx, vals, s0, s1, s2);
            q[c] = M::params32(vals, A.h);
            x0[c] = (float)s0;
            x1[c] = (float)s1;
            x2[c] = (float)s2;
#pragma unroll
            for (int k = 0; k < 9; ++k) v[c][k] = (k % 4 == 0) ? 1.0f : 0.0f;
            acc[c][0].init();
            acc[c][1].init();
            acc[c][2].init();
        }
        for (int s = 0; s < A.skip; ++s)
#pragma unroll
            for (int c = 0; c < CPL; ++c) M::step32(x0[c], x1[c], x2[c], q[c]);

        for (int b = 0; b < nb; ++b) {
            // unrolled so the loop-carried tangent components keep their
            // registers (one step per iteration rotates them with moves)
#pragma unroll 2
            for (int s = 0; s < A.nNT; ++s) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const float x2o = x2[c];
                    M::step32(x0[c], x1[c], x2[c], q[c]);
                    M::tan32(v[c][0], v[c][1], v[c][2], x0[c], x2o, q[c]);
                    M::tan32(v[c][3], v[c][4], v[c][5], x0[c], x2o, q[c]);
                    M::tan32(v[c][6], v[c][7], v[c][8], x0[c], x2o, q[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < CPL; ++c) gs3_f(v[c], acc[c]);
        }
        const double T = (double)nb * A.nNT * A.h;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            if (!live[c]) continue;
            double* o = A.out + ((long long)tile * tileCells + cellOf[c]) * 3;
            o[0] = acc[c][0].value() / T;
            o[1] = acc[c][1].value() / T;
            o[2] = acc[c][2].value() / T;
        }
    }
}

// Three lanes per cell: lane 3k+j of a wavefront carries tangent vector j
// of the wave's cell k (21 cells per wavefront, lane 63 idle), and each of
// the three lanes steps the cell's state itself (the same operations, so
// the same values).  At each renormalisation the three lanes exchange their
// vectors by cross-lane reads and all three run the one-lane kernel's
// Gram-Schmidt on the same nine values, so every result is bit-identical to
// k_ls_tiles; the lane holding vector j writes exponent j.  A work item is
// 21 cells per wavefront of the block.
template <class M>
__global__ void k_ls_tiles3(const TileArgs A)
{
    __shared__ int s_item;
    const int totalItems = A.nTiles * A.itemsPerTile;
    const int tileCells = A.tileW * A.tileH;
    const int lane = (int)threadIdx.x & 63;
    const int j = lane % 3;
    const int base = lane - j;
    const int slot = ((int)threadIdx.x >> 6) * 21 + lane / 3;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(A.counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();   // every lane has read s_item before lane 0 rewrites it
        if (item >= totalItems) return;
        const int tile = A.order[item / A.itemsPerTile];
        const int part = item % A.itemsPerTile;
        const int nb = A.nBlocks[tile];
        const int cell = part * A.itemCells + slot;
        const bool live = lane < 63 && cell < tileCells;
        const int ci = live ? cell % A.tileW : 0;
        const int cj = live ? cell / A.tileW : 0;
        const int idx = (A.originJ[tile] + cj) * A.g.nPts + (A.originI[tile] + ci);
        double vals[kMaxVal], s0, s1, s2;
        cell_setup(A.g, idx, vals, s0, s1, s2);
        const typename M::P32 q = M::params32(vals, A.h);
        float x0 = (float)s0, x1 = (float)s1, x2 = (float)s2;
        float w0 = j == 0 ? 1.0f : 0.0f, w1 = j == 1 ? 1.0f : 0.0f, w2 = j == 2 ? 1.0f : 0.0f;
        LogNormSum acc[3];
        acc[0].init();
        acc[1].init();
        acc[2].init();
        for (int s = 0; s < A.skip; ++s) M::step32(x0, x1, x2, q);
        for (int b = 0; b < nb; ++b) {
#pragma unroll 2
            for (int s = 0; s < A.nNT; ++s) {
                const float x2o = x2;
                M::step32(x0, x1, x2, q);
                M::tan32(w0, w1, w2, x0, x2o, q);
            }
            float v[9];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[3 * k] = __shfl(w0, base + k, 64);
                v[3 * k + 1] = __shfl(w1, base + k, 64);
                v[3 * k + 2] = __shfl(w2, base + k, 64);
            }
            gs3_f(v, acc);
            w0 = j == 0 ? v[0] : (j == 1 ? v[3] : v[6]);
            w1 = j == 0 ? v[1] : (j == 1 ? v[4] : v[7]);
            w2 = j == 0 ? v[2] : (j == 1 ? v[5] : v[8]);
        }
        if (live) {
            const double T = (double)nb * A.nNT * A.h;
            const double r = j == 0 ? acc[0].value() : (j == 1 ? acc[1].value() : acc[2].value());
            A.out[((long long)tile * tileCells + cell) * 3 + j] = r / T;
        }
    }
}

// FP64 state and tangent vectors, one lane per cell: the tangent map of the
// FP64 step (M::tan64), the reference's Gram-Schmidt (gs3, with divisions
// and square roots, Library/cudaLibrary.cu:913-957) every nNT steps and a
// sum of log norms -- the operations of the FP64 variational oracle in its
// order.  The reference against which the FP32 tiles' error is measured.
template <class M>
__global__ void k_ls_tiles64(const TileArgs A)
{
    __shared__ int s_item;
    const int totalItems = A.nTiles * A.itemsPerTile;
    const int tileCells = A.tileW * A.tileH;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(A.counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();   // every lane has read s_item before lane 0 rewrites it
        if (item >= totalItems) return;
        const int tile = A.order[item / A.itemsPerTile];
        const int part = item % A.itemsPerTile;
        const int nb = A.nBlocks[tile];
        const int cell = part * A.itemCells + (int)threadIdx.x;
        const bool live = cell < tileCells;
        const int ci = live ? cell % A.tileW : 0;
        const int cj = live ? cell / A.tileW : 0;
        const int idx = (A.originJ[tile] + cj) * A.g.nPts + (A.originI[tile] + ci);
        double vals[kMaxVal], x0, x1, x2;
        cell_setup(A.g, idx, vals, x0, x1, x2);
        const typename M::P64 p = M::params(vals);
        const double h = A.h;
        double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double res[3] = {0, 0, 0};
        for (int s = 0; s < A.skip; ++s) M::step(x0, x1, x2, p, h);
        for (int b = 0; b < nb; ++b) {
            for (int s = 0; s < A.nNT; ++s) {
                const double x2o = x2;
                M::step(x0, x1, x2, p, h);
                M::tan64(v[0], v[1], v[2], x0, x2o, p, h);
                M::tan64(v[3], v[4], v[5], x0, x2o, p, h);
                M::tan64(v[6], v[7], v[8], x0, x2o, p, h);
            }
            double w[9], den[3];
            gs3(v, w, den);
#pragma unroll
            for (int k = 0; k < 3; ++k) res[k] += log(den[k]);
#pragma unroll
            for (int k = 0; k < 9; ++k) v[k] = w[k];
        }
        if (live) {
            const double T = (double)nb * A.nNT * h;
            double* o = A.out + ((long long)tile * tileCells + cell) * 3;
            o[0] = res[0] / T;
            o[1] = res[1] / T;
            o[2] = res[2] / T;
        }
    }
}

__global__ void k_zero_counter(int* c) { *c = 0; }

// variant: kTileFp32 (one lane per cell, cpl cells per lane), kTileFp32x3
// (three lanes per cell), kTileFp64 (FP64 state, one lane per cell)
hipError_t launch_ls_tiles(int model, int variant, int cpl, int block, int grid,
                           const TileArgs& A, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, A.counter);
#define TILE_LAUNCH(K) hipLaunchKernelGGL(K, dim3(grid), dim3(block), 0, st, A)
    if (model == kLinear) {
        if (variant == kTileFp64) TILE_LAUNCH((k_ls_tiles64<Linear>));
        else if (variant == kTileFp32x3) TILE_LAUNCH((k_ls_tiles3<Linear>));
        else if (cpl == 2) TILE_LAUNCH((k_ls_tiles<Linear, 2>));
        else TILE_LAUNCH((k_ls_tiles<Linear, 1>));
    } else {
        if (variant == kTileFp64) TILE_LAUNCH((k_ls_tiles64<Rossler>));
        else if (variant == kTileFp32x3) TILE_LAUNCH((k_ls_tiles3<Rossler>));
        else if (cpl == 2) TILE_LAUNCH((k_ls_tiles<Rossler, 2>));
        else TILE_LAUNCH((k_ls_tiles<Rossler, 1>));
    }
#undef TILE_LAUNCH
    return hipGetLastError();
}

// kernel handle for the occupancy query of the persistent grid
const void* tile_kernel_ptr(int model, int variant, int cpl)
{
    if (model == kLinear) {
        if (variant == kTileFp64) return (const void*)&k_ls_tiles64<Linear>;
        if (variant == kTileFp32x3) return (const void*)&k_ls_tiles3<Linear>;
        return cpl == 2 ? (const void*)&k_ls_tiles<Linear, 2> : (const void*)&k_ls_tiles<Linear, 1>;
    }
    if (variant == kTileFp64) return (const void*)&k_ls_tiles64<Rossler>;
    if (variant == kTileFp32x3) return (const void*)&k_ls_tiles3<Rossler>;
    return cpl == 2 ? (const void*)&k_ls_tiles<Rossler, 2> : (const void*)&k_ls_tiles<Rossler, 1>;
}

}  // namespace bifhip
