// k_tiles.hip -- Lyapunov spectrum over uneven refinement tiles, gfx950.
//
// One persistent launch covers every tile.  Work items are slices of a tile
// (blockDim * cellsPerLane cells); blocks pull items from an atomic counter
// in longest-tile-first order, so tiles of 200..2000 time units balance
// across the 256 CUs without a host round trip.  Per lane: the state and
// three tangent vectors in FP32 registers, advanced with the variational
// equation of the discrete step; every nNT steps the vectors are promoted to
// FP64, orthonormalised with the reference's classical Gram-Schmidt
// (Library/cudaLibrary.cu:913-957) and log(norm) accumulates in FP64.
#include "bifhip_dev.h"

namespace bifhip {

__device__ __forceinline__ void gs3_f(float (&v)[9], double (&acc)[3])
{
    double a[9], b[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = v[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) b[i * 3 + j] = a[i * 3 + j];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double num = 0, dd = 0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                num += a[i * 3 + m] * b[j * 3 + m];
                dd += b[j * 3 + m] * b[j * 3 + m];
            }
            const double f = dd == 0 ? 0 : num / dd;
#pragma unroll
            for (int m = 0; m < 3; ++m) b[i * 3 + m] -= f * b[j * 3 + m];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double d = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) d += b[i * 3 + j] * b[i * 3 + j];
        d = sqrt(d);
        const double inv = d == 0 ? 0 : 1 / d;
#pragma unroll
        for (int j = 0; j < 3; ++j) v[i * 3 + j] = (float)(b[i * 3 + j] * inv);
        acc[i] += log(d);
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
        double acc[CPL][3];
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
            acc[c][0] = acc[c][1] = acc[c][2] = 0;
        }
        for (int s = 0; s < A.skip; ++s)
#pragma unroll
            for (int c = 0; c < CPL; ++c) M::step32(x0[c], x1[c], x2[c], q[c]);

        for (int b = 0; b < nb; ++b) {
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
            o[0] = acc[c][0] / T;
            o[1] = acc[c][1] / T;
            o[2] = acc[c][2] / T;
        }
    }
}

__global__ void k_zero_counter(int* c) { *c = 0; }

hipError_t launch_ls_tiles(int model, int cpl, int block, int grid, const TileArgs& A,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, A.counter);
#define TILE_CASE(MOD, C) \
    hipLaunchKernelGGL((k_ls_tiles<MOD, C>), dim3(grid), dim3(block), 0, st, A)
    if (model == kLinear) {
        if (cpl == 2) TILE_CASE(Linear, 2); else TILE_CASE(Linear, 1);
    } else {
        if (cpl == 2) TILE_CASE(Rossler, 2); else TILE_CASE(Rossler, 1);
    }
#undef TILE_CASE
    return hipGetLastError();
}

// kernel handle for the occupancy query of the persistent grid
const void* tile_kernel_ptr(int model, int cpl)
{
    if (model == kLinear)
        return cpl == 2 ? (const void*)&k_ls_tiles<Linear, 2> : (const void*)&k_ls_tiles<Linear, 1>;
    return cpl == 2 ? (const void*)&k_ls_tiles<Rossler, 2> : (const void*)&k_ls_tiles<Rossler, 1>;
}

}  // namespace bifhip
