// host_tiles.cpp -- uneven refinement tiles: generator, work queue order,
// autotuner and the graph-captured persistent launch (k_tiles.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "bifhip.h"
#include "host_internal.h"

namespace bifhip {
hipError_t launch_ls_tiles(int model, int variant, int cpl, int block, int grid,
                           const TileArgs& A, hipStream_t st);
const void* tile_kernel_ptr(int model, int variant, int cpl);
}  // namespace bifhip

using namespace bifhip;
using namespace bifhip::host;

namespace {

struct TileConfig {
    int block = 0;        // 0: not chosen yet (autotune on first use)
    int cpl = 1;          // cells per lane (one lane per cell)
    int blocksPerCU = 0;  // 0: occupancy maximum
    int lanes = 1;        // lanes per cell: 1, or 3 (one tangent vector per lane)
};
TileConfig g_cfg;
hipStream_t g_tstream = nullptr;

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<unsigned char> key;
    void reset()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
};
GraphCache g_graph;
// the last autotuner run: (block, lanes per cell, cells per lane, ms) per candidate
std::vector<std::array<double, 4>> tuneLog;

double u01(std::mt19937_64& r) { return (double)(r() >> 11) * (1.0 / 9007199254740992.0); }

int tile_stream(hipStream_t* s)
{
    if (!g_tstream) HIP_TRY(hipStreamCreateWithFlags(&g_tstream, hipStreamNonBlocking));
    *s = g_tstream;
    return BIFHIP_OK;
}

int num_cus(int* n)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, dev));
    return BIFHIP_OK;
}

template <class K>
int occupancy(K kernel, int block, int* blocks)
{
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, kernel, block, 0));
    if (*blocks < 1) *blocks = 1;
    return BIFHIP_OK;
}

// the kernel variant a configuration runs at the context's tile precision
int tile_variant(int lanes)
{
    if (g.tilePrecision == BIFHIP_PREC_FP64) return kTileFp64;
    return lanes == 3 ? kTileFp32x3 : kTileFp32;
}

// persistent grid = resident capacity of the device for this configuration
int grid_for(int variant, int block, int cpl, int bpcu, int* grid)
{
    int cus = 0;
    RC_TRY(num_cus(&cus));
    int per = bpcu;
    if (per <= 0) RC_TRY(occupancy(tile_kernel_ptr(g.model, variant, cpl), block, &per));
    *grid = std::max(1, cus * per);
    return BIFHIP_OK;
}

struct TileJob {
    TileArgs A;
    int variant, cpl, block, grid;
};

int run_job(const TileJob& J, hipStream_t s, bool graph)
{
    if (!graph) {
        HIP_TRY(launch_ls_tiles(g.model, J.variant, J.cpl, J.block, J.grid, J.A, s));
        return BIFHIP_OK;
    }
    std::vector<unsigned char> key(sizeof(TileJob) + sizeof(int));
    std::memcpy(key.data(), &J, sizeof(TileJob));
    std::memcpy(key.data() + sizeof(TileJob), &g.model, sizeof(int));
    if (!g_graph.exec || g_graph.key != key) {
        g_graph.reset();
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_ls_tiles(g.model, J.variant, J.cpl, J.block, J.grid, J.A, s);
        hipGraph_t gr = nullptr;
        hipError_t ce = hipStreamEndCapture(s, &gr);
        HIP_TRY(le);
        HIP_TRY(ce);
        g_graph.graph = gr;
        HIP_TRY(hipGraphInstantiate(&g_graph.exec, gr, nullptr, nullptr, 0));
        g_graph.key = key;
    }
    HIP_TRY(hipGraphLaunch(g_graph.exec, s));
    return BIFHIP_OK;
}

}  // namespace

extern "C" {

int bifhip_tiles_generate(int nTiles, int tileW, int tileH, int nPts,
                          unsigned long long seedOrigins, unsigned long long seedLengths,
                          double tMinTime, double tMaxTime, double NT, int* originI,
                          int* originJ, int* nBlocks)
{
    if (nTiles < 0 || tileW < 1 || tileH < 1 || nPts < tileW || nPts < tileH || !originI ||
        !originJ || !nBlocks || !(NT > 0) || !(tMaxTime >= tMinTime) || !(tMinTime >= 0))
        return fail(BIFHIP_E_ARG, "bad tile generator arguments");
    // origins: half uniform over the lattice, half Gaussian clusters around
    // the period-doubling cascade and the onset of chaos of the Rossler
    // (a, c) box (centres as fractions of the a and c ranges)
    static const double centres[3][2] = {{0.375, 0.20}, {0.375, 0.27}, {0.60, 0.45}};
    const double sigma = 0.05;
    std::mt19937_64 ro(seedOrigins);
    const int spanI = nPts - tileW, spanJ = nPts - tileH;
    for (int t = 0; t < nTiles; ++t) {
        double fi, fj;
        if (u01(ro) < 0.5) {
            fi = u01(ro);
            fj = u01(ro);
        } else {
            const int k = std::min(2, (int)(u01(ro) * 3));
            const double r = std::sqrt(-2.0 * std::log(1.0 - u01(ro)));
            const double th = 2.0 * M_PI * u01(ro);
            fi = centres[k][0] + sigma * r * std::cos(th);
            fj = centres[k][1] + sigma * r * std::sin(th);
        }
        fi = std::min(1.0, std::max(0.0, fi));
        fj = std::min(1.0, std::max(0.0, fj));
        originI[t] = (int)std::lround(fi * spanI);
        originJ[t] = (int)std::lround(fj * spanJ);
    }
    std::mt19937_64 rl(seedLengths);
    for (int t = 0; t < nTiles; ++t) {
        const double T = tMinTime + u01(rl) * (tMaxTime - tMinTime);
        nBlocks[t] = std::max(1, (int)(T / NT));
    }
    return BIFHIP_OK;
}

int bifhip_set_tile_config(int blockSize, int cellsPerLane, int blocksPerCU)
{
    if (blockSize != 0 && blockSize != 64 && blockSize != 128 && blockSize != 256)
        return fail(BIFHIP_E_ARG, "block size must be 0 (autotune), 64, 128 or 256");
    if (cellsPerLane != 1 && cellsPerLane != 2)
        return fail(BIFHIP_E_ARG, "cells per lane must be 1 or 2");
    if (blocksPerCU < 0) return fail(BIFHIP_E_ARG, "blocks per CU must be >= 0");
    g_cfg.block = blockSize;
    g_cfg.cpl = cellsPerLane;
    g_cfg.blocksPerCU = blocksPerCU;
    if (cellsPerLane != 1) g_cfg.lanes = 1;
    g_graph.reset();
    return BIFHIP_OK;
}

int bifhip_set_tile_lanes(int lanesPerCell)
{
    if (lanesPerCell != 1 && lanesPerCell != 3)
        return fail(BIFHIP_E_ARG, "lanes per cell must be 1 or 3");
    g_cfg.lanes = lanesPerCell;
    if (lanesPerCell == 3) g_cfg.cpl = 1;
    g_graph.reset();
    return BIFHIP_OK;
}

int bifhip_get_tile_lanes(int* lanesPerCell)
{
    if (lanesPerCell) *lanesPerCell = g_cfg.lanes;
    return BIFHIP_OK;
}

int bifhip_tile_autotune_log(double* out, int maxEntries)
{
    const int n = (int)tuneLog.size();
    if (out)
        for (int i = 0; i < std::min(n, maxEntries); ++i)
            for (int k = 0; k < 4; ++k) out[4 * i + k] = tuneLog[i][k];
    return n;
}

int bifhip_set_tile_precision(int precision)
{
    if (precision != BIFHIP_PREC_FP32 && precision != BIFHIP_PREC_FP64)
        return fail(BIFHIP_E_ARG, "unknown tile precision %d", precision);
    g.tilePrecision = precision;
    g_graph.reset();
    return BIFHIP_OK;
}

int bifhip_get_tile_config(int* blockSize, int* cellsPerLane, int* blocksPerCU)
{
    if (blockSize) *blockSize = g_cfg.block;
    if (cellsPerLane) *cellsPerLane = g_cfg.cpl;
    if (blocksPerCU) *blocksPerCU = g_cfg.blocksPerCU;
    return BIFHIP_OK;
}

int bifhip_set_graph(int on)
{
    g.useGraph = on != 0;
    g_graph.reset();
    return BIFHIP_OK;
}

int bifhip_set_partial_resolve(int on)
{
    g.partialResolve = on != 0;
    return BIFHIP_OK;
}

int bifhip_set_chaos_confidence(double threshold)
{
    if (!(threshold >= 0.5 && threshold <= 1.0))
        return fail(BIFHIP_E_ARG, "chaos confidence threshold %g outside [0.5, 1]", threshold);
    g.chaosConfidence = threshold;
    return BIFHIP_OK;
}

int bifhip_set_lle32_margin(double margin)
{
    if (std::isnan(margin)) return fail(BIFHIP_E_ARG, "FP32 LLE margin is NaN");
    g.lle32Margin = margin;
    return BIFHIP_OK;
}

int bifhip_set_regime_reading(int reading)
{
    if (reading != BIFHIP_READING_CLASS_AND_CHAOS && reading != BIFHIP_READING_CHAOS &&
        reading != BIFHIP_READING_CLASS)
        return fail(BIFHIP_E_ARG, "unknown regime reading %d", reading);
    g.reading = reading;
    return BIFHIP_OK;
}

int bifhip_set_quad_confidence(double threshold)
{
    if (!(threshold >= 0.0 && threshold <= 2.0))
        return fail(BIFHIP_E_ARG, "quad confidence threshold %g outside [0, 2]", threshold);
    g.quadConfidence = threshold;
    return BIFHIP_OK;
}

static int tiles_impl(int nTiles, int tileW, int tileH, const int* originI, const int* originJ,
                      const int* nBlocks, int nPts, const double* ranges, const int* mut,
                      const double* ic, int nIC, const double* values, int nVal, double h,
                      double NT, double transientTime, double* d_out, bool tune)
{
    RC_TRY(ensure_device());
    if (nTiles < 1 || tileW < 1 || tileH < 1 || !originI || !originJ || !nBlocks || !d_out)
        return fail(BIFHIP_E_ARG, "bad tile arguments");
    if (g.model == BIFHIP_MODEL_MEMRISTIVE)
        return fail(BIFHIP_E_ARG, "the tile path supports the Rossler and linear models");
    Grid G;
    RC_TRY(make_grid(G, nPts, 2, ranges, mut, ic, nIC, values, nVal));
    for (int t = 0; t < nTiles; ++t) {
        if (originI[t] < 0 || originJ[t] < 0 || originI[t] + tileW > nPts ||
            originJ[t] + tileH > nPts || nBlocks[t] < 0)
            return fail(BIFHIP_E_ARG, "tile %d lies outside the %d x %d lattice", t, nPts, nPts);
    }
    const size_t nNT = (size_t)(NT / h);
    if (nNT < 1 || nNT > (size_t)INT32_MAX) return fail(BIFHIP_E_ARG, "bad NT/h");
    // longest tile first: the queue hands the big items out early
    std::vector<int> order(nTiles);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return nBlocks[x] > nBlocks[y]; });
    int *d_tab, *d_counter;
    RC_TRY(pool_get("tiles_tab", (size_t)nTiles * 4, &d_tab));
    RC_TRY(pool_get("tiles_counter", 1, &d_counter));
    std::vector<int> tab((size_t)nTiles * 4);
    std::memcpy(tab.data(), order.data(), sizeof(int) * nTiles);
    std::memcpy(tab.data() + nTiles, originI, sizeof(int) * nTiles);
    std::memcpy(tab.data() + 2 * nTiles, originJ, sizeof(int) * nTiles);
    std::memcpy(tab.data() + 3 * nTiles, nBlocks, sizeof(int) * nTiles);
    hipStream_t s;
    RC_TRY(tile_stream(&s));
    HIP_TRY(hipMemcpyAsync(d_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));

    TileArgs A;
    std::memset(&A, 0, sizeof(A));
    A.g = G;
    A.nTiles = nTiles;
    A.tileW = tileW;
    A.tileH = tileH;
    A.order = d_tab;
    A.originI = d_tab + nTiles;
    A.originJ = d_tab + 2 * nTiles;
    A.nBlocks = d_tab + 3 * nTiles;
    A.h = h;
    A.nNT = (int)nNT;
    A.skip = (int)(transientTime / h);
    A.out = d_out;
    A.counter = d_counter;

    auto job_for = [&](int block, int cpl, int lanes, int bpcu, TileJob& J) -> int {
        J.A = A;
        J.variant = tile_variant(lanes);
        J.cpl = J.variant == kTileFp32 ? cpl : 1;
        J.block = block;
        // three lanes per cell: 21 cells per wavefront
        J.A.itemCells = J.variant == kTileFp32x3 ? (block / 64) * 21 : block * J.cpl;
        J.A.itemsPerTile = (tileW * tileH + J.A.itemCells - 1) / J.A.itemCells;
        return grid_for(J.variant, block, J.cpl, bpcu, &J.grid);
    };

    if (tune || g_cfg.block == 0) {
        // Autotuner: time each (block, lanes per cell, cells per lane) on a
        // strided sample of the queue (every 8th tile, lengths cut to a
        // tenth) and keep the fastest.  The sample keeps the length mix of
        // the real job.  The FP64 variant has one lane and one cell per lane.
        const int ns = std::max(1, nTiles / 8);
        std::vector<int> so(ns), sb(ns);
        for (int k = 0; k < ns; ++k) {
            so[k] = order[(size_t)k * nTiles / ns];
            sb[k] = std::max(1, nBlocks[so[k]] / 10);
        }
        int* d_s;
        RC_TRY(pool_get("tiles_tune", (size_t)nTiles * 4, &d_s));
        std::vector<int> st((size_t)nTiles * 4, 0);
        for (int k = 0; k < ns; ++k) {
            st[k] = so[k];
            st[3 * (size_t)nTiles + so[k]] = sb[k];
        }
        std::memcpy(st.data() + nTiles, originI, sizeof(int) * nTiles);
        std::memcpy(st.data() + 2 * nTiles, originJ, sizeof(int) * nTiles);
        HIP_TRY(hipMemcpy(d_s, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
        const int blocks[3] = {64, 128, 256};
        // (lanes per cell, cells per lane)
        const int shapes[3][2] = {{1, 1}, {1, 2}, {3, 1}};
        const int nShapes = g.tilePrecision == BIFHIP_PREC_FP64 ? 1 : 3;
        float best = 1e30f;
        TileConfig pick{256, 1, 0, 1};
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        tuneLog.clear();
        for (int bi = 0; bi < 3; ++bi) {
            for (int si = 0; si < nShapes; ++si) {
                const int lanes = shapes[si][0], cpl = shapes[si][1];
                TileJob J;
                RC_TRY(job_for(blocks[bi], cpl, lanes, 0, J));
                J.A.nTiles = ns;
                J.A.order = d_s;
                J.A.originI = d_s + nTiles;
                J.A.originJ = d_s + 2 * nTiles;
                J.A.nBlocks = d_s + 3 * nTiles;
                // warm, then time
                RC_TRY(run_job(J, s, false));
                HIP_TRY(hipEventRecord(e0, s));
                RC_TRY(run_job(J, s, false));
                HIP_TRY(hipEventRecord(e1, s));
                HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (g.verbose)
                    printf("tile autotune: block %d lanes %d cpl %d grid %d: %.3f ms\n",
                           blocks[bi], lanes, cpl, J.grid, ms);
                tuneLog.push_back({(double)blocks[bi], (double)lanes, (double)cpl, ms});
                if (ms < best) {
                    best = ms;
                    pick = {blocks[bi], cpl, 0, lanes};
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        g_cfg = pick;
        g_graph.reset();
    }

    TileJob J;
    std::memset(&J, 0, sizeof(J));
    RC_TRY(job_for(g_cfg.block, g_cfg.cpl, g_cfg.lanes, g_cfg.blocksPerCU, J));
    {
        Timed t("ls_tiles", s);
        RC_TRY(run_job(J, s, g.useGraph != 0));
    }
    HIP_TRY(hipStreamSynchronize(s));
    resolve_timers();
    return BIFHIP_OK;
}

int bifhip_ls_tiles_device(int nTiles, int tileW, int tileH, const int* originI,
                           const int* originJ, const int* nBlocks, int nPts, const double* ranges,
                           const int* indicesOfMutVars, const double* initialConditions,
                           int amountOfInitialConditions, const double* values,
                           int amountOfValues, double h, double NT, double transientTime,
                           double* d_out)
{
    return tiles_impl(nTiles, tileW, tileH, originI, originJ, nBlocks, nPts, ranges,
                      indicesOfMutVars, initialConditions, amountOfInitialConditions, values,
                      amountOfValues, h, NT, transientTime, d_out, false);
}

int bifhip_ls_tiles_autotune(int nTiles, int tileW, int tileH, const int* originI,
                             const int* originJ, const int* nBlocks, int nPts,
                             const double* ranges, const int* indicesOfMutVars,
                             const double* initialConditions, int amountOfInitialConditions,
                             const double* values, int amountOfValues, double h, double NT,
                             double transientTime, double* d_out)
{
    return tiles_impl(nTiles, tileW, tileH, originI, originJ, nBlocks, nPts, ranges,
                      indicesOfMutVars, initialConditions, amountOfInitialConditions, values,
                      amountOfValues, h, NT, transientTime, d_out, true);
}

}  // extern "C"
