// k_bif.hip -- bifurcation kernels for gfx950.
//
//  k_bif_peaks      one trajectory per lane, state in VGPRs: transient, then
//                   the recorded window fed sample by sample into a streaming
//                   restatement of peakFinder (Library/cudaLibrary.cu:220-266)
//                   so the 50 000-sample series never touches HBM.  Peaks
//                   (amplitude, interval) land in a per-cell row.
//  k_series         the reference's record pass (calculateDiscreteModelCUDA,
//                   Library/cudaLibrary.cu:78-117) writing the raw series;
//                   used by the peak-finder parity tests.
//  k_peak_finder    peakFinderCUDA (Library/cudaLibrary.cu:270-284) over
//                   series already in HBM, same streaming state machine.
//  k_dbscan_wave    dbscan (Library/cudaLibrary.cu:457-529), one wavefront
//                   per cell: the reference's sequential label/stack
//                   machine replayed on 64-bit masks, each neighbour row
//                   built with __ballot when the machine scans it.  Exact
//                   for P <= 256 peaks.
//  k_dbscan_block   the same machine, one workgroup per cell, label masks
//                   in LDS, for the rare cells with 257..16384 peaks.
//  k_dbscan_lane    the same algorithm one lane per cell on HBM scratch,
//                   for cells with still more peaks.
//  k_kde_wave       kde (Library/cudaLibrary.cu:356-422), one wavefront per
//                   cell: lanes evaluate 64 sample points at a time (each sum
//                   in the reference's peak order), the local-maximum state
//                   machine runs wave-uniformly over them.
#include "bifhip_dev.h"
#include "peak_stream.h"

namespace bifhip {

// samples tested together by the trajectory peak stream's batched path
#ifndef BIFHIP_PEAK_BATCH
#define BIFHIP_PEAK_BATCH 4
#endif
constexpr int kPeakBatch = BIFHIP_PEAK_BATCH;
// peaks staged in LDS per lane before they go to the row, 12 bytes each
// (amplitude and sample gap): 4 in the fused kernel (one cell per lane, 12
// KiB per workgroup), 4 per cell in the two-cell peak kernel (24 KiB: six
// workgroups of 256 lanes per CU, near the seven its registers allow; with
// 2 per cell its rows took 2.0x their bytes in HBM writes, with 4 1.08x)
constexpr int kStage = 4;
constexpr int kStage2 = 4;

template <class M, int WV>
__device__ __forceinline__ double pick(double x0, double x1, double x2)
{
    return WV == 0 ? x0 : (WV == 1 ? x1 : x2);
}

// One trajectory per lane.  counts[t] = P (after dropping the first peak)
// or -1 when the divergence check fired (the reference's flag value).
template <class M, int WV, bool CHECK>
__global__ __launch_bounds__(256) void k_bif_peaks(const BifArgs A, int* __restrict__ counts,
                                                   double* __restrict__ amps,
                                                   double* __restrict__ ivls)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;

    double v[kMaxVal];
    double x0, x1, x2;
    cell_setup(A.g, idx, v, x0, x1, x2);
    const typename M::P64 p = M::params(v);
    const double h = A.h;


    // transient, no divergence check (Library/cudaLibrary.cu:107-108)
    int s = 0;
    for (; s + 4 <= A.skip; s += 4) {
        M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
    }
    for (; s < A.skip; ++s) M::step(x0, x1, x2, p, h);

    // direct stores: this one-chain kernel serves small maps, the 1D diagram
    // and the overflow recompute, where it is latency-bound and an LDS stage
    // cost 3% (512^2 map); the two-cell and fused kernels stage their peaks
    // (their rows are the bulk of the peak-row traffic)
    PeakSink sink;
    sink.amps = amps ? amps + (long long)t * A.rowStride : nullptr;
    sink.ivls = ivls ? ivls + (long long)t * A.rowStride : nullptr;
    sink.cap = A.cap;
    sink.hpf = A.hpf;

    TrajPeakStream ps;
    ps.init();
    bool bad = false;
    const int N = A.N;
    const int pre = A.pre;
    int i = 0;
    for (; i < N && i < 2; ++i) {
        ps.prime(pick<M, WV>(x0, x1, x2), i);
        for (int j = 0; j < pre - 1; ++j) M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
        if (CHECK && diverged(pick<M, WV>(x0, x1, x2), A.maxValue)) { bad = true; break; }
    }
    if (!CHECK && pre == 1) {
        constexpr int K = kPeakBatch;
        for (; i + K <= N; i += K) {
            double va[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                va[k] = pick<M, WV>(x0, x1, x2);
                M::step(x0, x1, x2, p, h);
            }
            ps.push_batch(va, i, sink);
        }
        for (; i < N; ++i) {
            ps.push(pick<M, WV>(x0, x1, x2), i, sink);
            M::step(x0, x1, x2, p, h);
        }
    } else {
        for (; i < N && !bad; ++i) {
            ps.push(pick<M, WV>(x0, x1, x2), i, sink);
            for (int j = 0; j < pre - 1; ++j) M::step(x0, x1, x2, p, h);
            M::step(x0, x1, x2, p, h);
            if (CHECK && diverged(pick<M, WV>(x0, x1, x2), A.maxValue)) bad = true;
        }
    }
    ps.finish(sink);
    counts[t] = bad ? -1 : ps.count();
}

template <int WV>
__device__ __forceinline__ float pick32(float x0, float x1, float x2)
{
    return WV == 0 ? x0 : (WV == 1 ? x1 : x2);
}

// FP32 form of k_bif_peaks (bifhip_set_precision(BIFHIP_PREC_FP32)): the
// state in FP32, stepped with the model's fused FP32 map (M::step32, as the
// FP32 LLE kernel); the peak stream compares the float samples (the same
// outcomes as on the samples widened to double) and widens the peaks it
// writes.  Every update of the FP32 map
// contains the component itself, so NaN stays absorbing (the trajectory
// form of the stream applies).  No reference counterpart: the counts are
// judged by their agreement with the FP64 map.
template <class M, int WV, bool CHECK>
__global__ __launch_bounds__(256) void k_bif_peaks32(const BifArgs A, int* __restrict__ counts,
                                                     double* __restrict__ amps,
                                                     double* __restrict__ ivls)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;

    double v[kMaxVal];
    double s0, s1, s2;
    cell_setup(A.g, idx, v, s0, s1, s2);
    const typename M::P32 q = M::params32(v, A.h);
    float x0 = (float)s0, x1 = (float)s1, x2 = (float)s2;

    int s = 0;
    for (; s + 4 <= A.skip; s += 4) {
        M::step32(x0, x1, x2, q);
        M::step32(x0, x1, x2, q);
        M::step32(x0, x1, x2, q);
        M::step32(x0, x1, x2, q);
    }
    for (; s < A.skip; ++s) M::step32(x0, x1, x2, q);

    // direct stores, as k_bif_peaks (one chain per lane: an LDS stage cost 2%)
    PeakSink sink;
    sink.amps = amps ? amps + (long long)t * A.rowStride : nullptr;
    sink.ivls = ivls ? ivls + (long long)t * A.rowStride : nullptr;
    sink.cap = A.cap;
    sink.hpf = A.hpf;

    TrajPeakStream32 ps;
    ps.init();
    bool bad = false;
    const int N = A.N;
    const int pre = A.pre;
    int i = 0;
    for (; i < N && i < 2; ++i) {
        ps.prime(pick32<WV>(x0, x1, x2), i);
        for (int j = 0; j < pre - 1; ++j) M::step32(x0, x1, x2, q);
        M::step32(x0, x1, x2, q);
        if (CHECK && diverged((double)pick32<WV>(x0, x1, x2), A.maxValue)) { bad = true; break; }
    }
    if (!CHECK && pre == 1) {
        constexpr int K = kPeakBatch;
        for (; i + K <= N; i += K) {
            float va[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                va[k] = pick32<WV>(x0, x1, x2);
                M::step32(x0, x1, x2, q);
            }
            ps.push_batch(va, i, sink);
        }
        for (; i < N; ++i) {
            ps.push(pick32<WV>(x0, x1, x2), i, sink);
            M::step32(x0, x1, x2, q);
        }
    } else {
        for (; i < N && !bad; ++i) {
            ps.push(pick32<WV>(x0, x1, x2), i, sink);
            for (int j = 0; j < pre - 1; ++j) M::step32(x0, x1, x2, q);
            M::step32(x0, x1, x2, q);
            if (CHECK && diverged((double)pick32<WV>(x0, x1, x2), A.maxValue)) bad = true;
        }
    }
    ps.finish(sink);
    counts[t] = bad ? -1 : ps.count();
}

// below this many cells a launch keeps one cell per lane (occupancy first)
constexpr int kPairCells = 1 << 20;

// Same map without the divergence check, two cells per lane (cells 2t and
// 2t+1 of the launch) so that two independent trajectories interleave in
// the instruction stream (Rossler::step2); each cell's arithmetic is that
// of k_bif_peaks.  An odd last lane integrates its single cell twice and
// stores it once.

template <class M, int WV>
__global__ __launch_bounds__(256) void k_bif_peaks2(const BifArgs A, int* __restrict__ counts,
                                                    double* __restrict__ amps,
                                                    double* __restrict__ ivls)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = 2 * t;
    if (c0 >= A.nCells) return;
    const bool two = c0 + 1 < A.nCells;
    const int c1 = two ? c0 + 1 : c0;
    const int i0 = A.cellList ? A.cellList[c0] : A.cellBase + c0;
    const int i1 = A.cellList ? A.cellList[c1] : A.cellBase + c1;

    double v[kMaxVal];
    double x0, x1, x2, y0, y1, y2;
    cell_setup(A.g, i0, v, x0, x1, x2);
    const typename M::P64 p = M::params(v);
    cell_setup(A.g, i1, v, y0, y1, y2);
    const typename M::P64 q = M::params(v);
    const double h = A.h;

    int s = 0;
    for (; s + 2 <= A.skip; s += 2) {
        M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
        M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
    }
    for (; s < A.skip; ++s) M::step2(x0, x1, x2, p, y0, y1, y2, q, h);

    // two cells per lane, kStage2 peaks staged per cell
    __shared__ double s_stage[256 * 2 * kStage2];
    __shared__ int s_gap[256 * 2 * kStage2];
    PeakSink sa, sb;
    sa.amps = amps ? amps + (long long)c0 * A.rowStride : nullptr;
    sa.ivls = ivls ? ivls + (long long)c0 * A.rowStride : nullptr;
    sb.amps = amps && two ? amps + (long long)c1 * A.rowStride : nullptr;
    sb.ivls = ivls && two ? ivls + (long long)c1 * A.rowStride : nullptr;
    sa.cap = sb.cap = A.cap;
    sa.hpf = sb.hpf = A.hpf;
    sa.stage = (amps || ivls) ? s_stage + threadIdx.x * 2 * kStage2 : nullptr;
    sb.stage = sa.stage ? sa.stage + kStage2 : nullptr;
    sa.stageGap = s_gap + threadIdx.x * 2 * kStage2;
    sb.stageGap = sa.stageGap + kStage2;
    sa.stageN = sb.stageN = kStage2;
    TrajPeakStream pa, pb;
    pa.init();
    pb.init();
    const int N = A.N;
    int i = 0;
    for (; i < N && i < 2; ++i) {
        pa.prime(pick<M, WV>(x0, x1, x2), i);
        pb.prime(pick<M, WV>(y0, y1, y2), i);
        for (int j = 0; j < A.pre - 1; ++j) M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
        M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
    }
    if (A.pre == 1) {
        // K samples per iteration, tested together (push_batch)
        constexpr int K = kPeakBatch;
        for (; i + K <= N; i += K) {
            double va[K], vb[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                va[k] = pick<M, WV>(x0, x1, x2);
                vb[k] = pick<M, WV>(y0, y1, y2);
                M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
            }
            pa.push_batch(va, i, sa);
            pb.push_batch(vb, i, sb);
        }
        for (; i < N; ++i) {
            pa.push(pick<M, WV>(x0, x1, x2), i, sa);
            pb.push(pick<M, WV>(y0, y1, y2), i, sb);
            M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
        }
    } else {
        for (; i < N; ++i) {
            pa.push(pick<M, WV>(x0, x1, x2), i, sa);
            pb.push(pick<M, WV>(y0, y1, y2), i, sb);
            for (int j = 0; j < A.pre - 1; ++j) M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
            M::step2(x0, x1, x2, p, y0, y1, y2, q, h);
        }
    }
    pa.finish(sa);
    pb.finish(sb);
    counts[c0] = pa.count();
    if (two) counts[c1] = pb.count();
}

// Raw recorded series (the reference's data[] block) for parity tests.
template <class M>
__global__ __launch_bounds__(256) void k_series(const BifArgs A, int wv, double* __restrict__ data,
                                                int* __restrict__ flags)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal];
    double x0, x1, x2;
    cell_setup(A.g, idx, v, x0, x1, x2);
    const typename M::P64 p = M::params(v);
    const double h = A.h;

    for (int s = 0; s < A.skip; ++s) M::step(x0, x1, x2, p, h);
    double* row = data + (long long)t * A.rowStride;
    int flag = 0;
    for (int i = 0; i < A.N; ++i) {
        row[i] = wv == 0 ? x0 : (wv == 1 ? x1 : x2);
        for (int j = 0; j < A.pre - 1; ++j) M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
        if (A.maxValue != 0 && diverged(wv == 0 ? x0 : (wv == 1 ? x1 : x2), A.maxValue)) {
            flag = -1;
            break;
        }
    }
    flags[t] = flag;
}

// peakFinderCUDA over series in HBM: counts[c] == -1 -> 0, else the peak
// count; peaks/intervals written from the start of the cell's block.
// outPeaks may alias data (the 2D drivers find peaks in place): the write
// slot never passes the read index, so no __restrict__ here.
__global__ __launch_bounds__(256) void k_peak_finder(const double* data, int sizeOfBlock,
                                                     int nBlocks, int* __restrict__ counts,
                                                     double* outPeaks, double* outIntervals,
                                                     double hpf)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nBlocks) return;
    if (counts[c] == -1) { counts[c] = 0; return; }
    const double* d = data + (long long)c * sizeOfBlock;
    PeakSink sink;
    sink.amps = outPeaks ? outPeaks + (long long)c * sizeOfBlock : nullptr;
    sink.ivls = outIntervals ? outIntervals + (long long)c * sizeOfBlock : nullptr;
    sink.cap = sizeOfBlock;
    sink.hpf = hpf;
    sink.stage = nullptr;   // in place over the series: direct stores
    sink.stageN = 1;
    PeakStream ps;
    ps.init();
    if (sizeOfBlock >= 1) ps.p1 = d[0];
    if (sizeOfBlock >= 2) { ps.p2 = ps.p1; ps.p1 = d[1]; }
    for (int i = 2; i < sizeOfBlock; ++i) ps.push(d[i], i, sink);
    counts[c] = ps.count();
}

// ------------------------------------------------------------- DBSCAN
//
// Neighbour test of distance() (Library/cudaLibrary.cu:445-453):
//   hypotf((float)dx, (float)dy) < eps
// glibc's hypotf is (float)sqrt((double)fx*fx + (double)fy*fy) and the
// squares are exact in double, so the test is s < T with
//   s = (double)fx*fx + (double)fy*fy   (one rounding)
//   T = least double with (float)sqrt(T) >= eps   (computed on the host)
// Identical points give s = 0 < T exactly when eps > 0, as in the reference.
__device__ __forceinline__ bool near(double xa, double ya, double xb, double yb, double T)
{
    const float fx = (float)(xb - xa);
    const float fy = (float)(yb - ya);
    // both squares are exact in double, so the fused form rounds the sum
    // once, exactly as the separate product and sum do
    const double sy = (double)fy * (double)fy;
    return fma((double)fx, (double)fx, sy) < T;
}

constexpr int kDbR = 4;                  // register sets: P <= 64*kDbR
constexpr int kDbIsoSeeds = 2;           // isolated seeds before the isolated-peak pass
constexpr int kDbMaxP = kWave * kDbR;
// most cells per launch of a one-wavefront-per-cell kernel (2^30 work-items)
constexpr int kWaveCellsPerLaunch = 1 << 24;
// k_dbscan_wave workgroup size: one wavefront.  A cell's machine takes from
// a few dozen to thousands of steps, and a workgroup keeps its LDS until its
// last wavefront ends, so with four cells per workgroup the finished
// wavefronts' slots stayed idle behind the slowest (2.61 against 2.26 ms per
// 1M cells of the Rossler lattice; two per workgroup 2.39 ms).
constexpr int kDbBlock = kWave;

// Returned by dbscan_wave_cell if its machine has not finished after 4P+64
// steps.  It finishes in at most 3P+4 (each step seeds, pops or labels), so
// the bound only guarantees that the loop ends; no count can equal it.
constexpr int kDbUnfinished = -0x7fffffff - 1;

// Isolated peaks of one cell: bit k%64 of Z[k/64] is set when peak k < P has
// no neighbour among the valid peaks j < P-1, j != k.  Lanes hold the peaks
// (a, b); anchor j is read from the lane that holds it.  An anchor is tested
// exactly only when another lane's amplitude lies within B of its own:
// |dx| >= B = sqrt(T)(1 + 2^-20) makes (float)dx squared, and so the
// distance sum, exceed T, so such a pair is never near.  With T <= 0 no pair
// is near and every peak is isolated.
template <int NS>
__device__ __forceinline__ void isolated_peaks(int P, double T, const double (&a)[NS],
                                               const double (&b)[NS], const int (&kw)[NS],
                                               uint64_t (&Z)[NS])
{
    const int last = P - 1;
    int cnt[NS];                       // near anchors j != k of the lane's peak
#pragma unroll
    for (int w = 0; w < NS; ++w) cnt[w] = 0;
    if (T > 0.0) {
        const double B = sqrt(T) * (1.0 + 0x1p-20);
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int jend = min(kWave, last - r * kWave);
            uint64_t self = 1;                     // the anchor's own lane in word r
            for (int lj = 0; lj < jend; ++lj, self <<= 1) {
                const double aj = readlane_d(a[r], lj);
                uint64_t other = 0;
#pragma unroll
                for (int w = 0; w < NS; ++w) {
                    const uint64_t m = __builtin_amdgcn_ballot_w64(!(fabs(a[w] - aj) >= B));
                    other |= w == r ? (m & ~self) : m;
                }
                if (other != 0) {
                    const double bj = readlane_d(b[r], lj);
                    const int j = r * kWave + lj;
#pragma unroll
                    for (int w = 0; w < NS; ++w)
                        cnt[w] += (near(aj, bj, a[w], b[w], T) && kw[w] != j) ? 1 : 0;
                }
            }
        }
    }
#pragma unroll
    for (int w = 0; w < NS; ++w) Z[w] = __builtin_amdgcn_ballot_w64(kw[w] < P && cnt[w] == 0);
}

// Hashed amplitude buckets for the isolated-peak test: per wavefront
// kDbBuckets chain heads (peak index + 1, 0 = empty) and one link per peak.
constexpr int kDbBuckets = 1024;
constexpr int kDbWalkCap = 64;          // walk steps before the all-pairs test

// Value of peak j (per lane) held in lane j%64 of register set j/64, read
// with ds_bpermute; every lane of the wave must be active.
template <int NS>
__device__ __forceinline__ double gather_peak(const double (&v)[NS], int j)
{
    const int addr = (j & (kWave - 1)) << 2;
    double x = 0.0;
#pragma unroll
    for (int w = 0; w < NS; ++w) {
        const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v[w]));
        const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v[w]));
        if ((j >> 6) == w) x = __hiloint2double(hi, lo);
    }
    return x;
}

// The isolated-peak test of isolated_peaks with hashed buckets instead of
// all pairs.  A near pair has |dx| < B (see isolated_peaks), so its
// amplitudes fall in the same or adjacent buckets of width 2B:
// floor(a / 2B) differs by at most one (where a and a' are distinct and
// near, ulp(a) < B keeps a / 2B below 2^51, so its rounding moves it by at
// most 1/4; where ulp(a) >= B only equal amplitudes are near).  Bucket
// floor(a / 2B) mod kDbBuckets holds a chain of the valid anchors j < P-1
// with finite amplitudes; distant amplitudes that share a slot only add
// candidates, each tested exactly.  Peak k walks the chains of its slot and
// the two beside it until it meets an anchor j != k near it.  Non-finite
// amplitudes are near nothing: they are isolated and never anchors.
// Returns false, with Z incomplete, when some lane's walk is longer than
// kDbWalkCap steps; the caller then runs the all-pairs test.  Requires T > 0 and every lane of the wave active.
template <int NS>
__device__ __forceinline__ bool isolated_peaks_hashed(int P, double T, const double (&a)[NS],
                                                      const double (&b)[NS],
                                                      const int (&kw)[NS], int lane,
                                                      int* __restrict__ heads,
                                                      unsigned short* __restrict__ links,
                                                      uint64_t (&Z)[NS])
{
    const int last = P - 1;
    const double invW = 0.5 / (sqrt(T) * (1.0 + 0x1p-20));
    int slot[NS];                      // -1: no bucket (k >= P or non-finite)
#pragma unroll
    for (int w = 0; w < NS; ++w) {
        const double qd = floor(a[w] * invW);
        const bool ok = kw[w] < P && isfinite(a[w]) && isfinite(qd);
        slot[w] = ok ? (int)(qd - (double)kDbBuckets * floor(qd * (1.0 / kDbBuckets))) : -1;
    }
    // empty chains, then every anchor pushed on its chain
#pragma unroll
    for (int i = 0; i < kDbBuckets / (4 * kWave); ++i)
        reinterpret_cast<int4*>(heads)[i * kWave + lane] = int4{0, 0, 0, 0};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int w = 0; w < NS; ++w)
        if (slot[w] >= 0 && kw[w] < last)
            links[kw[w]] = (unsigned short)atomicExch(&heads[slot[w]], kw[w] + 1);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // walk, one 64-peak word at a time: bucket d of {slot-1, slot, slot+1}
    // (d = 0..2), node j (-1: this chain is exhausted)
#pragma unroll
    for (int w = 0; w < NS; ++w) {
        bool done = slot[w] < 0, hit = false;
        int d = 0;
        int j = done ? -1 : heads[(slot[w] + kDbBuckets - 1) & (kDbBuckets - 1)] - 1;
        for (int step = 0;; ++step) {
            if (__builtin_amdgcn_ballot_w64(!done) == 0) break;
            if (step == kDbWalkCap) return false;
            if (!done && j < 0) {
                // chain exhausted: the next bucket's head, or the end
                if (++d == 3) done = true;
                else j = heads[(slot[w] + d + kDbBuckets - 1) & (kDbBuckets - 1)] - 1;
            }
            const int jj = j < 0 ? 0 : j;
            const double aj = gather_peak<NS>(a, jj);
            const double bj = gather_peak<NS>(b, jj);
            if (!done && j >= 0) {
                if (j != kw[w] && near(aj, bj, a[w], b[w], T)) {
                    hit = true;
                    done = true;
                } else {
                    j = (int)links[j] - 1;
                }
            }
        }
        Z[w] = __builtin_amdgcn_ballot_w64(kw[w] < P && !hit);
    }
    return true;
}

// Peak k's value v held in lane k%64 of register set k/64 (k wave-uniform).
template <int NS>
__device__ __forceinline__ double peak_value(const double (&v)[NS], int k)
{
    double x = readlane_d(v[0], k & (kWave - 1));
#pragma unroll
    for (int w = 1; w < NS; ++w)
        if ((k >> 6) == w) x = readlane_d(v[w], k & (kWave - 1));
    return x;
}

// DBSCAN of one cell by one wavefront, specialised on the number NS of
// 64-peak words (P <= 64*NS).  Peak k lives in lane k%64 of register set
// k/64.  The label machine's sets are kept per peak, in the lane that holds
// the peak, as keys that make each search one vector compare:
//   uk  the peak's index while it is unlabelled and below P-1 (the only
//       peaks a scan can absorb), else -1 (whether P-1 is still unlabelled
//       is the flag lastFree, whether it is isolated the flag zlast)
//   zf  the peak is isolated (once known)
//   sk  the peak's index while its stack slot holds its own index, else
//       INT_MAX
// Updates (absorb a row, clear a popped slot or a seed, label a run) are
// vector selects; each search (the next seed, the highest set slot below
// nn, a count) is one ballot of one compare per word and a bit scan, so the
// scalar unit, which the CU's waves share, only finds bits and steers the
// loop.
//  - a neighbour row is built, one ballot per word, only when the machine
//    scans it: for a cluster seed or a popped index; nothing is scanned
//    once every absorbable index is labelled.  A period-1 orbit scans one
//    row, a period-2 orbit about a quarter of them.
//  - after kDbIsoSeeds seeds with empty rows (a chaotic cell, whose peaks
//    are mostly isolated) the isolated peaks are found at once
//    (isolated_peaks_hashed, else isolated_peaks), and from then on runs of
//    isolated seeds are labelled without scans.
template <int NS>
__device__ __forceinline__ int dbscan_wave_cell(const double* __restrict__ ra,
                                                const double* __restrict__ rb, int P, double T,
                                                int lane, int* __restrict__ heads,
                                                unsigned short* __restrict__ links)
{
    constexpr int kNoSlot = 0x7fffffff;
    // lanes past the last peak get an infinite amplitude: never near and
    // never a candidate of the isolated-peak test
    double a[NS], b[NS];
    int kw[NS], uk[NS], sk[NS];
    bool zf[NS];
#pragma unroll
    for (int w = 0; w < NS; ++w) {
        kw[w] = w * kWave + lane;
        a[w] = kw[w] < P ? ra[kw[w]] : __builtin_inf();
        b[w] = kw[w] < P ? rb[kw[w]] : 0.0;
    }
    const int last = P - 1;
#pragma unroll
    for (int w = 0; w < NS; ++w) {
        uk[w] = kw[w] < last ? kw[w] : -1;
        sk[w] = kNoSlot;
        zf[w] = false;
    }
    bool lastFree = true, zlast = false;
    int isoSeeds = 0;                  // seeds with empty rows before zf is known
    // The reference's loop (i = 0 .. P-1, a stack pop jumping to i = slot or
    // 0, then i + 1) runs as two phases.  Steps that provably change nothing
    // but i and nn are taken in bulk, with the state the reference reaches
    // after them:
    //  - rem counts the unlabelled indices below P-1, the only ones a scan
    //    can absorb.  At rem == 0 pops and scans change no label, P-1 is
    //    never absorbed (scans test k < P-1) and the walk still reaches it
    //    (a popped slot is at most P-2), so it opens one more cluster
    //    exactly when unlabelled: the machine stops there.
    //  - walk (empty stack): the unlabelled indices >= i are visited in
    //    order.  An isolated index (empty row) only opens a cluster of its
    //    own, so once zf is known every isolated unlabelled index before the
    //    next non-isolated one is labelled at once (with no zf the next
    //    unlabelled index is the seed).  A seed at P-1 ends the reference's
    //    loop after its scan.
    //  - pops: slot nn-1 is popped.  A pop of an unset slot sets i = 0 and
    //    scans row 0, which was scanned first (the walk starts at 0), so it
    //    has no unlabelled neighbour left: the run of unset slots above the
    //    highest set slot hs below nn is popped at once, leaving i = 1, and
    //    hs is popped next (or, with no set slot, the walk resumes at 1).
    // Scan of the row of peak (ai, bi): the unlabelled peaks near it leave
    // uk and enter sk; returns how many.  A lane's distance sum is made
    // infinite where its peak is labelled, so one compare tests both.
    auto scan = [&](double ai, double bi) {
        int c = 0;
#pragma unroll
        for (int w = 0; w < NS; ++w) {
            const float fx = (float)(a[w] - ai);
            const float fy = (float)(b[w] - bi);
            const double sy = (double)fy * (double)fy;
            double sd = fma((double)fx, (double)fx, sy);
            sd = uk[w] >= 0 ? sd : __builtin_inf();
            const bool got = sd < T;
            uk[w] = got ? -1 : uk[w];
            sk[w] = got ? kw[w] : sk[w];
            c += __popcll(__builtin_amdgcn_ballot_w64(got));
        }
        return c;
    };
    int nn = 0, cluster = 0, i = 0, rem = last;
    int guard = 4 * P + 64;
    for (;;) {
        // walk: nn == 0, rem > 0, i < P
        if (--guard < 0) return kDbUnfinished;
        unsigned qu = ~0u;                 // next non-isolated unlabelled >= i
#pragma unroll
        for (int w = 0; w < NS; ++w) {
            const int open = zf[w] ? -1 : uk[w];
            const uint64_t m = __builtin_amdgcn_ballot_w64(open >= i);
            const unsigned f = m != 0 ? (unsigned)__builtin_ctzll(m) : ~0u;
            qu = min(qu, f | (unsigned)(w * kWave));
        }
        const int q = qu != ~0u ? (int)qu : (lastFree && !zlast ? last : P);
        if (isoSeeds < 0) {
            // the isolated unlabelled indices in [i, q) each open a cluster
            // of their own; P-1 is among them when q == P
            int c = 0;
#pragma unroll
            for (int w = 0; w < NS; ++w) {
                const int iso = zf[w] ? uk[w] : -1;
                const bool inRun = (unsigned)(iso - i) < (unsigned)(q - i);
                uk[w] = inRun ? -1 : uk[w];
                c += __popcll(__builtin_amdgcn_ballot_w64(inRun));
            }
            cluster += c;
            rem -= c;
            if (q == P && lastFree && zlast) {
                ++cluster;
                lastFree = false;
            }
            if (rem == 0) break;
        }
        if (q >= P) return cluster - 1;
        ++cluster;
        if (q < last) {
#pragma unroll
            for (int w = 0; w < NS; ++w) uk[w] = uk[w] == q ? -1 : uk[w];
            if (--rem == 0) break;
        } else {
            lastFree = false;
        }
        // the seed's row: its unlabelled neighbours k < P-1 (q itself is
        // labelled, so k != q holds through uk)
        nn = scan(peak_value<NS>(a, q), peak_value<NS>(b, q));
        rem -= nn;
        if (rem == 0) break;
        if (q == last) return cluster - 1;     // i = P ends the reference's loop
        if (nn == 0 && isoSeeds >= 0 && ++isoSeeds == kDbIsoSeeds) {
            uint64_t Z[NS];
            if (!(T > 0.0) ||
                !isolated_peaks_hashed<NS>(P, T, a, b, kw, lane, heads, links, Z))
                isolated_peaks<NS>(P, T, a, b, kw, Z);
#pragma unroll
            for (int w = 0; w < NS; ++w) {
                zf[w] = ((Z[w] >> lane) & 1ull) != 0;
                if (w == (last >> 6)) zlast = ((Z[w] >> (last & 63)) & 1ull) != 0;
            }
            isoSeeds = -1;                 // known
        }
        i = q + 1;
        // pops
        while (nn >= 1) {
            if (--guard < 0) return kDbUnfinished;
            int hs = -1;                   // highest set slot below nn
#pragma unroll
            for (int w = 0; w < NS; ++w) {
                const uint64_t m = __builtin_amdgcn_ballot_w64(sk[w] < nn);
                if (m != 0) hs = w * kWave + 63 - __builtin_clzll(m);
            }
            if (hs < 0) {
                nn = 0;
                i = 1;
                break;
            }
#pragma unroll
            for (int w = 0; w < NS; ++w) sk[w] = sk[w] == hs ? kNoSlot : sk[w];
            // hs is labelled, so the scan cannot put it back in sk
            const int c = scan(peak_value<NS>(a, hs), peak_value<NS>(b, hs));
            nn = hs + c;
            rem -= c;
            if (rem == 0) break;
            i = hs + 1;
        }
        if (rem == 0) break;
    }
    // rem == 0: no scan can label anything; P-1 opens a last cluster when
    // still unlabelled
    return cluster - 1 + (lastFree ? 1 : 0);
}

// One wavefront per cell (dbscan_wave_cell holds the peaks in registers
// and builds the neighbour rows it scans).  The reference's machine:
//   U = unlabelled k (label == 0), S = stack slots holding their own index
//   (slot k is written with value k; popping slot NumNeibor-1 yields that
//   slot's index if set, else 0; the popped value becomes the loop
//   variable, then ++).  Result cluster-1.
// Cells whose peaks do not fit (P > min(cap, 256), P <= maxPeaks) are
// appended to overflow[] as listBase + cell and left to the workgroup and
// lane kernels.  Eight waves per SIMD are asked for: the machine is a chain
// of dependent compare / mask steps, and more waves hide its latency.  The
// register budget that allows them (78 scalar registers, some spilled to
// vector lanes) costs less than it gains: without the bound the kernel used
// 112 scalar registers and ran 8% slower.
__global__ __launch_bounds__(kDbBlock, 8) void k_dbscan_wave(int nCells, const int* __restrict__ counts,
                                                     const double* __restrict__ amps,
                                                     const double* __restrict__ ivls,
                                                     long long rowStride, int cap, int maxPeaks,
                                                     double T, int* __restrict__ out,
                                                     int* __restrict__ overflow,
                                                     int* __restrict__ nOverflow, int listBase)
{
    const int cell = __builtin_amdgcn_readfirstlane(
        (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave));
    if (cell >= nCells) return;          // wave-uniform
    const int lane = lane_id();
    const int P = __builtin_amdgcn_readfirstlane(counts[cell]);
    // each wavefront's bucket heads and peak links (isolated_peaks_hashed)
    __shared__ __attribute__((aligned(16))) int s_heads[kDbBlock / kWave * kDbBuckets];
    __shared__ unsigned short s_links[kDbBlock / kWave * kDbMaxP];
    int* heads = s_heads + (threadIdx.x / kWave) * kDbBuckets;
    unsigned short* links = s_links + (threadIdx.x / kWave) * kDbMaxP;
    int res;
    if (P <= 0) res = 0;                 // -1 (diverged) and 0
    else if (P == 1) res = 1;
    else if (P > maxPeaks) res = 0;
    else if (P > cap || P > kDbMaxP) {
        // handled by the workgroup / lane kernels
        if (lane == 0) {
            const int k = atomicAdd(nOverflow, 1);
            overflow[k] = listBase + cell;
        }
        return;
    } else {
        const double* ra = amps + (long long)cell * rowStride;
        const double* rb = ivls + (long long)cell * rowStride;
        const int nsets = (P + kWave - 1) / kWave;
        if (nsets == 1) res = dbscan_wave_cell<1>(ra, rb, P, T, lane, heads, links);
        else if (nsets == 2) res = dbscan_wave_cell<2>(ra, rb, P, T, lane, heads, links);
        else if (nsets == 3) res = dbscan_wave_cell<3>(ra, rb, P, T, lane, heads, links);
        else res = dbscan_wave_cell<4>(ra, rb, P, T, lane, heads, links);
    }
    if (lane == 0) out[cell] = res;
}

// Lane-per-cell fallback for large P, a direct restatement of the reference
// loop on an int scratch row (labels [0,P), slots [P,2P)).  Row r = list[t]
// (or t); result to out[outIndex[t]] (or out[r]).  Rows must hold all P
// peaks and P must not exceed maxPeaks (the caller filters).
__global__ __launch_bounds__(64) void k_dbscan_lane(int nList, const int* __restrict__ list,
                                                    const int* __restrict__ counts,
                                                    const double* __restrict__ amps,
                                                    const double* __restrict__ ivls,
                                                    long long rowStride, double T,
                                                    int* __restrict__ scratch,
                                                    long long scratchStride,
                                                    int* __restrict__ out,
                                                    const int* __restrict__ outIndex)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nList) return;
    const int cell = list ? list[t] : t;
    const int o = outIndex ? outIndex[t] : cell;
    const int P = counts[cell];
    if (P > rowStride || 2LL * P > scratchStride) {  // counts larger than the rows: invalid input
        out[o] = -1;
        return;
    }
    const double* A = amps + (long long)cell * rowStride;
    const double* B = ivls + (long long)cell * rowStride;
    int* H = scratch + (long long)t * scratchStride;
    for (int i = 0; i < 2 * P; ++i) H[i] = 0;
    // rem = unlabelled indices below P-1 (the only ones a scan can absorb);
    // at rem == 0 scans are no-ops and only index P-1 can open a cluster
    int nn = 0, cluster = 0, rem = P - 1;
    bool lastFree = true;
    for (int i = 0; i < P; i++) {
        bool scan = false;
        if (nn >= 1) {
            i = H[P + nn - 1];
            H[P + nn - 1] = 0;
            nn = nn - 1;
            scan = true;
        } else if (H[i] == 0) {
            nn = 0;
            ++cluster;
            H[i] = cluster;
            if (i < P - 1) --rem;
            else lastFree = false;
            scan = true;
        } else if (rem == 0) {
            if (lastFree) ++cluster;
            break;
        }
        if (scan && rem > 0) {
            const double ai = A[i], bi = B[i];
            for (int k = 0; k < P - 1; k++) {
                if (i != k && H[k] == 0 && near(ai, bi, A[k], B[k], T)) {
                    H[k] = cluster;
                    H[P + k] = k;
                    ++nn;
                    --rem;
                }
            }
        }
    }
    out[o] = cluster - 1;
}

// Workgroup-per-cell DBSCAN for cells with more peaks than the wave kernel
// holds (P <= kBlkMaxP).  Same machine as k_dbscan_wave with the label
// state as bitmasks in LDS: U = unlabelled, S = stack slots holding their
// own index.  A scan of row i tests every unlabelled k < P-1 in parallel
// (each k is owned by one thread; its U/S bits change with LDS atomics)
// and pushes the matches; the order of the pushes within a scan does not
// matter because slot k only ever holds k.  Peaks are staged in LDS when
// P <= kBlkLdsP, else read from the rows.  Once no index below P-1 is
// unlabelled, scans cannot change anything and are skipped; the walk then
// ends after at most the last index (never absorbed by a scan) opens one
// more cluster.  Row r = list[t] (or t); result to out[outIndex[t]]
// (or out[r]).
constexpr int kBlkLdsP = 3072;
constexpr int kBlkMaxP = 16384;

__global__ __launch_bounds__(256) void k_dbscan_block(int nList, const int* __restrict__ list,
                                                      const int* __restrict__ counts,
                                                      const double* __restrict__ amps,
                                                      const double* __restrict__ ivls,
                                                      long long rowStride, double T,
                                                      int* __restrict__ out,
                                                      const int* __restrict__ outIndex)
{
    __shared__ double s_a[kBlkLdsP], s_b[kBlkLdsP];
    __shared__ uint64_t s_U[kBlkMaxP / 64], s_S[kBlkMaxP / 64];
    __shared__ int s_cnt[2], s_next;
    const int t = blockIdx.x;
    if (t >= nList) return;                        // block-uniform
    const int row = list ? list[t] : t;
    const int o = outIndex ? outIndex[t] : row;
    const int P = counts[row];
    if (P <= 0 || P > rowStride || P > kBlkMaxP) { // not handled here (caller filters)
        if (threadIdx.x == 0) out[o] = P == 1 ? 1 : (P <= 0 ? 0 : -1);
        return;
    }
    if (P == 1) {
        if (threadIdx.x == 0) out[o] = 1;
        return;
    }
    const double* ra = amps + (long long)row * rowStride;
    const double* rb = ivls + (long long)row * rowStride;
    const bool inLds = P <= kBlkLdsP;
    if (inLds)
        for (int k = threadIdx.x; k < P; k += blockDim.x) {
            s_a[k] = ra[k];
            s_b[k] = rb[k];
        }
    const double* A = inLds ? s_a : ra;
    const double* B = inLds ? s_b : rb;
    const int nWords = (P + 63) / 64;
    for (int w = threadIdx.x; w < nWords; w += blockDim.x) {
        const int n = P - w * 64;
        s_U[w] = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
        s_S[w] = 0;
    }
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();

    // uniform state, computed redundantly by every thread
    int nn = 0, cluster = 0, i = 0, rem = P - 1, parity = 0;
    bool lastFree = true;
    while (i < P) {
        if (rem == 0) {
            // pending pops change no label; only index P-1 can still open
            // a cluster (block-uniform exit: rem is the same in every thread)
            if (lastFree) ++cluster;
            break;
        }
        if (nn >= 1) {
            const int slot = nn - 1;
            const uint64_t bit = 1ull << (slot & 63);
            const bool set = (s_S[slot >> 6] & bit) != 0;
            __syncthreads();                       // all have read before the clear
            if (threadIdx.x == 0) s_S[slot >> 6] &= ~bit;
            i = set ? slot : 0;
            nn = nn - 1;
        } else {
            if (threadIdx.x == 0) s_next = P;
            __syncthreads();
            for (int w = (i >> 6) + (int)threadIdx.x; w < nWords; w += blockDim.x) {
                uint64_t m = s_U[w];
                if (w == (i >> 6)) m &= ~0ull << (i & 63);
                if (m) {
                    atomicMin(&s_next, w * 64 + __builtin_ctzll(m));
                    break;
                }
            }
            __syncthreads();
            const int nxt = s_next;
            if (nxt >= P) break;
            i = nxt;
            ++cluster;
            if (threadIdx.x == 0) s_U[i >> 6] &= ~(1ull << (i & 63));
            if (i < P - 1) --rem;
            else lastFree = false;
        }
        if (rem > 0) {
            __syncthreads();                       // U/S updates of thread 0 visible
            const double ai = A[i], bi = B[i];
            int found = 0;
            for (int k = threadIdx.x; k < P - 1; k += blockDim.x) {
                const uint64_t bit = 1ull << (k & 63);
                if (k != i && (s_U[k >> 6] & bit) && near(ai, bi, A[k], B[k], T)) {
                    atomicAnd((unsigned long long*)&s_U[k >> 6], ~(unsigned long long)bit);
                    atomicOr((unsigned long long*)&s_S[k >> 6], (unsigned long long)bit);
                    ++found;
                }
            }
            if (found) atomicAdd(&s_cnt[parity], found);
            if (threadIdx.x == 0) s_cnt[parity ^ 1] = 0;   // next scan's counter
            __syncthreads();
            const int c = s_cnt[parity];
            nn += c;
            rem -= c;
            parity ^= 1;
        }
        ++i;
    }
    if (threadIdx.x == 0) out[o] = cluster - 1;
}

// ---------------------------------------------------------------- KDE
//
// expf of the reference's KDE kernel evaluated as exp in double rounded to
// float (correctly rounded except within ~1 double ulp of a float midpoint;
// glibc's expf differs from correct rounding on ~1e-4 of its inputs, so
// KDE counts are reported as agreement, not claimed bit-exact).
__device__ __forceinline__ double kde_term(double pm, double delt, double H)
{
    const double t = (pm - delt) / H;
    const float arg = (float)(-((t * t) / 2));
    return (double)(float)exp((double)arg);
}

// A term is exactly +0 once |pm - delt| > 15|H|: then |t| > 14.99, the
// float argument is below -112, and exp of it (< 2^-161) rounds to +0 in
// float (expf underflows to 0 below -103.97).  Adding +0 leaves the sum
// unchanged, so such terms may be skipped without changing any sum.  NaN
// differences never compare greater and are evaluated; infinite ones give
// +0 either way; H = 0 skips exactly the terms that are +0 (d / 0 = +-inf).
__device__ __forceinline__ double kde_skip_distance(double H) { return 15.0 * fabs(H); }

// k_kde_wave workgroup size: one wavefront, so a finished cell's slot is
// refilled at once (cells differ in their evaluated terms by orders of
// magnitude; 3.78 against 4.31 ms per 512^2 map with four per workgroup)
constexpr int kKdeBlock = kWave;

// Lane l receives v of lane l - k (lanes below k receive lane 0's); every
// lane of the wave must be active.
__device__ __forceinline__ double lane_shift_up_d(double v, int k)
{
    const int src = max((int)lane_id() - k, 0);
    const int lo = __builtin_amdgcn_ds_bpermute(src * 4, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src * 4, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Rows hold up to cap peaks; cells with cap < P <= maxPeaks are appended to
// overflow[] (listBase + cell) for recomputation.  Output to
// out[outIndex[cell]] when outIndex is given.
__global__ __launch_bounds__(kKdeBlock) void k_kde_wave(int nCells, const int* __restrict__ counts,
                                                  const double* __restrict__ amps,
                                                  long long rowStride, int cap, int maxPeaks,
                                                  int kdeSampling, double I1, double I2,
                                                  double H, int* __restrict__ out,
                                                  const int* __restrict__ outIndex,
                                                  int* __restrict__ overflow,
                                                  int* __restrict__ nOverflow, int listBase)
{
    const int cell = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (cell >= nCells) return;
    const int lane = lane_id();
    const int P = __builtin_amdgcn_readfirstlane(counts[cell]);
    int res;
    if (P <= 0) res = 0;                         // -1 (diverged) and 0
    else if (P == 1 || P == 2) res = 1;
    else if (P > maxPeaks) res = maxPeaks;
    else if (P > cap) {
        if (lane == 0) {
            const int k = atomicAdd(nOverflow, 1);
            overflow[k] = listBase + cell;
        }
        return;
    } else {
        const double* pk = amps + (long long)cell * rowStride;
        const double k1 = (double)(kdeSampling * P);
        const double k2 = (I2 - I1) / (k1 - 1);
        const double X = kde_skip_distance(H);
        double pp = 0, p = 0, cur = 0;     // the machine's last three sums
        bool strange = false;
        int result = 0;
        for (int base = 0; (double)base < k1 - 1; base += kWave) {
            const int w = base + lane;
            // lanes 0 .. wend-1 hold sample points w < k1 - 1
            const int wend = min(kWave, (int)ceil(k1 - 1) - base);
            const double delt = w * k2 + I1;
            // delt is monotone in w, so with finite ends every point of the
            // batch lies in [lo, hi].  A peak with pm - hi > X (or pm - lo
            // < -X) then has pm - delt beyond X for every lane, and adds +0
            // to every sum (kde_skip_distance): the wave skips it.
            const double d0 = readlane_d(delt, 0), d1 = readlane_d(delt, wend - 1);
            const bool bounded = isfinite(d0) && isfinite(d1) && isfinite(X);
            const double lo = fmin(d0, d1), hi = fmax(d0, d1);
            double val = 0;
            for (int c0 = 0; c0 < P; c0 += kWave) {
                const int j = c0 + lane;
                const double pj = j < P ? pk[j] : 0.0;
                const bool far = bounded && ((pj - hi) > X || (pj - lo) < -X);
                uint64_t live = __builtin_amdgcn_ballot_w64(j < P && !far);
                while (live != 0) {            // the live peaks, in the reference's order
                    const int l = __builtin_ctzll(live);
                    live &= live - 1;
                    val += kde_term(readlane_d(pj, l), delt, H);
                }
            }
            // The reference's local-maximum machine over this batch's sums.
            // With no plateau open (strange) and none opening here, it only
            // counts the points with pp < p > cur: lane l's (pp, p, cur) are
            // the sums at w-2, w-1, w, from the lanes below it or the state.
            const double v1 = lane_shift_up_d(val, 1), v2 = lane_shift_up_d(val, 2);
            const double pl = lane >= 1 ? v1 : cur;
            const double ppl = lane >= 2 ? v2 : (lane == 1 ? cur : p);
            const bool eval = lane < wend && w >= 2 && pl > ppl;
            if (!strange && __builtin_amdgcn_ballot_w64(eval && pl == val) == 0) {
                result += __popcll(__builtin_amdgcn_ballot_w64(eval && pl > val));
                if (wend >= 3) {
                    pp = readlane_d(val, wend - 3);
                    p = readlane_d(val, wend - 2);
                } else if (wend == 2) {
                    pp = cur;
                    p = readlane_d(val, 0);
                } else {
                    pp = p;
                    p = cur;
                }
                cur = readlane_d(val, wend - 1);
                continue;
            }
            // a plateau is open or opens in this batch: step the machine
            // sum by sum
            for (int l = 0; l < wend; ++l) {
                const int ww = base + l;
                pp = p;
                p = cur;
                cur = readlane_d(val, l);
                if (ww < 2) continue;
                if (strange) {
                    if (p == cur) continue;
                    else if (p < cur) { strange = false; continue; }
                    else if (p > cur) { strange = false; ++result; continue; }
                } else if (p > pp && p > cur) { ++result; continue; }
                else if (p > pp && p == cur) { strange = true; continue; }
            }
        }
        if (p < cur) ++result;
        res = result;
    }
    if (lane == 0) out[outIndex ? outIndex[cell] : cell] = res;
}

// ------------------------------------------- fused LLE + peak-count map
//
// The LLE's reference trajectory x starts from the same state, with the
// same parameters and step, as the bifurcation trajectory, so it passes
// through every state the peak window records (e.g. transient 500 + 1000
// time units of x cover the peak map's samples at 1000..1500).  One lane
// therefore produces both results with the operation sequence of each
// reference kernel: LLEKernelCUDA (Library/cudaLibrary.cu:666-785) for the
// exponent, calculateDiscreteModelCUDA + peakFinder (:78-117, :220-266)
// for the peaks -- bit for bit, with 250 000 instead of 400 000 steps per
// cell at the benchmark's settings.  No divergence check (maxValue == 0);
// preScaller 1; the window must lie inside x's span (checked on the host).
template <class M, int WV>
__global__ __launch_bounds__(256) void k_lle_peaks(const FusedArgs F, double* __restrict__ lle,
                                                   int* __restrict__ counts,
                                                   double* __restrict__ amps,
                                                   double* __restrict__ ivls)
{
    const BifArgs& A = F.pk;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nCells) return;
    const int idx = A.cellList ? A.cellList[t] : A.cellBase + t;
    double v[kMaxVal];
    double x0, x1, x2;
    cell_setup(A.g, idx, v, x0, x1, x2);
    const typename M::P64 p = M::params(v);
    const double h = A.h, eps = F.eps;


    __shared__ double s_stage[256 * kStage];   // 4 peaks per lane
    __shared__ int s_gap[256 * kStage];
    PeakSink sink;
    sink.amps = amps + (long long)t * A.rowStride;
    sink.ivls = ivls + (long long)t * A.rowStride;
    sink.cap = A.cap;
    sink.hpf = A.hpf;
    sink.stage = s_stage + threadIdx.x * kStage;
    sink.stageGap = s_gap + threadIdx.x * kStage;
    sink.stageN = kStage;
    TrajPeakStream ps;
    ps.init();
    const int w0 = A.skip, w1 = A.skip + A.N;   // window of recorded states [w0, w1)

    // sample the state reached after k steps of x, if k is in the window
    auto record = [&](int k) {
        if (k >= w0 && k < w1) {
            const double d = WV == 0 ? x0 : (WV == 1 ? x1 : x2);
            const int i = k - w0;
            if (i >= 2) ps.push(d, i, sink);
            else ps.prime(d, i);
        }
    };

    // transient of the LLE: x alone; states inside the window are recorded
    int k = 0;
    const int preEnd = min(F.skipL, max(0, w0));
    for (; k + 4 <= preEnd; k += 4) {
        M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
        M::step(x0, x1, x2, p, h);
    }
    for (; k < preEnd; ++k) M::step(x0, x1, x2, p, h);
    for (; k < F.skipL; ++k) {
        record(k);
        M::step(x0, x1, x2, p, h);
    }
    double z0 = 1, z1 = 0, z2 = 0;
    if (F.dirs) {
        const double* d = F.dirs + (long long)t * 3;
        z0 = d[0]; z1 = d[1]; z2 = d[2];
    }
    double y0 = z0 * eps + x0, y1 = z1 * eps + x1, y2 = z2 * eps + x2;
    double result = 0;
    for (int b = 0; b < F.nBlocks; ++b) {
        if (k + F.nNT <= w0 || k >= w1) {
            // no recorded state in this block: plain interleaved stepping
            int i = 0;
            for (; i + 2 <= F.nNT; i += 2) {
                M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
                M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
            }
            for (; i < F.nNT; ++i) M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
            k += F.nNT;
        } else if (k >= w0 + 2 && k + F.nNT <= w1) {
            // every state of this block is a recorded sample past the first
            // two; K samples per iteration, tested together (push_batch)
            constexpr int K = kPeakBatch;
            int i = 0;
            for (; i + K <= F.nNT; i += K, k += K) {
                double va[K];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) {
                    va[kk] = WV == 0 ? x0 : (WV == 1 ? x1 : x2);
                    M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
                }
                ps.push_batch(va, k - w0, sink);
            }
            for (; i < F.nNT; ++i, ++k) {
                ps.push(WV == 0 ? x0 : (WV == 1 ? x1 : x2), k - w0, sink);
                M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
            }
        } else {
            for (int i = 0; i < F.nNT; ++i, ++k) {
                record(k);
                M::step2(x0, x1, x2, p, y0, y1, y2, p, h);
            }
        }
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
    lle[t] = result / F.tMax;
    ps.finish(sink);
    counts[t] = ps.count();
}

hipError_t launch_lle_peaks(int model, int wv, const FusedArgs& F, double* lle, int* counts,
                            double* amps, double* ivls, hipStream_t st)
{
    if (F.pk.nCells <= 0) return hipSuccess;
    const dim3 blk(256), grd((F.pk.nCells + 255) / 256);
#define FUSED_CASE(MOD, W) \
    hipLaunchKernelGGL((k_lle_peaks<MOD, W>), grd, blk, 0, st, F, lle, counts, amps, ivls)
    if (model == kLinear) {
        if (wv == 0) FUSED_CASE(Linear, 0); else if (wv == 1) FUSED_CASE(Linear, 1); else FUSED_CASE(Linear, 2);
    } else {
        if (wv == 0) FUSED_CASE(Rossler, 0); else if (wv == 1) FUSED_CASE(Rossler, 1); else FUSED_CASE(Rossler, 2);
    }
#undef FUSED_CASE
    return hipGetLastError();
}

// --------------------------------------------------- integer post-ops

// Cluster-count clamp min(v, clamp) applied in place (clamp > 0).
__global__ void k_clamp_i32(int n, int* __restrict__ v, int clamp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && v[i] > clamp) v[i] = clamp;
}

// The used prefix of each peak row, packed for one contiguous copy:
// dst[c*stride + j] = src[c*rowCap + j] for j < stride.
__global__ void k_pack_rows(long long nCells, int rowCap, int stride,
                            const double* __restrict__ src, double* __restrict__ dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nCells * stride) return;
    const long long c = i / stride;
    const int j = (int)(i - c * stride);
    dst[i] = src[c * rowCap + j];
}

// Fill kernel (used instead of hipMemset so graph-captured work stays
// kernel-only).
__global__ void k_fill_i32(long long n, int* __restrict__ v, int value)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = value;
}

// --------------------------------------------------------- launchers

template <class M>
static hipError_t launch_bif_m(const BifArgs& A, int wv, bool check, bool fp32, int* counts,
                               double* amps, double* ivls, hipStream_t st)
{
    const dim3 blk(256), grd((A.nCells + 255) / 256);
    if (A.nCells <= 0) return hipSuccess;
    if (fp32) {
#define BIF32_CASE(W, C) hipLaunchKernelGGL((k_bif_peaks32<M, W, C>), grd, blk, 0, st, A, counts, amps, ivls)
        if (check) {
            if (wv == 0) BIF32_CASE(0, true); else if (wv == 1) BIF32_CASE(1, true); else BIF32_CASE(2, true);
        } else {
            if (wv == 0) BIF32_CASE(0, false); else if (wv == 1) BIF32_CASE(1, false); else BIF32_CASE(2, false);
        }
#undef BIF32_CASE
        return hipGetLastError();
    }
    if (!check && A.nCells >= kPairCells) {
        // enough cells to fill the device at two per lane
        const dim3 grd2((A.nCells / 2 + 1 + 255) / 256);
#define BIF2_CASE(W) hipLaunchKernelGGL((k_bif_peaks2<M, W>), grd2, blk, 0, st, A, counts, amps, ivls)
        if (wv == 0) BIF2_CASE(0); else if (wv == 1) BIF2_CASE(1); else BIF2_CASE(2);
#undef BIF2_CASE
        return hipGetLastError();
    }
#define BIF_CASE(W, C) hipLaunchKernelGGL((k_bif_peaks<M, W, C>), grd, blk, 0, st, A, counts, amps, ivls)
    if (check) {
        if (wv == 0) BIF_CASE(0, true); else if (wv == 1) BIF_CASE(1, true); else BIF_CASE(2, true);
    } else {
        if (wv == 0) BIF_CASE(0, false); else if (wv == 1) BIF_CASE(1, false); else BIF_CASE(2, false);
    }
#undef BIF_CASE
    return hipGetLastError();
}

// Timer family of the kernel launch_bif runs for these arguments:
// "bif_peaks2" for the two-cell FP64 kernel, else "bif_peaks".
const char* bif_kernel_family(int model, int precision, long long nCells, bool check)
{
    const bool two = model != kMemristive && precision != 1 && !check && nCells >= kPairCells;
    return two ? "bif_peaks2" : "bif_peaks";
}

hipError_t launch_gen_bif(int model, const BifArgs& A, int wv, bool check, int* counts,
                          double* amps, double* ivls, double* last, hipStream_t st);
hipError_t launch_gen_series(int model, const BifArgs& A, int wv, double* data, int* flags,
                             hipStream_t st);

// precision 1: the FP32 peak kernel (Rossler and linear models only)
hipError_t launch_bif(int model, int precision, const BifArgs& A, int wv, bool check, int* counts,
                      double* amps, double* ivls, hipStream_t st)
{
    const bool fp32 = precision == 1;
    if (model == kMemristive) {
        if (fp32) return hipErrorInvalidValue;
        return launch_gen_bif(model, A, wv, check, counts, amps, ivls, nullptr, st);
    }
    if (model == kLinear) return launch_bif_m<Linear>(A, wv, check, fp32, counts, amps, ivls, st);
    return launch_bif_m<Rossler>(A, wv, check, fp32, counts, amps, ivls, st);
}

hipError_t launch_series(int model, const BifArgs& A, int wv, double* data, int* flags,
                         hipStream_t st)
{
    if (A.nCells <= 0) return hipSuccess;
    if (model == kMemristive) return launch_gen_series(model, A, wv, data, flags, st);
    const dim3 blk(256), grd((A.nCells + 255) / 256);
    if (model == kLinear)
        hipLaunchKernelGGL(k_series<Linear>, grd, blk, 0, st, A, wv, data, flags);
    else
        hipLaunchKernelGGL(k_series<Rossler>, grd, blk, 0, st, A, wv, data, flags);
    return hipGetLastError();
}

hipError_t launch_peak_finder(const double* data, int sizeOfBlock, int nBlocks, int* counts,
                              double* outPeaks, double* outIntervals, double hpf, hipStream_t st)
{
    if (nBlocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_peak_finder, dim3((nBlocks + 255) / 256), dim3(256), 0, st, data,
                       sizeOfBlock, nBlocks, counts, outPeaks, outIntervals, hpf);
    return hipGetLastError();
}

hipError_t launch_dbscan_wave(int nCells, const int* counts, const double* amps,
                              const double* ivls, long long rowStride, int cap, int maxPeaks,
                              double T, int* out, int* overflow, int* nOverflow, int listBase,
                              hipStream_t st)
{
    // launches of at most kWaveCellsPerLaunch cells keep the grid's work-item
    // count (64 per cell) within 32 bits
    for (int c0 = 0; c0 < nCells; c0 += kWaveCellsPerLaunch) {
        const int n = min(nCells - c0, kWaveCellsPerLaunch);
        const long long threads = (long long)n * kWave;
        hipLaunchKernelGGL(k_dbscan_wave, dim3((unsigned)((threads + kDbBlock - 1) / kDbBlock)),
                           dim3(kDbBlock), 0, st, n, counts + c0, amps + c0 * rowStride,
                           ivls + c0 * rowStride, rowStride, cap, maxPeaks, T, out + c0,
                           overflow, nOverflow, listBase + c0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_dbscan_lane(int nList, const int* list, const int* counts, const double* amps,
                              const double* ivls, long long rowStride, double T, int* scratch,
                              long long scratchStride, int* out, const int* outIndex,
                              hipStream_t st)
{
    if (nList <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dbscan_lane, dim3((nList + 63) / 64), dim3(64), 0, st, nList, list,
                       counts, amps, ivls, rowStride, T, scratch, scratchStride, out, outIndex);
    return hipGetLastError();
}

hipError_t launch_dbscan_block(int nList, const int* list, const int* counts, const double* amps,
                               const double* ivls, long long rowStride, double T, int* out,
                               const int* outIndex, hipStream_t st)
{
    // one workgroup of 256 per entry: launches of at most 2^22 entries keep
    // the grid's work-item count within 32 bits
    constexpr int kChunk = 1 << 22;
    for (int c0 = 0; c0 < nList; c0 += kChunk) {
        const int n = min(nList - c0, kChunk);
        const long long r0 = list ? 0 : c0;      // without a list, entry t is cell t
        hipLaunchKernelGGL(k_dbscan_block, dim3(n), dim3(256), 0, st, n,
                           list ? list + c0 : nullptr, counts + r0, amps + r0 * rowStride,
                           ivls + r0 * rowStride, rowStride, T,
                           (list || outIndex) ? out : out + c0,
                           outIndex ? outIndex + c0 : nullptr);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int dbscan_block_max_peaks() { return kBlkMaxP; }

hipError_t launch_kde_wave(int nCells, const int* counts, const double* amps, long long rowStride,
                           int cap, int maxPeaks, int kdeSampling, double I1, double I2, double H,
                           int* out, const int* outIndex, int* overflow, int* nOverflow,
                           int listBase, hipStream_t st)
{
    for (int c0 = 0; c0 < nCells; c0 += kWaveCellsPerLaunch) {
        const int n = min(nCells - c0, kWaveCellsPerLaunch);
        const long long threads = (long long)n * kWave;
        hipLaunchKernelGGL(k_kde_wave, dim3((unsigned)((threads + kKdeBlock - 1) / kKdeBlock)),
                           dim3(kKdeBlock), 0, st,
                           n, counts + c0, amps + c0 * rowStride, rowStride, cap, maxPeaks,
                           kdeSampling, I1, I2, H, outIndex ? out : out + c0,
                           outIndex ? outIndex + c0 : nullptr, overflow, nOverflow,
                           listBase + c0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_pack_rows(long long nCells, int rowCap, int stride, const double* src,
                            double* dst, hipStream_t st)
{
    const long long n = nCells * stride;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nCells,
                       rowCap, stride, src, dst);
    return hipGetLastError();
}

hipError_t launch_clamp(int n, int* v, int clamp, hipStream_t st)
{
    if (n <= 0 || clamp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clamp_i32, dim3((n + 255) / 256), dim3(256), 0, st, n, v, clamp);
    return hipGetLastError();
}

// dst[i] = src[idx[i]] (cell indices of listed entries)
__global__ void k_gather_i32(int n, const int* __restrict__ idx, const int* __restrict__ src,
                             int* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

hipError_t launch_gather(int n, const int* idx, const int* src, int* dst, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, idx,
                       src, dst);
    return hipGetLastError();
}

hipError_t launch_fill(long long n, int* v, int value, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, v,
                       value);
    return hipGetLastError();
}

}  // namespace bifhip
