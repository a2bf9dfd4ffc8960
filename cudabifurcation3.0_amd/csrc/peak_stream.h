// peak_stream.h -- streaming restatement of peakFinder for gfx950 kernels.
#pragma once

#include "bifhip_dev.h"

namespace bifhip {

// ------------------------------------------------------------ peak stream
//
// Streaming form of peakFinder (Library/cudaLibrary.cu:220-266).  The batch
// scan evaluates candidate i (d[i] > d[i-1] && d[i] >= d[i+1]) when d[i+1]
// is known; on a plateau it compares d[j] with d[j+1]; after a peak or a
// rising plateau exit at j the next candidate is j+2.  Per lane the machine
// is in one of three states:
//   seek      evaluate candidate t-1 on sample t
//   plateau   compare d[t-1] with d[t]
//   skip      candidate t-1 is skipped (the loop's ++i after i = j+1)
// The first peak found only provides the time origin of the first interval
// (peakFinder drops it); intervals are (j_k - j_{k-1}) * hpf.
//
// The plateau and skip states are wavefront-uniform lane masks (SGPRs), so
// a sample costs one FP64 compare (p1 >= v) and a few scalar ops; the
// per-lane part runs only on a sample where some lane of the wavefront has
// a candidate or sits on a plateau.
//
// TRAJ: samples come from an integrated trajectory of a model whose every
// component update contains the component itself (Rossler, the linear test
// model), so a NaN sample is never followed by a non-NaN one.  Then
// d[i] > d[i-1] equals !(d[i-1] >= d[i]) -- the previous sample's compare --
// for every pair that can occur (with d[i-1] NaN, d[i] is NaN and the
// candidate fails on d[i] >= d[i+1] in both forms).  Two consequences:
//   * a peak at j has d[j] > d[j+1], so the rising test of candidate j+1
//     fails by itself: only a rising plateau exit needs the skip mask;
//   * push_batch() tests K samples with K compares and one branch, and runs
//     the per-sample machine only for a batch holding a candidate or a
//     pending plateau / skip.
// Without TRAJ (series from memory, general models) the rising test is an
// explicit compare.
//
// T: the sample type.  FP32 trajectories stream float samples: widening a
// float to double is exact and keeps the order (NaN included), so every
// comparison gives what it gives on the widened samples, and only the peaks
// written to the rows are widened.

// Where a lane's peaks go.  With a stage (per-lane LDS buffers of stageN
// amplitudes and stageN sample gaps, stageN a power of two), peaks collect
// there and reach the row stageN at a time, as back-to-back stores into one
// stretch of the row: a lane's peaks arrive hundreds of steps apart, and
// single 8-byte stores into rows 2 KiB apart leave partly written sectors to
// be evicted from L2 (about 3.8x the bytes).  An interval is staged as its
// integer sample gap and written as (double)gap * hpf, the operation
// peakFinder's interval is, so the stage holds 12 bytes per peak.
struct PeakSink {
    double* amps;     // row base (nullptr: count only)
    double* ivls;     // row base (nullptr: no intervals)
    int cap;          // slots in the row
    double hpf;       // peakFinder's h (0.01 by default)
    double* stage = nullptr;   // LDS amplitudes of this lane (nullptr: direct stores)
    int* stageGap = nullptr;   // LDS sample gaps of this lane
    int stageN = 1;

    __device__ __forceinline__ void put(int slot, double amp, int gap) const
    {
        if (!stage) {
            if (amps) amps[slot] = amp;
            if (ivls) ivls[slot] = (double)gap * hpf;
            return;
        }
        const int k = slot & (stageN - 1);
        stage[k] = amp;
        stageGap[k] = gap;
        if (k == stageN - 1) flush(slot - k, stageN);
    }
    __device__ __forceinline__ void flush(int base, int n) const
    {
        for (int k = 0; k < n; ++k) {
            if (amps) amps[base + k] = stage[k];
            if (ivls) ivls[base + k] = (double)stageGap[k] * hpf;
        }
    }
};

template <bool TRAJ, class T = double>
struct PeakStreamT {
    T p1, p2;         // d[t-1], d[t-2] (p2 unused with TRAJ)
    int found, lastj;
    // wavefront-uniform lane masks
    uint64_t plat;    // lanes on a plateau
    uint64_t skip;    // lanes whose next candidate is skipped
    uint64_t geprev;  // TRAJ: lanes with d[t-2] >= d[t-1]

    __device__ __forceinline__ void init()
    {
        p1 = p2 = 0;
        found = 0;
        lastj = 0;
        plat = skip = geprev = 0;
    }

    // samples 0 and 1 fill the window
    __device__ __forceinline__ void prime(T v, int i)
    {
        if (TRAJ && i == 1) geprev = __builtin_amdgcn_ballot_w64(p1 >= v);
        p2 = p1;
        p1 = v;
    }

    __device__ __forceinline__ void emit(T amp, int j, const PeakSink& s)
    {
        if (found >= 1) {
            const int slot = found - 1;
            if (slot < s.cap) s.put(slot, (double)amp, j - lastj);
        }
        lastj = j;
        ++found;
    }

    // the stage's last, partly filled group (full groups were written when
    // they filled; slots from the row's capacity on are never staged)
    __device__ __forceinline__ void finish(const PeakSink& s) const
    {
        if (!s.stage) return;
        const int n = min(found > 1 ? found - 1 : 0, s.cap);
        const int k = n & (s.stageN - 1);
        if (k) s.flush(n - k, k);
    }

    // One sample of the machine: d[t-1] = prev, d[t] = v, rising candidates
    // (before the skip / plateau masks) in c.
    __device__ __forceinline__ void machine(T prev, T v, uint64_t c, int t,
                                            const PeakSink& s)
    {
        const uint64_t cand = c & ~skip & ~plat;
        uint64_t nskip = 0;
        if ((cand | plat) != 0) {
            const uint64_t gt = __builtin_amdgcn_ballot_w64(prev > v);
            const uint64_t lt = __builtin_amdgcn_ballot_w64(prev < v);
            const uint64_t em = (cand | plat) & gt;      // strict fall: a peak at t-1
            nskip = plat & lt;                           // rising plateau exit
            if (!TRAJ) nskip |= em;
            plat = (cand & ~gt) | (plat & ~gt & ~lt);    // equal: (still) on a plateau
            if ((em >> __lane_id()) & 1) emit(prev, t - 1, s);
        }
        skip = nskip;
    }

    // sample v has index t >= 2
    __device__ __forceinline__ void push(T v, int t, const PeakSink& s)
    {
        const uint64_t ge = __builtin_amdgcn_ballot_w64(p1 >= v);
        const uint64_t rise = TRAJ ? ~geprev : __builtin_amdgcn_ballot_w64(p1 > p2);
        // lanes that left a divergent loop hold no pending state
        plat &= __builtin_amdgcn_read_exec();
        const uint64_t c = ge & rise;
        if ((c | plat | skip) != 0) machine(p1, v, c, t, s);
        if (TRAJ) geprev = ge;
        else p2 = p1;
        p1 = v;
    }

    // K consecutive samples v[0..K-1] with indices t..t+K-1 (t >= 2), all
    // lanes of the wavefront active (TRAJ only).
    template <int K>
    __device__ __forceinline__ void push_batch(const T (&v)[K], int t, const PeakSink& s)
    {
        static_assert(TRAJ, "batched samples need the trajectory form");
        uint64_t ge[K], c[K];
        ge[0] = __builtin_amdgcn_ballot_w64(p1 >= v[0]);
#pragma unroll
        for (int k = 1; k < K; ++k) ge[k] = __builtin_amdgcn_ballot_w64(v[k - 1] >= v[k]);
        c[0] = ge[0] & ~geprev;
        uint64_t any = c[0];
#pragma unroll
        for (int k = 1; k < K; ++k) {
            c[k] = ge[k] & ~ge[k - 1];
            any |= c[k];
        }
        if ((any | plat | skip) != 0) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if ((c[k] | plat | skip) != 0) machine(k == 0 ? p1 : v[k - 1], v[k], c[k], t + k, s);
        }
        geprev = ge[K - 1];
        p1 = v[K - 1];
    }

    __device__ __forceinline__ int count() const { return found > 1 ? found - 1 : 0; }
};

using PeakStream = PeakStreamT<false>;
using TrajPeakStream = PeakStreamT<true>;
using TrajPeakStream32 = PeakStreamT<true, float>;

}  // namespace bifhip
