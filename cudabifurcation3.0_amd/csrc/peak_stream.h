// peak_stream.h -- streaming restatement of peakFinder for gfx950 kernels.
#pragma once

#include "bifhip_dev.h"

namespace bifhip {

// ------------------------------------------------------------ peak stream
//
// Streaming form of peakFinder.  The batch scan evaluates candidate i when
// d[i+1] is known; a plateau scan compares d[j] with d[j+1]; after a peak
// or a rising plateau exit at j the next candidate is j+2.  Modes:
//   kSeek  evaluate candidate t-1 on sample t
//   kPlat  inside a plateau: compare d[t-1] with d[t]
//   kSkip  candidate t-1 is skipped (the loop's ++i after i = j+1)
// The first peak found only provides the time origin of the first interval
// (peakFinder drops it); intervals are (j_k - j_{k-1}) * hpf.
enum : int { kSeek = 0, kPlat = 1, kSkip = 2 };

struct PeakSink {
    double* amps;     // row base (nullptr: count only)
    double* ivls;     // row base (nullptr: no intervals)
    int cap;          // slots in the row
    double hpf;       // peakFinder's h (0.01 by default)
};

struct PeakStream {
    double p1, p2;
    int mode, found, lastj;
    // Wavefront-uniform: bit l set when lane l's machine is not in kSeek.
    // Kept as a scalar mask so the per-sample test needs only the two
    // candidate compares; refreshed after every machine run.
    uint64_t pend;

    __device__ __forceinline__ void init()
    {
        p1 = p2 = 0;
        mode = kSeek;
        found = 0;
        lastj = 0;
        pend = 0;
    }

    __device__ __forceinline__ void emit(double amp, int j, const PeakSink& s)
    {
        if (found >= 1) {
            const int slot = found - 1;
            if (slot < s.cap) {
                if (s.amps) s.amps[slot] = amp;
                if (s.ivls) s.ivls[slot] = (double)(j - lastj) * s.hpf;
            }
        }
        lastj = j;
        ++found;
    }

    // sample v has index t (t >= 2 is checked by the caller's loop shape).
    // A lane in kSeek without a candidate (the common case: a peak comes
    // every few hundred samples) only shifts its window, so the state
    // machine runs only when some lane of the wavefront needs it.
    __device__ __forceinline__ void push(double v, int t, const PeakSink& s)
    {
        // each compare feeds a ballot directly as a lane mask
        const uint64_t cand = __builtin_amdgcn_ballot_w64(p1 > p2) &
                              __builtin_amdgcn_ballot_w64(p1 >= v);
        if ((cand | pend) != 0) {
            machine(v, t, s);
            // lanes not active here (left a divergent loop) keep their bits
            const bool busy = mode != kSeek;
            pend = (pend & ~__builtin_amdgcn_ballot_w64(true)) | __builtin_amdgcn_ballot_w64(busy);
        }
        p2 = p1;
        p1 = v;
    }

    __device__ __forceinline__ void machine(double v, int t, const PeakSink& s)
    {
        if (mode == kSeek) {
            if (p1 > p2 && p1 >= v) {
                if (p1 > v) { emit(p1, t - 1, s); mode = kSkip; }
                else mode = kPlat;
            }
        } else if (mode == kPlat) {
            if (p1 < v) mode = kSkip;
            else if (p1 > v) { emit(p1, t - 1, s); mode = kSkip; }
        } else {
            mode = kSeek;
        }
    }

    __device__ __forceinline__ int count() const { return found > 1 ? found - 1 : 0; }
};

}  // namespace bifhip
