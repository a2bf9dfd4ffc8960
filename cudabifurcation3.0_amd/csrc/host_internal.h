// host_internal.h -- host-side state shared by the C ABI translation units
// (host_api.cpp, host_tiles.cpp, host_mlp.cpp).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "bifhip.h"
#include "bifhip_dev.h"

namespace bifhip {
namespace host {

// Per-process library context (not thread-safe, like the reference).
struct Ctx {
    std::string outPath = "mat.csv";
    int model = BIFHIP_MODEL_ROSSLER;
    int precision = BIFHIP_PREC_FP64;
    int pert = BIFHIP_PERT_REFERENCE;
    long long dirChunk = -1;
    long long maxCellsPerLaunch = 0;
    int peakCap = 256;
    int clusterClamp = 0;
    int verbose = 0;
    int timing = 0;
    int useGraph = 1;          // replay tile launches and training epochs from HIP graphs
    int sweepIC = 0;           // swept entries index the initial conditions (IC drivers)
    int partialResolve = 1;    // surrogate: class-only cells re-solve the peak count alone
    double chaosConfidence = 0.975;  // surrogate: chaos confidence below this re-solves the LLE
    double lle32Margin = 0.02;  // surrogate: FP32 regime re-solves this far from the threshold
    int reading = BIFHIP_READING_CLASS_AND_CHAOS;  // surrogate: what a cell's regime is
    double quadConfidence = 0.0;  // surrogate: coarse-quad chaos rule (0: off)
    int tilePrecision = BIFHIP_PREC_FP32;  // uneven tiles: FP32 state (FP64: the reference)
    std::string lastError;
    hipStream_t stream = nullptr;
};
extern Ctx g;

// Record an error message; returns code.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::bifhip::host::fail((int)e_, "%s: %s (%s:%d)", #expr,                    \
                                        hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define RC_TRY(expr)                       \
    do {                                   \
        int rc_ = (expr);                  \
        if (rc_ != BIFHIP_OK) return rc_;  \
    } while (0)

// HIP events around a launch family when timing is on (bifhip_kernel_time);
// `cells` is the work the launch covers, summed per family
// (bifhip_kernel_cells).
void add_pending_timer(const char* name, hipEvent_t a, hipEvent_t b, double cells);
void resolve_timers();

struct Timed {
    const char* name;
    hipStream_t st;
    double cells;
    hipEvent_t a = nullptr, b = nullptr;
    explicit Timed(const char* n, hipStream_t s = nullptr, double c = 0)
        : name(n), st(s ? s : g.stream), cells(c)
    {
        if (!g.timing) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        (void)hipEventRecord(a, st);
    }
    ~Timed()
    {
        if (!a) return;
        (void)hipEventRecord(b, st);
        add_pending_timer(name, a, b, cells);
    }
};

// Device buffer that frees itself.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n)
    {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = n;
        if (n == 0) return hipSuccess;
        return hipMalloc(&p, n);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Named device workspaces reused across calls (freed by bifhip_release).
hipError_t pool_get_raw(const char* key, size_t bytes, void** out);

template <class T>
int pool_get(const char* key, size_t count, T** out)
{
    void* p = nullptr;
    HIP_TRY(pool_get_raw(key, count * sizeof(T), &p));
    *out = static_cast<T*>(p);
    return BIFHIP_OK;
}

int ensure_device();
// Frees the surrogate training's pinned staging buffers (bifhip_release).
int release_training_staging();
int make_grid(Grid& G, int nPts, int dimension, const double* ranges, const int* mut,
              const double* ic, int nIC, const double* values, int nVal);
int finish();

}  // namespace host
}  // namespace bifhip
