// host_api.cpp -- C ABI of libbifhip.so: context, drivers, CSV writers.
//
// The drivers keep the reference's host-function surface
// (Library/hostLibrary.cu) -- argument meaning, derived step counts and
// output files -- and drive the gfx950 kernels in k_bif.hip / k_lyap.hip.
// Unlike the reference they never stage a full trajectory in HBM: the
// integrator streams samples straight into the peak finder, so memory per
// cell is a small peak row and a whole map usually fits in one launch.
#include "bifhip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <system_error>
#include <thread>
#include <numeric>
#include <vector>

#include "bifhip_dev.h"
#include "host_internal.h"

// kernel launchers (k_bif.hip, k_lyap.hip) -------------------------------
namespace bifhip {
hipError_t launch_bif(int model, int precision, const BifArgs& A, int wv, bool check, int* counts,
                      double* amps, double* ivls, hipStream_t st);
hipError_t launch_series(int model, const BifArgs& A, int wv, double* data, int* flags,
                         hipStream_t st);
hipError_t launch_pack_rows(long long nCells, int rowCap, int stride, const double* src,
                            double* dst, hipStream_t st);
hipError_t launch_peak_finder(const double* data, int sizeOfBlock, int nBlocks, int* counts,
                              double* outPeaks, double* outIntervals, double hpf, hipStream_t st);
hipError_t launch_dbscan_wave(int nCells, const int* counts, const double* amps,
                              const double* ivls, long long rowStride, int cap, int maxPeaks,
                              double T, int* out, int* overflow, int* nOverflow, int listBase,
                              hipStream_t st);
hipError_t launch_dbscan_lane(int nList, const int* list, const int* counts, const double* amps,
                              const double* ivls, long long rowStride, double T, int* scratch,
                              long long scratchStride, int* out, const int* outIndex,
                              hipStream_t st);
hipError_t launch_dbscan_block(int nList, const int* list, const int* counts, const double* amps,
                               const double* ivls, long long rowStride, double T, int* out,
                               const int* outIndex, hipStream_t st);
int dbscan_block_max_peaks();
const char* bif_kernel_family(int model, int precision, long long nCells, bool check);
hipError_t launch_gen_bif(int model, const BifArgs& A, int wv, bool check, int* counts,
                          double* amps, double* ivls, double* last, hipStream_t st);
hipError_t launch_osc_flags(int n, const int* counts, const double* amps, long long rowStride,
                            const double* last, double thr, unsigned char* flags, hipStream_t st);
hipError_t launch_osc_combine(int n, const int* clusters, const unsigned char* flags,
                              double* out, hipStream_t st);
hipError_t launch_denis(int model, const Grid& g, int T, double h, double hS, int iters, int wv,
                        long long nPoints, double* data, hipStream_t st);
hipError_t launch_kde_wave(int nCells, const int* counts, const double* amps, long long rowStride,
                           int cap, int maxPeaks, int kdeSampling, double I1, double I2, double H,
                           int* out, const int* outIndex, int* overflow, int* nOverflow,
                           int listBase, hipStream_t st);
hipError_t launch_lle_peaks(int model, int wv, const FusedArgs& F, double* lle, int* counts,
                            double* amps, double* ivls, hipStream_t st);
hipError_t launch_clamp(int n, int* v, int clamp, hipStream_t st);
hipError_t launch_fill(long long n, int* v, int value, hipStream_t st);
hipError_t launch_gather(int n, const int* idx, const int* src, int* dst, hipStream_t st);
hipError_t launch_lle(int model, int precision, const LyapArgs& A, bool check, double* out,
                      hipStream_t st);
hipError_t launch_ls(int model, const LyapArgs& A, bool check, double* out, hipStream_t st);
hipError_t launch_gs(const double* a, double* b, double* den, int n, int count, hipStream_t st);
}  // namespace bifhip

// ------------------------------------------------------------- context

namespace bifhip {
namespace host {

Ctx g;

// per-kernel-family timing: pending event pairs resolved after each call
struct TimerSlot {
    double ms = 0;
    long long launches = 0;
    double cells = 0;
};
std::map<std::string, TimerSlot> g_times;
struct PendingEv {
    std::string name;
    hipEvent_t a, b;
    double cells;
};
std::vector<PendingEv> g_pending;

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g.lastError = buf;
    return code;
}

void add_pending_timer(const char* name, hipEvent_t a, hipEvent_t b, double cells)
{
    g_pending.push_back({name, a, b, cells});
}

void resolve_timers()
{
    for (auto& p : g_pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& s = g_times[p.name];
            s.ms += ms;
            s.launches += 1;
            s.cells += p.cells;
        }
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    g_pending.clear();
}

// cached workspaces reused across calls (released by bifhip_release)
struct Pool {
    std::map<std::string, void*> ptr;
    std::map<std::string, size_t> size;
    hipError_t get(const char* key, size_t bytes, void** out)
    {
        auto it = ptr.find(key);
        if (it != ptr.end() && size[key] >= bytes) {
            *out = it->second;
            return hipSuccess;
        }
        if (it != ptr.end()) {
            (void)hipFree(it->second);
            ptr.erase(it);
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
        if (e != hipSuccess) return e;
        ptr[key] = p;
        size[key] = bytes ? bytes : 16;
        *out = p;
        return hipSuccess;
    }
    void release()
    {
        for (auto& kv : ptr) (void)hipFree(kv.second);
        ptr.clear();
        size.clear();
    }
};
Pool g_pool;

hipError_t pool_get_raw(const char* key, size_t bytes, void** out)
{
    return g_pool.get(key, bytes, out);
}

int ensure_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(BIFHIP_E_NODEV, "no HIP device visible");
    return BIFHIP_OK;
}

// getValueByIdx on the host (Library/cudaLibrary.cu:210-216)
double value_by_idx(int idx, int nPts, double s, double f, int k)
{
    float p = powf((float)nPts, (float)k);
    float q = (float)idx / p;
    int t = (int)q;
    return s + ((double)(t % nPts) * ((double)(f - s) / (double)(nPts - 1)));
}

int model_dim(int model) { return model == BIFHIP_MODEL_MEMRISTIVE ? 5 : 3; }

int make_grid(Grid& G, int nPts, int dimension, const double* ranges, const int* mut,
              const double* ic, int nIC, const double* values, int nVal)
{
    std::memset(&G, 0, sizeof(G));
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1 (got %d)", nPts);
    if (dimension < 1 || dimension > 2) return fail(BIFHIP_E_ARG, "dimension must be 1 or 2");
    if (!ranges || !mut || !ic || !values) return fail(BIFHIP_E_ARG, "null input array");
    if (nIC != model_dim(g.model))
        return fail(BIFHIP_E_ARG, "model %d has %d state variables, got %d initial conditions",
                    g.model, model_dim(g.model), nIC);
    const int minVal = g.model == BIFHIP_MODEL_MEMRISTIVE ? 8 : 3;
    if (nVal < minVal || nVal > kMaxVal)
        return fail(BIFHIP_E_ARG, "amountOfValues must be in [%d, %d] for model %d (got %d)",
                    minVal, kMaxVal, g.model, nVal);
    const long long cells = dimension == 1 ? (long long)nPts : (long long)nPts * nPts;
    if (cells > (1LL << 30)) return fail(BIFHIP_E_LIMIT, "grid of %lld cells is too large", cells);
    G.nPts = nPts;
    G.dimension = dimension;
    G.nIC = nIC;
    G.nVal = nVal;
    for (int i = 0; i < nIC; ++i) G.ic[i] = ic[i];
    for (int i = 0; i < nVal; ++i) G.values[i] = values[i];
    G.icSweep = g.sweepIC;
    const int nMut = g.sweepIC ? nIC : nVal;   // what indicesOfMutVars indexes
    for (int k = 0; k < dimension; ++k) {
        if (mut[k] < 0 || mut[k] >= nMut)
            return fail(BIFHIP_E_ARG, "indicesOfMutVars[%d] = %d out of range", k, mut[k]);
        G.mut[k] = mut[k];
        G.pk[k] = powf((float)nPts, (float)k);
        G.start[k] = ranges[2 * k];
        G.step[k] = (double)(ranges[2 * k + 1] - ranges[2 * k]) / (double)(nPts - 1);
    }
    return BIFHIP_OK;
}

long long cells_of(int nPts, int dimension)
{
    return dimension == 1 ? (long long)nPts : (long long)nPts * nPts;
}

// Least double T with (float)sqrt(T) >= eps: the exact DBSCAN neighbour
// threshold (hypotf(a,b) < eps  <=>  (double)a*a + (double)b*b < T).
double dbscan_threshold(double eps)
{
    if (!(eps > 0)) return 0.0;  // never a neighbour (also NaN eps)
    if (std::isinf(eps)) return INFINITY;
    auto ok = [&](double s) { return (double)(float)std::sqrt(s) >= eps; };
    // search over the ordered bit patterns of non-negative doubles
    uint64_t lo = 0, hi;
    double inf = INFINITY;
    std::memcpy(&hi, &inf, sizeof(hi));
    // invariant: ok(hi) true (sqrt(inf) = inf >= eps), lo may be false
    if (ok(0.0)) return 0.0;
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        double m;
        std::memcpy(&m, &mid, sizeof(m));
        if (ok(m)) hi = mid; else lo = mid;
    }
    double T;
    std::memcpy(&T, &hi, sizeof(T));
    return T;
}

long long launch_chunk(long long cells)
{
    long long c = g.maxCellsPerLaunch > 0 ? g.maxCellsPerLaunch : cells;
    if (c > (1LL << 26)) c = 1LL << 26;
    return c < 1 ? 1 : c;
}

// ------------------------------------------------------- peak-map core

struct PeakMapSpec {
    Grid G;
    int dimension;
    long long cells;
    double h, maxValue;
    int skip, N, pre, wv;
    int precision;    // BIFHIP_PREC_FP64, or FP32 for the peak maps that honour it
};

int peak_spec(PeakMapSpec& S, int dimension, double tMax, int nPts, double h, int nIC,
              const double* ic, const double* ranges, const int* mut, int wv, double maxValue,
              double transientTime, const double* values, int nVal, int pre)
{
    RC_TRY(make_grid(S.G, nPts, dimension, ranges, mut, ic, nIC, values, nVal));
    if (wv < 0 || wv >= nIC) return fail(BIFHIP_E_ARG, "writableVar %d out of range", wv);
    if (pre == 0) return fail(BIFHIP_E_ARG, "preScaller must not be 0 (tMax/h/preScaller)");
    // derived counts exactly as Library/hostLibrary.cu:22-23, :417-418, :588-589
    const int N = (int)(tMax / h / pre);
    const int skip = (int)(transientTime / h);
    if (N < 0 || skip < 0) return fail(BIFHIP_E_ARG, "negative step counts");
    S.dimension = dimension;
    S.cells = cells_of(nPts, dimension);
    S.h = h;
    S.maxValue = maxValue;
    S.skip = skip;
    S.N = N;
    S.pre = pre < 1 ? 1 : pre;
    S.wv = wv;
    S.precision = BIFHIP_PREC_FP64;
    return BIFHIP_OK;
}

// The 1D diagram and the 2D cluster maps follow bifhip_set_precision; the
// surrogate's re-solves always integrate in FP64.
int peak_precision(PeakMapSpec& S)
{
    if (g.precision != BIFHIP_PREC_FP32) return BIFHIP_OK;
    if (g.model == BIFHIP_MODEL_MEMRISTIVE)
        return fail(BIFHIP_E_ARG, "FP32 peak maps support the Rossler and linear models");
    S.precision = BIFHIP_PREC_FP32;
    return BIFHIP_OK;
}

BifArgs bif_args(const PeakMapSpec& S, int cellBase, const int* list, int n, int cap)
{
    BifArgs A;
    A.g = S.G;
    A.h = S.h;
    A.maxValue = S.maxValue;
    A.hpf = 0.01;  // peakFinder's default h (Library/cudaLibrary.cuh:204)
    A.cellBase = cellBase;
    A.cellList = list;
    A.nCells = n;
    A.skip = S.skip;
    A.N = S.N;
    A.pre = S.pre;
    A.cap = cap;
    A.rowStride = cap;
    return A;
}

// Integrate cells [base, base+n) (or the listed ones) into peak rows.
int integrate_peaks(const PeakMapSpec& S, int base, const int* d_list, int n, int cap,
                    int* d_counts, double* d_amps, double* d_ivls)
{
    const long long chunk = launch_chunk(n);
    for (long long off = 0; off < n; off += chunk) {
        const int m = (int)std::min<long long>(chunk, n - off);
        BifArgs A = bif_args(S, base + (int)off, d_list ? d_list + off : nullptr, m, cap);
        Timed t(bif_kernel_family(g.model, S.precision, m, S.maxValue != 0), nullptr, m);
        HIP_TRY(launch_bif(g.model, S.precision, A, S.wv, S.maxValue != 0, d_counts + off,
                           d_amps ? d_amps + off * cap : nullptr,
                           d_ivls ? d_ivls + off * cap : nullptr, g.stream));
    }
    return BIFHIP_OK;
}

// ------------------------------------------ chunked integrate / cluster
//
// Peak maps run in passes of at most `pass` cells whose peak rows fit in
// ~8 GiB.  Two row sets alternate: while pass k+1 integrates on the library
// stream, pass k is clustered on a side stream.  The integrators are bound
// by the vector ALUs and the clustering by the scalar unit, so the two
// share the CUs well.  Events order the work: clustering k waits for
// integration k, integration k+2 waits for clustering k (the same rows),
// and the library stream waits for the last clustering before the caller
// sees the results.
struct PeakRows {
    int* counts;
    double* amps;
    double* ivls;
};

namespace {
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t integrated[2] = {nullptr, nullptr};
    hipEvent_t clustered[2] = {nullptr, nullptr};
};
SideStream g_side;
}  // namespace

static int side_stream()
{
    if (g_side.s) return BIFHIP_OK;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipEventCreateWithFlags(&g_side.integrated[b], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&g_side.clustered[b], hipEventDisableTiming));
    }
    // non-blocking: the library stream is the null stream
    HIP_TRY(hipStreamCreateWithFlags(&g_side.s, hipStreamNonBlocking));
    return BIFHIP_OK;
}

// Two row sets of `pass` cells x `cap` slots (pool keys <tag>_<what><b>);
// no interval rows when ivls is false.
static int peak_rows(const char* tag, long long pass, int cap, bool ivls, PeakRows rows[2])
{
    char key[64];
    for (int b = 0; b < 2; ++b) {
        std::snprintf(key, sizeof(key), "%s_counts%d", tag, b);
        RC_TRY(pool_get(key, (size_t)pass, &rows[b].counts));
        std::snprintf(key, sizeof(key), "%s_amps%d", tag, b);
        RC_TRY(pool_get(key, (size_t)(pass * cap), &rows[b].amps));
        rows[b].ivls = nullptr;
        if (ivls) {
            std::snprintf(key, sizeof(key), "%s_ivls%d", tag, b);
            RC_TRY(pool_get(key, (size_t)(pass * cap), &rows[b].ivls));
        }
    }
    return BIFHIP_OK;
}

// integrate(off, m, rows) enqueues on g.stream; cluster(off, m, rows, st)
// enqueues on st.  join false leaves the last clustering running on the side
// stream (the caller synchronises it before reading the results).  tail > 0
// makes the last pass `tail` cells (when more than that remain), so that
// only a small clustering pass is left with no integration beside it.
template <class Integrate, class Cluster>
static int chunk_pipeline(long long n, long long pass, const PeakRows rows[2],
                          Integrate integrate, Cluster cluster, bool join = true,
                          long long tail = 0)
{
    RC_TRY(side_stream());
    int k = 0;
    for (long long off = 0, m = 0; off < n; off += m, ++k) {
        const long long r = n - off;
        m = std::min(pass, r);
        if (tail > 0 && r > tail && r <= pass) m = r - tail;
        const int b = k & 1;
        if (k >= 2) HIP_TRY(hipStreamWaitEvent(g.stream, g_side.clustered[b], 0));
        RC_TRY(integrate(off, (int)m, rows[b]));
        HIP_TRY(hipEventRecord(g_side.integrated[b], g.stream));
        HIP_TRY(hipStreamWaitEvent(g_side.s, g_side.integrated[b], 0));
        RC_TRY(cluster(off, (int)m, rows[b], g_side.s));
        HIP_TRY(hipEventRecord(g_side.clustered[b], g_side.s));
    }
    if (join && k > 0) HIP_TRY(hipStreamWaitEvent(g.stream, g_side.clustered[(k - 1) & 1], 0));
    return BIFHIP_OK;
}

// A peak-count job split in two so independent work can be enqueued while
// its last clustering pass still runs on the side stream: begin enqueues the
// passes; finish waits for the side stream, reads the overflow count and
// recomputes the overflowing cells.  One job at a time (the rows and
// overflow buffers are shared).
struct PeakCellsJob {
    PeakMapSpec S;
    int maxPeaks;
    double T;
    const int* h_list;
    const int* d_list;
    int base, n;
    int* d_peak;
    int* d_ovf;
    int* d_novf;
    bool pending = false;
};

static PeakCellsJob g_peakJob;   // the surrogate's pending peak-count job

// Entry points that use the shared rows first drop a job left pending by a
// call that failed between its two halves: its clustering is waited for and
// its results are not processed.
static int abandon_pending_peaks()
{
    if (!g_peakJob.pending) return BIFHIP_OK;
    g_peakJob.pending = false;
    HIP_TRY(hipStreamSynchronize(g_side.s));
    return BIFHIP_OK;
}

// Whole 2D/1D cluster map (DBSCAN when kind==0, KDE when kind==1) into d_out.
struct ClusterParams {
    int kind;
    int maxPeaks;
    double eps;
    int kdeSampling;
    double I1, I2, H;
};

// DBSCAN of cells whose rows hold every peak: entry t reads row
// r = list ? list[t] : t of counts/amps/ivls (rowStride slots) and writes
// out[outIndex ? outIndex[t] : r].  Cells with up to dbscan_block_max_peaks()
// peaks take the workgroup kernel, larger ones the lane kernel.
int dbscan_full_rows(int n, const int* d_list, int nRows, const int* d_counts,
                     const double* d_amps, const double* d_ivls, long long rowStride, double T,
                     int* d_out, const int* d_outIndex)
{
    if (n <= 0) return BIFHIP_OK;
    std::vector<int> cnt(nRows), lst(d_list ? n : 0), oi(d_outIndex ? n : 0);
    HIP_TRY(hipMemcpy(cnt.data(), d_counts, sizeof(int) * nRows, hipMemcpyDeviceToHost));
    if (d_list) HIP_TRY(hipMemcpy(lst.data(), d_list, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (d_outIndex)
        HIP_TRY(hipMemcpy(oi.data(), d_outIndex, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int> bigRow, bigOut;
    for (int t = 0; t < n; ++t) {
        const int r = d_list ? lst[t] : t;
        if (cnt[r] > dbscan_block_max_peaks()) {
            bigRow.push_back(r);
            bigOut.push_back(d_outIndex ? oi[t] : r);
        }
    }
    {
        Timed tm("dbscan", nullptr, n);
        HIP_TRY(launch_dbscan_block(n, d_list, d_counts, d_amps, d_ivls, rowStride, T, d_out,
                                    d_outIndex, g.stream));
    }
    if (!bigRow.empty()) {
        const int nb = (int)bigRow.size();
        DBuf drow, dout, scr;
        HIP_TRY(drow.alloc(sizeof(int) * nb));
        HIP_TRY(dout.alloc(sizeof(int) * nb));
        HIP_TRY(hipMemcpy(drow.p, bigRow.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dout.p, bigOut.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        HIP_TRY(scr.alloc(sizeof(int) * (size_t)nb * 2 * rowStride));
        Timed tm("dbscan", nullptr, nb);
        HIP_TRY(launch_dbscan_lane(nb, drow.as<int>(), d_counts, d_amps, d_ivls, rowStride, T,
                                   scr.as<int>(), 2LL * rowStride, d_out, dout.as<int>(),
                                   g.stream));
    }
    HIP_TRY(hipStreamSynchronize(g.stream));
    return BIFHIP_OK;
}

int cluster_map_device(const PeakMapSpec& S, const ClusterParams& C, int* d_out)
{
    RC_TRY(abandon_pending_peaks());
    const long long cells = S.cells;
    const int capMax = std::max(2, std::min(g.peakCap, S.N / 2 + 1));
    int cap = capMax;
    if ((long long)C.maxPeaks + 1 < cap) cap = (int)std::max(2LL, (long long)C.maxPeaks + 1);
    // cells per pass bounded so the peak rows stay within ~8 GiB
    const long long rowBytes = (long long)cap * 16 + 8;
    long long pass = std::max<long long>(1, (8LL << 30) / rowBytes);
    pass = std::min(pass, cells);
    if (g.maxCellsPerLaunch > 0) pass = std::min(pass, g.maxCellsPerLaunch);

    int *d_ovf, *d_novf;
    PeakRows rows[2];
    RC_TRY(peak_rows("m", pass, cap, C.kind == 0, rows));
    RC_TRY(pool_get("ovf", (size_t)cells, &d_ovf));
    RC_TRY(pool_get("novf", 1, &d_novf));
    HIP_TRY(launch_fill(1, d_novf, 0, g.stream));
    const double T = dbscan_threshold(C.eps);

    RC_TRY(chunk_pipeline(
        cells, pass, rows,
        [&](long long off, int n, const PeakRows& r) {
            return integrate_peaks(S, (int)off, nullptr, n, cap, r.counts, r.amps, r.ivls);
        },
        [&](long long off, int n, const PeakRows& r, hipStream_t st) {
            if (C.kind == 0) {
                Timed t("dbscan", st, n);
                HIP_TRY(launch_dbscan_wave(n, r.counts, r.amps, r.ivls, cap, cap, C.maxPeaks, T,
                                           d_out + off, d_ovf, d_novf, (int)off, st));
            } else {
                Timed t("kde", st);
                HIP_TRY(launch_kde_wave(n, r.counts, r.amps, cap, cap, C.maxPeaks,
                                        C.kdeSampling, C.I1, C.I2, C.H, d_out + off, nullptr,
                                        d_ovf, d_novf, (int)off, st));
            }
            return BIFHIP_OK;
        }));
    int novf = 0;
    HIP_TRY(hipMemcpyAsync(&novf, d_novf, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (novf > 0) {
        // Rare: cells with more peaks than the rows hold (or than the wave
        // kernel's 256 registers slots).  Recompute them with full rows.
        std::vector<int> list(novf);
        HIP_TRY(hipMemcpy(list.data(), d_ovf, sizeof(int) * novf, hipMemcpyDeviceToHost));
        const int cap2 = (int)std::max(2LL, std::min((long long)S.N / 2 + 1, (long long)C.maxPeaks + 1));
        DBuf cnt, am, iv;
        HIP_TRY(cnt.alloc(sizeof(int) * novf));
        HIP_TRY(am.alloc(sizeof(double) * (size_t)novf * cap2));
        HIP_TRY(iv.alloc(sizeof(double) * (size_t)novf * cap2));
        RC_TRY(integrate_peaks(S, 0, d_ovf, novf, cap2, cnt.as<int>(), am.as<double>(),
                               iv.as<double>()));
        if (C.kind == 0) {
            RC_TRY(dbscan_full_rows(novf, nullptr, novf, cnt.as<int>(), am.as<double>(),
                                    iv.as<double>(), cap2, T, d_out, d_ovf));
        } else {
            // rows now hold every peak (cap2 >= P), so this pass cannot overflow
            int *d_o2, *d_n2;
            RC_TRY(pool_get("ovf2", (size_t)novf, &d_o2));
            RC_TRY(pool_get("novf2", 1, &d_n2));
            HIP_TRY(launch_fill(1, d_n2, 0, g.stream));
            Timed t("kde");
            HIP_TRY(launch_kde_wave(novf, cnt.as<int>(), am.as<double>(), cap2, cap2, C.maxPeaks,
                                    C.kdeSampling, C.I1, C.I2, C.H, d_out, d_ovf, d_o2, d_n2, 0,
                                    g.stream));
        }
        HIP_TRY(hipStreamSynchronize(g.stream));
    }
    if (C.kind == 0 && g.clusterClamp > 0) HIP_TRY(launch_clamp((int)cells, d_out, g.clusterClamp, g.stream));
    return BIFHIP_OK;
}

void progress(const char* name, long long cells);

// Ostrovsky periodicity map (calculationOfPeriodicityByOstrovsky,
// Library/hostLibrary.cu:748-925) of a 2D lattice swept over initial
// conditions: rows hold every peak of a cell, so the five flags and the
// DBSCAN count see the reference's full peak list.
int ostrovsky_map_device(const PeakMapSpec& S, int maxPeaks, double eps, double threshold,
                         double* d_out)
{
    const long long cells = S.cells;
    const int cap = std::max(2, S.N / 2 + 1);
    const long long rowBytes = (long long)cap * 16 + 16;
    long long pass = std::max<long long>(1, (8LL << 30) / rowBytes);
    pass = std::min(pass, cells);
    if (g.maxCellsPerLaunch > 0) pass = std::min(pass, g.maxCellsPerLaunch);
    int *d_counts, *d_clusters, *d_ovf, *d_novf;
    double *d_amps, *d_ivls, *d_last;
    unsigned char* d_flags;
    RC_TRY(pool_get("o_counts", (size_t)pass, &d_counts));
    RC_TRY(pool_get("o_clusters", (size_t)pass, &d_clusters));
    RC_TRY(pool_get("o_amps", (size_t)(pass * cap), &d_amps));
    RC_TRY(pool_get("o_ivls", (size_t)(pass * cap), &d_ivls));
    RC_TRY(pool_get("o_last", (size_t)pass, &d_last));
    RC_TRY(pool_get("o_flags", (size_t)pass, &d_flags));
    RC_TRY(pool_get("o_ovf", (size_t)pass, &d_ovf));
    RC_TRY(pool_get("o_novf", 1, &d_novf));
    const double T = dbscan_threshold(eps);
    for (long long off = 0; off < cells; off += pass) {
        const int m = (int)std::min(pass, cells - off);
        BifArgs A = bif_args(S, (int)off, nullptr, m, cap);
        {
            Timed t("bif_peaks");
            HIP_TRY(launch_gen_bif(g.model, A, S.wv, S.maxValue != 0, d_counts, d_amps, d_ivls,
                                   d_last, g.stream));
        }
        HIP_TRY(launch_osc_flags(m, d_counts, d_amps, cap, d_last, threshold, d_flags,
                                 g.stream));
        HIP_TRY(launch_fill(1, d_novf, 0, g.stream));
        {
            Timed t("dbscan");
            HIP_TRY(launch_dbscan_wave(m, d_counts, d_amps, d_ivls, cap, cap, maxPeaks, T,
                                       d_clusters, d_ovf, d_novf, 0, g.stream));
        }
        int novf = 0;
        HIP_TRY(hipMemcpyAsync(&novf, d_novf, sizeof(int), hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        if (novf > 0)   // rows already hold every peak: cluster them in place
            RC_TRY(dbscan_full_rows(novf, d_ovf, m, d_counts, d_amps, d_ivls, cap, T, d_clusters,
                                    nullptr));
        HIP_TRY(launch_osc_combine(m, d_clusters, d_flags, d_out + off, g.stream));
    }
    return BIFHIP_OK;
}

int ostrovsky_impl(double tMax, int nPts, double h, int nIC, const double* ic,
                   const double* ranges, const int* mut, int wv, double maxValue, int maxPeaks,
                   double transientTime, const double* values, int nVal, int pre, double eps,
                   double threshold, double* d_out)
{
    RC_TRY(ensure_device());
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 2, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, transientTime,
                     values, nVal, pre));
    progress("OSTROVSKY", S.cells);
    RC_TRY(ostrovsky_map_device(S, maxPeaks, eps, threshold, d_out));
    return finish();
}

// ------------------------------------------------------- Lyapunov core

struct LyapSpec {
    Grid G;
    int dimension;
    long long cells;
    double h, eps, tMax, maxValue;
    int skip, nNT, nBlocks;
    long long dirChunk;
};

int lyap_spec(LyapSpec& S, int dimension, double tMax, double NT, int nPts, double h,
              double eps, const double* ic, int nIC, const double* ranges, const int* mut,
              double maxValue, double transientTime, const double* values, int nVal)
{
    RC_TRY(make_grid(S.G, nPts, dimension, ranges, mut, ic, nIC, values, nVal));
    // derived counts exactly as Library/hostLibrary.cu:1286-1288 and
    // LLEKernelCUDA's (size_t)(NT/h) (Library/cudaLibrary.cu:697)
    const size_t nNT = (size_t)(NT / h);
    const int nBlocks = (int)(tMax / NT);
    const int skip = (int)(transientTime / h);
    if (nNT > (size_t)INT32_MAX || skip < 0)
        return fail(BIFHIP_E_ARG, "invalid NT/h or transient");
    S.dimension = dimension;
    S.cells = cells_of(nPts, dimension);
    S.h = h;
    S.eps = eps;
    S.tMax = tMax;
    S.maxValue = maxValue;
    S.skip = skip;
    S.nNT = (int)nNT;
    S.nBlocks = nBlocks < 0 ? 0 : nBlocks;
    // reference chunking of the direction index: 200000 cells in LLE2D
    // (Library/hostLibrary.cu:1552), one chunk in the other drivers
    return BIFHIP_OK;
}

static void direction_of(bool ls, long long gi, long long chunk, int D, double* z)
{
    const long long local = chunk > 0 ? gi % chunk : gi;
    if (!ls) {
        const uint32_t id = (uint32_t)local;
        double zp = 0;
        for (int i = 0; i < D; ++i) {
            if (g.sweepIC) {
                // LLEKernelICCUDA, Library/cudaLibrary.cu:846
                const int idx = (int)local;
                const double arg = 0.2171828 * (i + 1) + idx + (0.2171828 + i * idx);
                z[i] = sinf((float)arg) * 0.5;
            } else {
                const int32_t arg = (int32_t)(id * ((uint32_t)i * id + 1u) + 1u);
                z[i] = 0.5 * (double)sinf((float)arg);
            }
            zp += z[i] * z[i];
        }
        zp = std::sqrt(zp);
        for (int i = 0; i < D; ++i) z[i] /= zp;
    } else {
        const int idx = (int)local;
        for (int j = 0; j < D; ++j) {
            double zp = 0;
            for (int i = 0; i < D; ++i) {
                const double arg = 0.2171828 * (i + 1) * (j + 1) + idx + (0.2171828 + i * j * idx);
                z[j * D + i] = sinf((float)arg) * 0.5;
                zp += z[j * D + i] * z[j * D + i];
            }
            zp = std::sqrt(zp);
            for (int i = 0; i < D; ++i) z[j * D + i] /= zp;
        }
    }
}

// Directions of n cells, cell t's flattened index cellIndex(t), into out
// (D or D*D values per cell).  Each cell's values depend only on its own
// index, so the table is filled by up to 16 host threads.
// Runs work(k) for k in [0, parts) on that many host threads; a part whose
// thread cannot be started runs on the calling thread.
template <class W>
static void run_parts(int parts, W work)
{
    if (parts <= 1) {
        work(0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(parts);
    int started = 0;
    for (; started < parts; ++started) {
        try {
            th.emplace_back(work, started);
        } catch (const std::system_error&) {
            break;
        }
    }
    for (int k = started; k < parts; ++k) work(k);
    for (auto& t : th) t.join();
}

// The table lives in pinned host memory that is reused across calls: the
// host threads fill it and one DMA copies it (dir_upload).  dir_table waits
// for the previous copy out of it before handing it out again.
struct DirStaging {
    double* p = nullptr;
    size_t cap = 0;
    hipEvent_t copied = nullptr;
    bool pending = false;
} g_dirStage;

static int dir_table(size_t n, double** out)
{
    DirStaging& D = g_dirStage;
    if (D.pending) {
        HIP_TRY(hipEventSynchronize(D.copied));
        D.pending = false;
    }
    if (!D.copied) HIP_TRY(hipEventCreateWithFlags(&D.copied, hipEventDisableTiming));
    if (n > D.cap) {
        if (D.p) HIP_TRY(hipHostFree(D.p));
        D.p = nullptr;
        D.cap = 0;
        HIP_TRY(hipHostMalloc((void**)&D.p, sizeof(double) * std::max<size_t>(n, 1),
                              hipHostMallocDefault));
        D.cap = n;
    }
    *out = D.p;
    return BIFHIP_OK;
}

static int dir_upload(double* d_dirs, size_t n, hipStream_t st)
{
    DirStaging& D = g_dirStage;
    HIP_TRY(hipMemcpyAsync(d_dirs, D.p, sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(D.copied, st));
    D.pending = true;
    return BIFHIP_OK;
}

int release_dir_staging()
{
    DirStaging& D = g_dirStage;
    if (D.pending) HIP_TRY(hipEventSynchronize(D.copied));
    D.pending = false;
    if (D.p) HIP_TRY(hipHostFree(D.p));
    D.p = nullptr;
    D.cap = 0;
    if (D.copied) HIP_TRY(hipEventDestroy(D.copied));
    D.copied = nullptr;
    return BIFHIP_OK;
}

// Fills the pinned table with the directions of n cells and copies it to
// d_dirs on st.
template <class Idx>
static int directions_of(bool ls, int n, long long chunk, Idx cellIndex, double* d_dirs,
                         hipStream_t st)
{
    const int D = model_dim(g.model);
    const int per = ls ? D * D : D;
    double* o = nullptr;
    RC_TRY(dir_table((size_t)n * per, &o));
    auto fill = [&](int t0, int t1) {
        for (int t = t0; t < t1; ++t) direction_of(ls, cellIndex(t), chunk, D, o + (size_t)t * per);
    };
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int T = std::max(1, std::min({16, hw, n / 8192}));
    run_parts(T, [&](int k) {
        fill((int)((long long)n * k / T), (int)((long long)n * (k + 1) / T));
    });
    return dir_upload(d_dirs, (size_t)n * per, st);
}

// Host table of reference perturbation directions for cells [off, off+n):
// LLE: 0.5*sinf((int)(idx*(i*idx+1)+1)) normalised (Library/cudaLibrary.cu:720-732);
//      sweeps over initial conditions: 0.5*sinf(0.2171828(i+1) + idx +
//      (0.2171828 + i*idx)) normalised (Library/cudaLibrary.cu:844-855)
// LS:  0.5*sinf(0.2171828(i+1)(j+1) + idx + (0.2171828 + i*j*idx)) rows
//      normalised (Library/cudaLibrary.cu:1025-1040)
int directions(bool ls, long long off, int n, long long chunk, double* d_dirs, hipStream_t st)
{
    return directions_of(ls, n, chunk, [off](int t) { return off + t; }, d_dirs, st);
}

int lyap_map_device(const LyapSpec& S, bool ls, long long defaultChunk, double* d_out)
{
    const long long cells = S.cells;
    const long long chunk = launch_chunk(cells);
    const long long dirChunk = g.dirChunk >= 0 ? g.dirChunk : defaultChunk;
    const bool refDirs = g.pert == BIFHIP_PERT_REFERENCE;
    const int D = model_dim(g.model);
    const int per = ls ? D * D : D;
    if (!ls && g.precision == BIFHIP_PREC_FP32 && g.model == BIFHIP_MODEL_MEMRISTIVE)
        return fail(BIFHIP_E_ARG, "the FP32 LLE path supports the Rossler and linear models");
    double* d_dirs = nullptr;
    if (refDirs) RC_TRY(pool_get("dirs", (size_t)std::min(chunk, cells) * per, &d_dirs));
    for (long long off = 0; off < cells; off += chunk) {
        const int n = (int)std::min(chunk, cells - off);
        LyapArgs A;
        A.g = S.G;
        A.h = S.h;
        A.eps = S.eps;
        A.tMax = S.tMax;
        A.maxValue = S.maxValue;
        A.cellBase = (int)off;
        A.cellList = nullptr;
        A.nCells = n;
        A.skip = S.skip;
        A.nNT = S.nNT;
        A.nBlocks = S.nBlocks;
        A.dirs = nullptr;
        if (refDirs) {
            RC_TRY(directions(ls, off, n, dirChunk, d_dirs, g.stream));
            A.dirs = d_dirs;
        }
        if (ls) {
            Timed t("ls");
            HIP_TRY(launch_ls(g.model, A, S.maxValue != 0, d_out + off * D, g.stream));
        } else {
            Timed t(g.precision == BIFHIP_PREC_FP32 ? "lle32" : "lle", nullptr, A.nCells);
            HIP_TRY(launch_lle(g.model, g.precision, A, S.maxValue != 0, d_out + off, g.stream));
        }
    }
    return BIFHIP_OK;
}

// ------------------------------------------- fused LLE + peak-count map

// LLE (LLE2D semantics) and DBSCAN peak count (bifurcation2DDBSCAN
// semantics) of n cells in one trajectory pass.  Cells are the flattened
// indices h_list[0..n) (or base..base+n when h_list is null); results go
// to d_lle[i], d_peak[i] for launch position i.
// Cells whose peaks did not fit their rows (entries pos = d_ovf[0..novf)
// of a listed or ranged batch): integrate them again with rows that hold
// every peak and cluster exactly, into d_peak[pos].
int recompute_overflow(const PeakMapSpec& S, int maxPeaks, double T, const int* h_list,
                       const int* d_list, int base, int novf, const int* d_ovf, int* d_peak)
{
    if (novf <= 0) return BIFHIP_OK;
    std::vector<int> pos(novf), flat(novf);
    HIP_TRY(hipMemcpy(pos.data(), d_ovf, sizeof(int) * novf, hipMemcpyDeviceToHost));
    const int cap2 = (int)std::max(2LL, std::min((long long)S.N / 2 + 1, (long long)maxPeaks + 1));
    DBuf dflat, dpos, cnt, am, iv;
    HIP_TRY(dflat.alloc(sizeof(int) * novf));
    HIP_TRY(dpos.alloc(sizeof(int) * novf));
    HIP_TRY(hipMemcpy(dpos.p, pos.data(), sizeof(int) * novf, hipMemcpyHostToDevice));
    if (!h_list && d_list) {
        // cell indices of the overflowing entries, gathered from the device list
        HIP_TRY(launch_gather(novf, dpos.as<int>(), d_list, dflat.as<int>(), g.stream));
    } else {
        for (int i = 0; i < novf; ++i) flat[i] = h_list ? h_list[pos[i]] : base + pos[i];
        HIP_TRY(hipMemcpy(dflat.p, flat.data(), sizeof(int) * novf, hipMemcpyHostToDevice));
    }
    HIP_TRY(cnt.alloc(sizeof(int) * novf));
    HIP_TRY(am.alloc(sizeof(double) * (size_t)novf * cap2));
    HIP_TRY(iv.alloc(sizeof(double) * (size_t)novf * cap2));
    RC_TRY(integrate_peaks(S, 0, dflat.as<int>(), novf, cap2, cnt.as<int>(), am.as<double>(),
                           iv.as<double>()));
    return dbscan_full_rows(novf, nullptr, novf, cnt.as<int>(), am.as<double>(), iv.as<double>(),
                            cap2, T, d_peak, dpos.as<int>());
}

int fused_cells_device(const PeakMapSpec& S, const LyapSpec& L, int maxPeaks, double epsDb,
                       const int* h_list, const int* d_list, int base, int n, double* d_lle,
                       int* d_peak)
{
    if (g.model == BIFHIP_MODEL_MEMRISTIVE)
        return fail(BIFHIP_E_ARG, "the fused map supports the Rossler and linear models");
    if (S.pre != 1) return fail(BIFHIP_E_ARG, "the fused map needs preScaller 1");
    RC_TRY(abandon_pending_peaks());
    if (S.maxValue != 0 || L.maxValue != 0)
        return fail(BIFHIP_E_ARG, "the fused map runs without divergence checks (maxValue 0)");
    const long long span = (long long)L.skip + (long long)L.nBlocks * L.nNT;
    if ((long long)S.skip + S.N > span)
        return fail(BIFHIP_E_ARG, "peak window [%d, %lld) is not inside the LLE trajectory (%lld steps)",
                    S.skip, (long long)S.skip + S.N, span);
    const int capMax = std::max(2, std::min(g.peakCap, S.N / 2 + 1));
    int cap = capMax;
    if ((long long)maxPeaks + 1 < cap) cap = (int)std::max(2LL, (long long)maxPeaks + 1);
    const long long rowBytes = (long long)cap * 16 + 8;
    long long pass = std::max<long long>(1, (8LL << 30) / rowBytes);
    pass = std::min<long long>(pass, std::max(1, n));
    if (g.maxCellsPerLaunch > 0) pass = std::min(pass, g.maxCellsPerLaunch);

    int *d_ovf, *d_novf;
    double* d_dirs = nullptr;
    PeakRows rows[2];
    RC_TRY(peak_rows("f", pass, cap, true, rows));
    RC_TRY(pool_get("f_ovf", (size_t)std::max(1, n), &d_ovf));
    RC_TRY(pool_get("f_novf", 1, &d_novf));
    HIP_TRY(launch_fill(1, d_novf, 0, g.stream));
    const double T = dbscan_threshold(epsDb);
    const bool refDirs = g.pert == BIFHIP_PERT_REFERENCE;
    const long long dirChunk = g.dirChunk >= 0 ? g.dirChunk : 200000;
    if (refDirs) RC_TRY(pool_get("f_dirs", (size_t)pass * 3, &d_dirs));

    auto integrate = [&](long long off, int m, const PeakRows& r) -> int {
        FusedArgs F;
        std::memset(&F, 0, sizeof(F));
        F.pk = bif_args(S, base + (int)off, d_list ? d_list + off : nullptr, m, cap);
        F.eps = L.eps;
        F.tMax = L.tMax;
        F.skipL = L.skip;
        F.nNT = L.nNT;
        F.nBlocks = L.nBlocks;
        if (refDirs) {
            // directions of the listed cells (their flattened indices)
            std::vector<int> seg;
            if (!h_list && d_list) {
                seg.resize((size_t)m);
                HIP_TRY(hipMemcpy(seg.data(), d_list + off, sizeof(int) * m, hipMemcpyDeviceToHost));
            }
            RC_TRY(directions_of(false, m, dirChunk, [&](int i) -> long long {
                return h_list ? h_list[off + i] : (d_list ? seg[i] : base + off + i);
            }, d_dirs, g.stream));
            F.dirs = d_dirs;
        }
        {
            Timed t("lle_peaks", nullptr, m);
            HIP_TRY(launch_lle_peaks(g.model, S.wv, F, d_lle + off, r.counts, r.amps, r.ivls,
                                     g.stream));
        }
        return BIFHIP_OK;
    };
    auto cluster = [&](long long off, int m, const PeakRows& r, hipStream_t st) -> int {
        Timed t("dbscan", st, m);
        HIP_TRY(launch_dbscan_wave(m, r.counts, r.amps, r.ivls, cap, cap, maxPeaks, T,
                                   d_peak + off, d_ovf, d_novf, (int)off, st));
        return BIFHIP_OK;
    };
    // with several passes the last is an eighth of a pass: the clustering of
    // the pass before runs beside its integration (a single pass is not
    // split: its short tail would integrate at low occupancy for as long as
    // the clustering it hides)
    const long long tail = n > pass ? std::max<long long>(1 << 16, pass / 8) : 0;
    RC_TRY(chunk_pipeline(n, pass, rows, integrate, cluster, true, tail));
    int novf = 0;
    HIP_TRY(hipMemcpyAsync(&novf, d_novf, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    RC_TRY(recompute_overflow(S, maxPeaks, T, h_list, d_list, base, novf, d_ovf, d_peak));
    if (g.clusterClamp > 0) HIP_TRY(launch_clamp(n, d_peak, g.clusterClamp, g.stream));
    return BIFHIP_OK;
}

static int peak_cells_begin(const PeakMapSpec& S, int maxPeaks, double epsDb, const int* h_list,
                            const int* d_list, int base, int n, int* d_peak, PeakCellsJob& J)
{
    J.pending = false;
    if (n <= 0) return BIFHIP_OK;
    const int capMax = std::max(2, std::min(g.peakCap, S.N / 2 + 1));
    int cap = capMax;
    if ((long long)maxPeaks + 1 < cap) cap = (int)std::max(2LL, (long long)maxPeaks + 1);
    const long long rowBytes = (long long)cap * 16 + 8;
    long long pass = std::max<long long>(1, (8LL << 30) / rowBytes);
    pass = std::min<long long>(pass, n);
    if (g.maxCellsPerLaunch > 0) pass = std::min(pass, g.maxCellsPerLaunch);
    int *d_ovf, *d_novf;
    PeakRows rows[2];
    RC_TRY(peak_rows("f", pass, cap, true, rows));
    RC_TRY(pool_get("f_ovf", (size_t)n, &d_ovf));
    RC_TRY(pool_get("f_novf", 1, &d_novf));
    HIP_TRY(launch_fill(1, d_novf, 0, g.stream));
    const double T = dbscan_threshold(epsDb);
    RC_TRY(chunk_pipeline(
        n, pass, rows,
        [&](long long off, int m, const PeakRows& r) {
            return integrate_peaks(S, base + (int)off, d_list ? d_list + off : nullptr, m, cap,
                                   r.counts, r.amps, r.ivls);
        },
        [&](long long off, int m, const PeakRows& r, hipStream_t st) {
            Timed t("dbscan", st, m);
            HIP_TRY(launch_dbscan_wave(m, r.counts, r.amps, r.ivls, cap, cap, maxPeaks, T,
                                       d_peak + off, d_ovf, d_novf, (int)off, st));
            return BIFHIP_OK;
        },
        false));
    J.S = S;
    J.maxPeaks = maxPeaks;
    J.T = T;
    J.h_list = h_list;
    J.d_list = d_list;
    J.base = base;
    J.n = n;
    J.d_peak = d_peak;
    J.d_ovf = d_ovf;
    J.d_novf = d_novf;
    J.pending = true;
    return BIFHIP_OK;
}

static int peak_cells_finish(PeakCellsJob& J)
{
    if (!J.pending) return BIFHIP_OK;
    J.pending = false;
    HIP_TRY(hipStreamSynchronize(g_side.s));   // the last clustering pass
    int novf = 0;
    HIP_TRY(hipMemcpyAsync(&novf, J.d_novf, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    RC_TRY(recompute_overflow(J.S, J.maxPeaks, J.T, J.h_list, J.d_list, J.base, novf, J.d_ovf,
                              J.d_peak));
    if (g.clusterClamp > 0) HIP_TRY(launch_clamp(J.n, J.d_peak, g.clusterClamp, g.stream));
    return BIFHIP_OK;
}

// DBSCAN peak count (bifurcation2DDBSCAN semantics) of listed cells:
// cells = h_list/d_list (flattened lattice indices) or base..base+n;
// d_peak[t] for entry t.
int peak_cells_device(const PeakMapSpec& S, int maxPeaks, double epsDb, const int* h_list,
                      const int* d_list, int base, int n, int* d_peak)
{
    RC_TRY(abandon_pending_peaks());
    PeakCellsJob J;
    RC_TRY(peak_cells_begin(S, maxPeaks, epsDb, h_list, d_list, base, n, d_peak, J));
    return peak_cells_finish(J);
}

// Plain FP64 DBSCAN peak count of listed cells (surrogate pipeline, cells
// whose LLE regime is predicted with confidence): flattened indices on an
// nPts x nPts lattice.
int surrogate_resolve_peaks(int nPts, double h, int nIC, const double* ic, const double* ranges,
                            const int* mut, const double* values, int nVal, double peakTMax,
                            double peakTransient, int wv, int maxPeaks, double dbEps,
                            const int* h_list, const int* d_list, int n, int* d_peak)
{
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 2, peakTMax, nPts, h, nIC, ic, ranges, mut, wv, 0.0, peakTransient,
                     values, nVal, 1));
    return peak_cells_device(S, maxPeaks, dbEps, h_list, d_list, 0, n, d_peak);
}

// The same, in two halves (one pending job at a time): the begin half
// enqueues the passes and returns with the last clustering pass still on the
// side stream, so independent work enqueued next overlaps it; the finish
// half completes the job before its results are read.
int surrogate_resolve_peaks_begin(int nPts, double h, int nIC, const double* ic,
                                  const double* ranges, const int* mut, const double* values,
                                  int nVal, double peakTMax, double peakTransient, int wv,
                                  int maxPeaks, double dbEps, const int* d_list, int n,
                                  int* d_peak)
{
    RC_TRY(abandon_pending_peaks());
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 2, peakTMax, nPts, h, nIC, ic, ranges, mut, wv, 0.0, peakTransient,
                     values, nVal, 1));
    return peak_cells_begin(S, maxPeaks, dbEps, nullptr, d_list, 0, n, d_peak, g_peakJob);
}

int surrogate_resolve_peaks_finish() { return peak_cells_finish(g_peakJob); }

// Plain LLE (LLE2D semantics) of n listed cells (flattened indices
// d_list[0..n) on an nPts x nPts lattice) into d_lle[0..n): the surrogate's
// re-solve of cells whose class is predicted with confidence.  precision
// BIFHIP_PREC_FP64 is the reference's FP64 kernel, BIFHIP_PREC_FP32 the
// FP32 bulk kernel (k_lle32).
int surrogate_resolve_lle(int nPts, double h, int nIC, const double* ic, const double* ranges,
                          const int* mut, const double* values, int nVal, double lleTMax,
                          double NT, double lleTransient, double lleEps, const int* d_list, int n,
                          double* d_lle, int precision)
{
    if (n <= 0) return BIFHIP_OK;
    if (g.model == BIFHIP_MODEL_MEMRISTIVE)
        return fail(BIFHIP_E_ARG, "the LLE re-solve supports the Rossler and linear models");
    LyapSpec L;
    RC_TRY(lyap_spec(L, 2, lleTMax, NT, nPts, h, lleEps, ic, nIC, ranges, mut, 0.0, lleTransient,
                     values, nVal));
    const long long chunk = launch_chunk(n);
    const bool refDirs = g.pert == BIFHIP_PERT_REFERENCE;
    const long long dirChunk = g.dirChunk >= 0 ? g.dirChunk : 200000;
    double* d_dirs = nullptr;
    std::vector<int> seg;
    if (refDirs) RC_TRY(pool_get("lr_dirs", (size_t)std::min<long long>(chunk, n) * 3, &d_dirs));
    for (long long off = 0; off < n; off += chunk) {
        const int m = (int)std::min<long long>(chunk, n - off);
        LyapArgs A;
        A.g = L.G;
        A.h = L.h;
        A.eps = L.eps;
        A.tMax = L.tMax;
        A.maxValue = 0.0;
        A.cellBase = 0;
        A.cellList = d_list + off;
        A.nCells = m;
        A.skip = L.skip;
        A.nNT = L.nNT;
        A.nBlocks = L.nBlocks;
        A.dirs = nullptr;
        if (refDirs) {
            // directions of the listed cells (their flattened indices)
            seg.resize((size_t)m);
            HIP_TRY(hipMemcpy(seg.data(), d_list + off, sizeof(int) * m, hipMemcpyDeviceToHost));
            RC_TRY(directions_of(false, m, dirChunk, [&](int i) -> long long { return seg[i]; },
                                 d_dirs, g.stream));
            A.dirs = d_dirs;
        }
        {
            Timed t(precision == BIFHIP_PREC_FP32 ? "lle32" : "lle", nullptr, m);
            HIP_TRY(launch_lle(g.model, precision, A, false, d_lle + off, g.stream));
        }
    }
    return BIFHIP_OK;
}

// Plain FP64 fused LLE + peak count of listed cells (surrogate pipeline):
// cells = h_list/d_list (flattened indices on an nPts x nPts lattice), or
// base..base+n when both lists are null.
int surrogate_resolve(int nPts, double h, int nIC, const double* ic, const double* ranges,
                      const int* mut, const double* values, int nVal, double peakTMax,
                      double peakTransient, int wv, int maxPeaks, double dbEps, double lleTMax,
                      double NT, double lleTransient, double lleEps, const int* h_list,
                      const int* d_list, int n, double* d_lle, int* d_peak)
{
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 2, peakTMax, nPts, h, nIC, ic, ranges, mut, wv, 0.0, peakTransient,
                     values, nVal, 1));
    LyapSpec L;
    RC_TRY(lyap_spec(L, 2, lleTMax, NT, nPts, h, lleEps, ic, nIC, ranges, mut, 0.0, lleTransient,
                     values, nVal));
    return fused_cells_device(S, L, maxPeaks, dbEps, h_list, d_list, 0, n, d_lle, d_peak);
}

// ------------------------------------------------------------ CSV output

// A value's text as the reference's stream writes it: printf's %.*g.
// std::to_chars in general format with a precision is specified to give
// printf's text (checked against snprintf on 2e7 values at precisions 1, 6,
// 12, 15 and 17) and is several times faster; non-finite values and other
// precisions keep snprintf.
constexpr int kFmtBuf = 384;

static int fmt_g(char* t, double v, int prec)
{
    if (std::isfinite(v) && prec >= 1 && prec <= 17) {
        const auto r = std::to_chars(t, t + kFmtBuf, v, std::chars_format::general, prec);
        return (int)(r.ptr - t);
    }
    const int n = std::snprintf(t, kFmtBuf, "%.*g", std::min(prec, 300), v);
    return std::min(n, kFmtBuf - 1);
}

static int fmt_d(char* t, int v)
{
    const auto r = std::to_chars(t, t + kFmtBuf, v);
    return (int)(r.ptr - t);
}

struct Out {
    FILE* f = nullptr;
    std::string buf;
    ~Out()
    {
        if (f) fclose(f);
    }
    bool open(const std::string& path, const char* mode)
    {
        f = fopen(path.c_str(), mode);
        return f != nullptr;
    }
    void num(double v, int prec)
    {
        char t[kFmtBuf];
        buf.append(t, (size_t)fmt_g(t, v, prec));
    }
    void num(int v)
    {
        char t[kFmtBuf];
        buf.append(t, (size_t)fmt_d(t, v));
    }
    void str(const char* s) { buf.append(s); }
    bool flush()
    {
        if (!f) return false;
        if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size()) return false;
        buf.clear();
        return true;
    }
    bool close()
    {
        bool ok = flush();
        if (f && fclose(f) != 0) ok = false;
        f = nullptr;
        return ok;
    }
};

void header(Out& o, const double* r, int prec)
{
    o.num(r[0], prec); o.str(" "); o.num(r[1], prec); o.str("\n");
    o.num(r[2], prec); o.str(" "); o.num(r[3], prec); o.str("\n");
}

// the reference's row logic: ", " before every value but the first, a
// line break once nPts values are on the row (Library/hostLibrary.cu:708-725)
struct RowWriter {
    int counter = 0;
    int nPts;
    explicit RowWriter(int n) : nPts(n) {}
    void sep(Out& o)
    {
        if (counter != 0) o.str(", ");
        if (counter == nPts) {
            o.str("\n");
            counter = 0;
        }
        ++counter;
    }
};

int file_error(const std::string& path)
{
    return fail(BIFHIP_E_FILE, "Output file open error: %s", path.c_str());
}

// Writes n values with the reference's row logic.  Value i's text is
// preceded by ", " when i > 0 and then by a line break when i % nPts == 0
// (RowWriter's output for nPts >= 1), so ranges of the index can be
// formatted independently: large maps are formatted by up to 16 host
// threads, a round of ranges at a time, and written in order.
template <class Fmt>
static int write_rows(const std::string& path, int nPts, const double* ranges, int hdr,
                      int prec, long long n, Fmt fmtValue)
{
    Out o;
    if (!o.open(path, "w")) return file_error(path);
    if (hdr) header(o, ranges, prec);
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int T = (int)std::max(1LL, std::min<long long>(std::min(16, hw), n / 16384));
    if (nPts <= 0 || T == 1) {
        RowWriter rw(nPts);
        char t[kFmtBuf];
        for (long long i = 0; i < n; ++i) {
            rw.sep(o);
            o.buf.append(t, (size_t)fmtValue(i, t));
            if (o.buf.size() > (1u << 22) && !o.flush()) return file_error(path);
        }
        if (!o.close()) return file_error(path);
        return BIFHIP_OK;
    }
    if (!o.flush()) return file_error(path);
    const long long part = std::min<long long>(1LL << 18, (n + T - 1) / T);  // per thread and round
    std::vector<std::string> text(T);
    auto fmtRange = [&](int k, long long i0, long long i1) {
        std::string& out = text[k];
        out.clear();
        char t[kFmtBuf];
        for (long long i = i0; i < i1; ++i) {
            if (i != 0) {
                out.append(", ", 2);
                if (i % nPts == 0) out.push_back('\n');
            }
            out.append(t, (size_t)fmtValue(i, t));
        }
    };
    for (long long r0 = 0; r0 < n; r0 += part * T) {
        run_parts(T, [&](int k) {
            const long long i0 = std::min(n, r0 + part * k), i1 = std::min(n, i0 + part);
            fmtRange(k, i0, i1);
        });
        for (int k = 0; k < T; ++k)
            if (!text[k].empty() &&
                fwrite(text[k].data(), 1, text[k].size(), o.f) != text[k].size())
                return file_error(path);
    }
    if (!o.close()) return file_error(path);
    return BIFHIP_OK;
}

int write_map_int(const std::string& path, int nPts, const double* ranges, int hdr, int prec,
                  const int* v, long long n)
{
    return write_rows(path, nPts, ranges, hdr, prec, n,
                      [v](long long i, char* t) { return fmt_d(t, v[i]); });
}

int write_map_double(const std::string& path, int nPts, const double* ranges, int hdr, int prec,
                     const double* v, long long n, int stride, int offset)
{
    return write_rows(path, nPts, ranges, hdr, prec, n, [=](long long i, char* t) {
        return fmt_g(t, v[i * stride + offset], prec);
    });
}

// One line per peak: "parameter, peak" at %.12g.  The points' lines are
// independent, so ranges of points are formatted by host threads and
// written in order.
int write_bif1d(const std::string& path, int nPts, const double* ranges, const int* counts,
                const double* peaks, int stride)
{
    Out o;
    if (!o.open(path, "w")) return file_error(path);
    long long lines = 0;
    for (int k = 0; k < nPts; ++k) lines += std::max(0, counts[k]);
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int T = (int)std::max(1LL, std::min<long long>(std::min(16, hw), lines / 8192));
    std::vector<std::string> text(T);
    auto fmtPoints = [&](int t, int k0, int k1) {
        std::string& out = text[t];
        char a[kFmtBuf], b[kFmtBuf];
        for (int k = k0; k < k1; ++k) {
            const int la = fmt_g(a, value_by_idx(k, nPts, ranges[0], ranges[1], 0), 12);
            for (int j = 0; j < counts[k]; ++j) {
                out.append(a, (size_t)la);
                out.append(", ", 2);
                out.append(b, (size_t)fmt_g(b, peaks[(long long)k * stride + j], 12));
                out.push_back('\n');
            }
        }
    };
    run_parts(T, [&](int t) {
        fmtPoints(t, (int)((long long)nPts * t / T), (int)((long long)nPts * (t + 1) / T));
    });
    for (int t = 0; t < T; ++t)
        if (!text[t].empty() && fwrite(text[t].data(), 1, text[t].size(), o.f) != text[t].size())
            return file_error(path);
    if (!o.close()) return file_error(path);
    return BIFHIP_OK;
}

int write_1d_values(const std::string& path, int nPts, const double* ranges, const double* v,
                    int per)
{
    Out o;
    if (!o.open(path, "w")) return file_error(path);
    for (int k = 0; k < nPts; ++k) {
        o.num(value_by_idx(k, nPts, ranges[0], ranges[1], 0), 6);
        for (int j = 0; j < per; ++j) {
            o.str(", ");
            o.num(v[(long long)k * per + j], 6);
        }
        o.str("\n");
    }
    if (!o.close()) return file_error(path);
    return BIFHIP_OK;
}

void progress(const char* name, long long cells)
{
    if (!g.verbose) return;
    printf("%s\n", name);
    printf("nPtsLimiter : %lld\n", cells);
    printf("Amount of iterations %d: \n", 1);
}

int finish()
{
    HIP_TRY(hipStreamSynchronize(g.stream));
    resolve_timers();
    return BIFHIP_OK;
}

// ----------------------------------------------------------- drivers

int bif1d_impl(double tMax, int nPts, double h, int nIC, const double* ic, const double* ranges,
               const int* mut, int wv, double maxValue, double transientTime,
               const double* values, int nVal, int pre, std::vector<int>& counts,
               std::vector<double>& peaks, int& stride)
{
    RC_TRY(ensure_device());
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 1, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, transientTime,
                     values, nVal, pre));
    RC_TRY(peak_precision(S));
    progress("BIFURCATION1D", S.cells);
    // rows of the library's peak capacity first (as the maps); a cell with
    // more peaks than that sends the call round again with rows that hold
    // every peak a record can have.  Only the used prefix of each row (the
    // largest count) comes back to the host, packed on the device (a
    // strided 2D copy to pageable memory took most of the call).
    const int fullCap = std::max(1, S.N / 2 + 1);
    int rowCap = std::min(fullCap, std::max(2, g.peakCap));
    DBuf c, a;
    HIP_TRY(c.alloc(sizeof(int) * (size_t)S.cells));
    counts.resize((size_t)S.cells);
    int used = 1;
    for (;;) {
        HIP_TRY(a.alloc(sizeof(double) * (size_t)S.cells * rowCap));
        RC_TRY(integrate_peaks(S, 0, nullptr, (int)S.cells, rowCap, c.as<int>(), a.as<double>(),
                               nullptr));
        HIP_TRY(hipMemcpyAsync(counts.data(), c.p, sizeof(int) * counts.size(),
                               hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        used = 1;
        for (int v : counts) used = std::max(used, v);
        if (used <= rowCap || rowCap == fullCap) break;
        rowCap = fullCap;
    }
    // peakFinderCUDA turns the divergence flag into 0 peaks
    for (auto& v : counts)
        if (v < 0) v = 0;
    stride = used;
    peaks.resize((size_t)S.cells * stride);
    DBuf packed;
    HIP_TRY(packed.alloc(sizeof(double) * peaks.size()));
    HIP_TRY(launch_pack_rows(S.cells, rowCap, stride, a.as<double>(), packed.as<double>(),
                             g.stream));
    HIP_TRY(hipMemcpyAsync(peaks.data(), packed.p, sizeof(double) * peaks.size(),
                           hipMemcpyDeviceToHost, g.stream));
    return finish();
}

int cluster_impl(int kind, double tMax, int nPts, double h, int nIC, const double* ic,
                 const double* ranges, const int* mut, int wv, double maxValue, int maxPeaks,
                 double transientTime, const double* values, int nVal, int pre, double eps,
                 int kdeSampling, double I1, double I2, double H, int* d_out)
{
    RC_TRY(ensure_device());
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 2, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, transientTime,
                     values, nVal, pre));
    RC_TRY(peak_precision(S));
    progress("BIFURCATION2D", S.cells);
    ClusterParams C{kind, maxPeaks, eps, kdeSampling, I1, I2, H};
    RC_TRY(cluster_map_device(S, C, d_out));
    return finish();
}

int cluster_to_host(int kind, double tMax, int nPts, double h, int nIC, const double* ic,
                    const double* ranges, const int* mut, int wv, double maxValue, int maxPeaks,
                    double transientTime, const double* values, int nVal, int pre, double eps,
                    int kdeSampling, double I1, double I2, double H, int* out)
{
    RC_TRY(ensure_device());
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    const long long cells = (long long)nPts * nPts;
    DBuf d;
    HIP_TRY(d.alloc(sizeof(int) * (size_t)cells));
    RC_TRY(cluster_impl(kind, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                        transientTime, values, nVal, pre, eps, kdeSampling, I1, I2, H,
                        d.as<int>()));
    HIP_TRY(hipMemcpy(out, d.p, sizeof(int) * (size_t)cells, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

int lyap_impl(bool ls, int dimension, double tMax, double NT, int nPts, double h, double eps,
              const double* ic, int nIC, const double* ranges, const int* mut, double maxValue,
              double transientTime, const double* values, int nVal, double* d_out)
{
    RC_TRY(ensure_device());
    LyapSpec S;
    RC_TRY(lyap_spec(S, dimension, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                     transientTime, values, nVal));
    progress(ls ? (dimension == 1 ? "LS1D" : "LS2D") : (dimension == 1 ? "LLE1D" : "LLE2D"),
             S.cells);
    // LLE2D seeds directions from a 200000-cell chunk-local index
    // (Library/hostLibrary.cu:1552); LLE2DIC's chunk is free-memory sized,
    // here one chunk
    const long long defChunk = (!ls && dimension == 2 && !g.sweepIC) ? 200000 : 0;
    RC_TRY(lyap_map_device(S, ls, defChunk, d_out));
    return finish();
}

int lyap_to_host(bool ls, int dimension, double tMax, double NT, int nPts, double h, double eps,
                 const double* ic, int nIC, const double* ranges, const int* mut,
                 double maxValue, double transientTime, const double* values, int nVal,
                 double* out)
{
    RC_TRY(ensure_device());
    if (nPts < 1 || dimension < 1 || dimension > 2) return fail(BIFHIP_E_ARG, "bad grid");
    const long long cells = cells_of(nPts, dimension);
    const size_t n = (size_t)cells * (ls ? nIC : 1);
    if (ls && nIC != model_dim(g.model))
        return fail(BIFHIP_E_ARG, "LS of model %d needs %d state variables (got %d)", g.model,
                    model_dim(g.model), nIC);
    DBuf d;
    HIP_TRY(d.alloc(sizeof(double) * n));
    RC_TRY(lyap_impl(ls, dimension, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                     transientTime, values, nVal, d.as<double>()));
    HIP_TRY(hipMemcpy(out, d.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

}  // namespace host
}  // namespace bifhip

using namespace bifhip;
using namespace bifhip::host;

// ================================================================ C ABI

extern "C" {

int bifhip_set_output_path(const char* path)
{
    if (!path || !*path) return fail(BIFHIP_E_ARG, "empty output path");
    g.outPath = path;
    return BIFHIP_OK;
}
const char* bifhip_get_output_path(void) { return g.outPath.c_str(); }

int bifhip_set_model(int model)
{
    if (model != BIFHIP_MODEL_ROSSLER && model != BIFHIP_MODEL_LINEAR &&
        model != BIFHIP_MODEL_MEMRISTIVE)
        return fail(BIFHIP_E_ARG, "unknown model %d", model);
    g.model = model;
    return BIFHIP_OK;
}

int bifhip_set_precision(int p)
{
    if (p != BIFHIP_PREC_FP64 && p != BIFHIP_PREC_FP32)
        return fail(BIFHIP_E_ARG, "unknown precision %d", p);
    g.precision = p;
    return BIFHIP_OK;
}

int bifhip_set_perturbation(int m)
{
    if (m != BIFHIP_PERT_REFERENCE && m != BIFHIP_PERT_AXIS_X)
        return fail(BIFHIP_E_ARG, "unknown perturbation mode %d", m);
    g.pert = m;
    return BIFHIP_OK;
}

int bifhip_set_direction_chunk(long long cells)
{
    if (cells < -1) return fail(BIFHIP_E_ARG, "direction chunk must be >= -1");
    g.dirChunk = cells;
    return BIFHIP_OK;
}

int bifhip_set_max_cells_per_launch(long long cells)
{
    if (cells < 0) return fail(BIFHIP_E_ARG, "max cells per launch must be >= 0");
    g.maxCellsPerLaunch = cells;
    return BIFHIP_OK;
}

int bifhip_set_peak_capacity(int cap)
{
    if (cap < 2) return fail(BIFHIP_E_ARG, "peak capacity must be >= 2");
    g.peakCap = cap;
    return BIFHIP_OK;
}

int bifhip_set_cluster_clamp(int clamp)
{
    if (clamp < 0) return fail(BIFHIP_E_ARG, "clamp must be >= 0");
    g.clusterClamp = clamp;
    return BIFHIP_OK;
}

int bifhip_set_verbose(int on)
{
    g.verbose = on;
    return BIFHIP_OK;
}

int bifhip_set_timing(int on)
{
    g.timing = on;
    return BIFHIP_OK;
}

int bifhip_kernel_time(const char* kernel, double* total_ms, long long* launches)
{
    if (!kernel) {
        g_times.clear();
        return BIFHIP_OK;
    }
    auto it = g_times.find(kernel);
    if (total_ms) *total_ms = it == g_times.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == g_times.end() ? 0 : it->second.launches;
    return BIFHIP_OK;
}

int bifhip_kernel_cells(const char* kernel, double* cells)
{
    if (!kernel || !cells) return fail(BIFHIP_E_ARG, "null argument");
    auto it = g_times.find(kernel);
    *cells = it == g_times.end() ? 0.0 : it->second.cells;
    return BIFHIP_OK;
}

int bifhip_release(void)
{
    RC_TRY(abandon_pending_peaks());
    if (g_side.s) HIP_TRY(hipStreamSynchronize(g_side.s));  // clustering still reading rows
    RC_TRY(release_training_staging());
    RC_TRY(release_dir_staging());
    g_pool.release();
    return BIFHIP_OK;
}

const char* bifhip_last_error(void) { return g.lastError.c_str(); }

int bifhip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* bifhip_version(void) { return "bifhip 0.1 gfx950"; }

int bifhip_device_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(BIFHIP_E_ARG, "null pointer");
    RC_TRY(ensure_device());
    *ptr = nullptr;
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 16));
    return BIFHIP_OK;
}

int bifhip_device_free(void* ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return BIFHIP_OK;
}

int bifhip_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BIFHIP_OK;
}

int bifhip_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

int bifhip_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return BIFHIP_OK;
}

int bifhip_device_memset(void* ptr, int value, size_t bytes)
{
    if (bytes) HIP_TRY(hipMemset(ptr, value, bytes));
    return BIFHIP_OK;
}

int bifhip_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    resolve_timers();
    return BIFHIP_OK;
}

double bifhip_value_by_idx(int idx, int nPts, double s, double f, int k)
{
    return value_by_idx(idx, nPts, s, f, k);
}

// ----------------------------------------------------------- file drivers

// Sweep target for the duration of one named driver call.
struct SweepGuard {
    int saved;
    explicit SweepGuard(bool ic) : saved(g.sweepIC) { g.sweepIC = ic ? 1 : 0; }
    ~SweepGuard() { g.sweepIC = saved; }
};

int bifhip_bifurcation1D(double tMax, int nPts, double h, int nIC, const double* ic,
                         const double* ranges, const int* mut, int wv, double maxValue,
                         double transientTime, const double* values, int nVal, int pre)
{
    SweepGuard sweep(false);
    std::vector<int> counts;
    std::vector<double> peaks;
    int stride = 0;
    RC_TRY(bif1d_impl(tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, transientTime, values,
                      nVal, pre, counts, peaks, stride));
    return write_bif1d(g.outPath, nPts, ranges, counts.data(), peaks.data(), stride);
}

int bifhip_bifurcation2DKDE(double tMax, int nPts, double h, int nIC, const double* ic,
                            const double* ranges, const int* mut, int wv, double maxValue,
                            int maxPeaks, double transientTime, const double* values, int nVal,
                            int pre, int kdeSampling, double I1, double I2, double H)
{
    SweepGuard sweep(false);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<int> out((size_t)nPts * nPts);
    RC_TRY(cluster_to_host(1, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                           transientTime, values, nVal, pre, 0.0, kdeSampling, I1, I2, H,
                           out.data()));
    // Library/hostLibrary.cu:474-480: setprecision(12) before the header
    return write_map_int(g.outPath, nPts, ranges, 1, 12, out.data(), (long long)out.size());
}

int bifhip_bifurcation2DDBSCAN(double tMax, int nPts, double h, int nIC, const double* ic,
                               const double* ranges, const int* mut, int wv, double maxValue,
                               int maxPeaks, double transientTime, const double* values,
                               int nVal, int pre, double eps)
{
    SweepGuard sweep(false);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<int> out((size_t)nPts * nPts);
    RC_TRY(cluster_to_host(0, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                           transientTime, values, nVal, pre, eps, 0, 0, 0, 0, out.data()));
    return write_map_int(g.outPath, nPts, ranges, 1, 6, out.data(), (long long)out.size());
}

int bifhip_LLE1D(double tMax, double NT, int nPts, double h, double eps, const double* ic,
                 int nIC, const double* ranges, const int* mut, int wv, double maxValue,
                 double transientTime, const double* values, int nVal)
{
    SweepGuard sweep(false);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<double> out((size_t)nPts);
    RC_TRY(lyap_to_host(false, 1, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out.data()));
    (void)wv;
    return write_1d_values(g.outPath, nPts, ranges, out.data(), 1);
}

int bifhip_LLE2D(double tMax, double NT, int nPts, double h, double eps, const double* ic,
                 int nIC, const double* ranges, const int* mut, int wv, double maxValue,
                 double transientTime, const double* values, int nVal)
{
    SweepGuard sweep(false);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<double> out((size_t)nPts * nPts);
    RC_TRY(lyap_to_host(false, 2, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out.data()));
    (void)wv;
    return write_map_double(g.outPath, nPts, ranges, 1, 6, out.data(), (long long)out.size(), 1,
                            0);
}

int bifhip_LS1D(double tMax, double NT, int nPts, double h, double eps, const double* ic,
                int nIC, const double* ranges, const int* mut, int wv, double maxValue,
                double transientTime, const double* values, int nVal)
{
    SweepGuard sweep(false);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<double> out((size_t)nPts * nIC);
    RC_TRY(lyap_to_host(true, 1, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out.data()));
    (void)wv;
    return write_1d_values(g.outPath, nPts, ranges, out.data(), nIC);
}

int bifhip_LS2D(double tMax, double NT, int nPts, double h, double eps, const double* ic,
                int nIC, const double* ranges, const int* mut, int wv, double maxValue,
                double transientTime, const double* values, int nVal)
{
    SweepGuard sweep(false);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    const long long cells = (long long)nPts * nPts;
    std::vector<double> out((size_t)cells * nIC);
    RC_TRY(lyap_to_host(true, 2, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out.data()));
    (void)wv;
    // Library/hostLibrary.cu:2003-2066: one file per exponent, path + "k.csv"
    for (int k = 0; k < nIC; ++k) {
        const std::string path = g.outPath + std::to_string(k + 1) + ".csv";
        RC_TRY(write_map_double(path, nPts, ranges, 1, 6, out.data(), cells, nIC, k));
    }
    return BIFHIP_OK;
}

// ------------------------------------- drivers sweeping initial conditions
//
// Same computations with the swept values written into the initial state
// (indicesOfMutVars index initialConditions), as the reference's *IC
// drivers; output formats as theirs.

// Library/hostLibrary.cu:243-393 (bifurcation1DIC): the bifurcation1D format.
int bifhip_bifurcation1DIC(double tMax, int nPts, double h, int nIC, const double* ic,
                           const double* ranges, const int* mut, int wv, double maxValue,
                           double transientTime, const double* values, int nVal, int pre)
{
    SweepGuard sweep(true);
    std::vector<int> counts;
    std::vector<double> peaks;
    int stride = 0;
    RC_TRY(bif1d_impl(tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, transientTime, values,
                      nVal, pre, counts, peaks, stride));
    return write_bif1d(g.outPath, nPts, ranges, counts.data(), peaks.data(), stride);
}

// Library/hostLibrary.cu:929-1090 (bifurcation2DKDEIC): no range header.
int bifhip_bifurcation2DKDEIC(double tMax, int nPts, double h, int nIC, const double* ic,
                              const double* ranges, const int* mut, int wv, double maxValue,
                              int maxPeaks, double transientTime, const double* values, int nVal,
                              int pre, int kdeSampling, double I1, double I2, double H)
{
    SweepGuard sweep(true);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<int> out((size_t)nPts * nPts);
    RC_TRY(cluster_to_host(1, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                           transientTime, values, nVal, pre, 0.0, kdeSampling, I1, I2, H,
                           out.data()));
    return write_map_int(g.outPath, nPts, ranges, 0, 6, out.data(), (long long)out.size());
}

// Library/hostLibrary.cu:1094-1266 (bifurcation2DDBSCANIC): with the header.
int bifhip_bifurcation2DDBSCANIC(double tMax, int nPts, double h, int nIC, const double* ic,
                                 const double* ranges, const int* mut, int wv, double maxValue,
                                 int maxPeaks, double transientTime, const double* values,
                                 int nVal, int pre, double eps)
{
    SweepGuard sweep(true);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<int> out((size_t)nPts * nPts);
    RC_TRY(cluster_to_host(0, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                           transientTime, values, nVal, pre, eps, 0, 0, 0, 0, out.data()));
    return write_map_int(g.outPath, nPts, ranges, 1, 6, out.data(), (long long)out.size());
}

// Library/hostLibrary.cu:1397-1520 (LLE1DIC): the LLE1D format.
int bifhip_LLE1DIC(double tMax, double NT, int nPts, double h, double eps, const double* ic,
                   int nIC, const double* ranges, const int* mut, int wv, double maxValue,
                   double transientTime, const double* values, int nVal)
{
    SweepGuard sweep(true);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<double> out((size_t)nPts);
    RC_TRY(lyap_to_host(false, 1, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out.data()));
    (void)wv;
    return write_1d_values(g.outPath, nPts, ranges, out.data(), 1);
}

// Library/hostLibrary.cu:1666-1798 (LLE2DIC): no range header.
int bifhip_LLE2DIC(double tMax, double NT, int nPts, double h, double eps, const double* ic,
                   int nIC, const double* ranges, const int* mut, int wv, double maxValue,
                   double transientTime, const double* values, int nVal)
{
    SweepGuard sweep(true);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    std::vector<double> out((size_t)nPts * nPts);
    RC_TRY(lyap_to_host(false, 2, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out.data()));
    (void)wv;
    return write_map_double(g.outPath, nPts, ranges, 0, 6, out.data(), (long long)out.size(), 1,
                            0);
}

// Library/hostLibrary.cu:748-925 (calculationOfPeriodicityByOstrovsky):
// the 2D map layout with the header, doubles at the stream's precision.
int bifhip_calculationOfPeriodicityByOstrovsky(double tMax, int nPts, double h, int nIC,
                                               const double* ic, const double* ranges,
                                               const int* mut, int wv, double maxValue,
                                               int maxPeaks, double transientTime,
                                               const double* values, int nVal, int pre,
                                               double eps, double ostrovskyThreshold)
{
    SweepGuard sweep(true);
    if (nPts < 1) return fail(BIFHIP_E_ARG, "nPts must be >= 1");
    const long long cells = (long long)nPts * nPts;
    std::vector<double> out((size_t)cells);
    {
        DBuf d;
        HIP_TRY(d.alloc(sizeof(double) * (size_t)cells));
        RC_TRY(ostrovsky_impl(tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                              transientTime, values, nVal, pre, eps, ostrovskyThreshold,
                              d.as<double>()));
        HIP_TRY(hipMemcpy(out.data(), d.p, sizeof(double) * (size_t)cells,
                          hipMemcpyDeviceToHost));
    }
    return write_map_double(g.outPath, nPts, ranges, 1, 6, out.data(), cells, 1, 0);
}

int bifhip_ostrovsky_to_buffer(double tMax, int nPts, double h, int nIC, const double* ic,
                               const double* ranges, const int* mut, int wv, double maxValue,
                               int maxPeaks, double transientTime, const double* values,
                               int nVal, int pre, double eps, double ostrovskyThreshold,
                               double* out)
{
    SweepGuard sweep(true);
    if (nPts < 1 || !out) return fail(BIFHIP_E_ARG, "bad arguments");
    const long long cells = (long long)nPts * nPts;
    DBuf d;
    HIP_TRY(d.alloc(sizeof(double) * (size_t)cells));
    RC_TRY(ostrovsky_impl(tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                          transientTime, values, nVal, pre, eps, ostrovskyThreshold,
                          d.as<double>()));
    HIP_TRY(hipMemcpy(out, d.p, sizeof(double) * (size_t)cells, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

// modelingOneSystemDenis (Library/hostLibrary.cu:162-239): derived counts
// (int)(tMax/h) output points, (int)(hSpecial/h) trajectories and
// (int)(tMax/hSpecial) coarse steps each; unwritten points are 0.
int bifhip_denis_to_buffer(double tMax, double h, double hSpecial, int nIC, const double* ic,
                           int wv, const double* values, int nVal, double* out, long long nOut)
{
    RC_TRY(ensure_device());
    if (!ic || !values || !out || !(h > 0) || !(hSpecial > 0))
        return fail(BIFHIP_E_ARG, "bad arguments");
    const long long nPoints = (long long)(int)(tMax / h);
    const int T = (int)(hSpecial / h);
    const int iters = (int)(tMax / hSpecial);
    if (nPoints < 0 || nOut < nPoints)
        return fail(BIFHIP_E_ARG, "output holds %lld points, %lld needed", nOut, nPoints);
    if (wv < 0 || wv >= nIC) return fail(BIFHIP_E_ARG, "writableVar %d out of range", wv);
    const double dummyRange[2] = {0, 0};
    const int dummyMut[1] = {0};
    Grid G;
    RC_TRY(make_grid(G, 1, 1, dummyRange, dummyMut, ic, nIC, values, nVal));
    G.dimension = 0;   // nothing swept
    DBuf d;
    HIP_TRY(d.alloc(sizeof(double) * (size_t)std::max(1LL, nPoints)));
    HIP_TRY(hipMemsetAsync(d.p, 0, sizeof(double) * (size_t)std::max(1LL, nPoints), g.stream));
    {
        Timed t("denis");
        HIP_TRY(launch_denis(g.model, G, T, h, hSpecial, iters, wv, nPoints, d.as<double>(),
                             g.stream));
    }
    HIP_TRY(hipMemcpyAsync(out, d.p, sizeof(double) * (size_t)nPoints, hipMemcpyDeviceToHost,
                           g.stream));
    return finish();
}

// Library/hostLibrary.cuh:59-67: "h*j, value" lines, 12 significant digits.
int bifhip_modelingOneSystemDenis(double tMax, double h, double hSpecial, int nIC,
                                  const double* ic, int wv, const double* values, int nVal)
{
    SweepGuard sweep(false);
    if (!(h > 0)) return fail(BIFHIP_E_ARG, "h must be positive");
    const long long nPoints = std::max(0LL, (long long)(int)(tMax / h));
    std::vector<double> out((size_t)std::max(1LL, nPoints));
    RC_TRY(bifhip_denis_to_buffer(tMax, h, hSpecial, nIC, ic, wv, values, nVal, out.data(),
                                  nPoints));
    FILE* f = std::fopen(g.outPath.c_str(), "w");
    if (!f) return file_error(g.outPath);
    for (long long j = 0; j < nPoints; ++j)
        std::fprintf(f, "%.12g, %.12g\n", h * (double)j, out[(size_t)j]);
    if (std::fclose(f) != 0) return file_error(g.outPath);
    return BIFHIP_OK;
}

int bifhip_set_sweep(int target)
{
    if (target != BIFHIP_SWEEP_PARAMETERS && target != BIFHIP_SWEEP_INITIAL_CONDITIONS)
        return fail(BIFHIP_E_ARG, "sweep target must be BIFHIP_SWEEP_PARAMETERS or "
                                  "BIFHIP_SWEEP_INITIAL_CONDITIONS");
    g.sweepIC = target == BIFHIP_SWEEP_INITIAL_CONDITIONS;
    return BIFHIP_OK;
}

// ----------------------------------------------------- host-buffer twins

int bifhip_bifurcation1D_to_buffer(double tMax, int nPts, double h, int nIC, const double* ic,
                                   const double* ranges, const int* mut, int wv, double maxValue,
                                   double transientTime, const double* values, int nVal, int pre,
                                   int* counts, double* peaks, int peakStride)
{
    if (!counts) return fail(BIFHIP_E_ARG, "null counts");
    // A call refused for its short stride keeps its diagram; the caller's
    // retry with the same arguments and the stride that max(counts) gives
    // copies it instead of integrating again.
    static struct {
        std::vector<unsigned char> key;
        std::vector<int> c;
        std::vector<double> p;
        int stride = 0;
    } kept;
    std::vector<unsigned char> key;
    auto put = [&](const void* q, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(q);
        key.insert(key.end(), b, b + n);
    };
    const int dimension = 1;
    put(&tMax, sizeof tMax); put(&nPts, sizeof nPts); put(&h, sizeof h); put(&nIC, sizeof nIC);
    put(&wv, sizeof wv); put(&maxValue, sizeof maxValue); put(&transientTime, sizeof transientTime);
    put(&nVal, sizeof nVal); put(&pre, sizeof pre); put(&g.model, sizeof g.model);
    put(&g.precision, sizeof g.precision); put(&g.sweepIC, sizeof g.sweepIC);
    put(&g.peakCap, sizeof g.peakCap);
    if (ic && nIC > 0) put(ic, sizeof(double) * (size_t)nIC);
    if (ranges) put(ranges, sizeof(double) * 2 * dimension);
    if (mut) put(mut, sizeof(int) * dimension);
    if (values && nVal > 0) put(values, sizeof(double) * (size_t)nVal);
    std::vector<int> c;
    std::vector<double> p;
    int stride = 0;
    if (!kept.key.empty() && kept.key == key) {
        c.swap(kept.c);
        p.swap(kept.p);
        stride = kept.stride;
    } else {
        RC_TRY(bif1d_impl(tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, transientTime,
                          values, nVal, pre, c, p, stride));
    }
    kept.key.clear();
    kept.c.clear();
    kept.p.clear();
    int most = 0;
    for (int k = 0; k < nPts; ++k) {
        counts[k] = c[k];
        most = std::max(most, c[k]);
    }
    if (peaks && most > peakStride) {
        kept.key.swap(key);
        kept.c.swap(c);
        kept.p.swap(p);
        kept.stride = stride;
        return fail(BIFHIP_E_LIMIT, "a cell has %d peaks > peakStride %d (counts are filled)",
                    most, peakStride);
    }
    if (peaks)
        for (int k = 0; k < nPts; ++k)
            for (int j = 0; j < c[k]; ++j)
                peaks[(long long)k * peakStride + j] = p[(long long)k * stride + j];
    return BIFHIP_OK;
}

int bifhip_bifurcation2DKDE_to_buffer(double tMax, int nPts, double h, int nIC, const double* ic,
                                      const double* ranges, const int* mut, int wv,
                                      double maxValue, int maxPeaks, double transientTime,
                                      const double* values, int nVal, int pre, int kdeSampling,
                                      double I1, double I2, double H, int* out)
{
    if (!out) return fail(BIFHIP_E_ARG, "null out");
    return cluster_to_host(1, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                           transientTime, values, nVal, pre, 0.0, kdeSampling, I1, I2, H, out);
}

int bifhip_bifurcation2DDBSCAN_to_buffer(double tMax, int nPts, double h, int nIC,
                                         const double* ic, const double* ranges, const int* mut,
                                         int wv, double maxValue, int maxPeaks,
                                         double transientTime, const double* values, int nVal,
                                         int pre, double eps, int* out)
{
    if (!out) return fail(BIFHIP_E_ARG, "null out");
    return cluster_to_host(0, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                           transientTime, values, nVal, pre, eps, 0, 0, 0, 0, out);
}

int bifhip_LLE_to_buffer(int dimension, double tMax, double NT, int nPts, double h, double eps,
                         const double* ic, int nIC, const double* ranges, const int* mut, int wv,
                         double maxValue, double transientTime, const double* values, int nVal,
                         double* out)
{
    (void)wv;
    if (!out) return fail(BIFHIP_E_ARG, "null out");
    return lyap_to_host(false, dimension, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut,
                        maxValue, transientTime, values, nVal, out);
}

int bifhip_LS_to_buffer(int dimension, double tMax, double NT, int nPts, double h, double eps,
                        const double* ic, int nIC, const double* ranges, const int* mut, int wv,
                        double maxValue, double transientTime, const double* values, int nVal,
                        double* out)
{
    (void)wv;
    if (!out) return fail(BIFHIP_E_ARG, "null out");
    return lyap_to_host(true, dimension, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                        transientTime, values, nVal, out);
}

// --------------------------------------------------- device-buffer twins

int bifhip_bifurcation2DDBSCAN_device(double tMax, int nPts, double h, int nIC, const double* ic,
                                      const double* ranges, const int* mut, int wv,
                                      double maxValue, int maxPeaks, double transientTime,
                                      const double* values, int nVal, int pre, double eps,
                                      int* d_out)
{
    if (!d_out) return fail(BIFHIP_E_ARG, "null d_out");
    return cluster_impl(0, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                        transientTime, values, nVal, pre, eps, 0, 0, 0, 0, d_out);
}

int bifhip_bifurcation2DKDE_device(double tMax, int nPts, double h, int nIC, const double* ic,
                                   const double* ranges, const int* mut, int wv,
                                   double maxValue, int maxPeaks, double transientTime,
                                   const double* values, int nVal, int pre, int kdeSampling,
                                   double I1, double I2, double H, int* d_out)
{
    if (!d_out) return fail(BIFHIP_E_ARG, "null d_out");
    return cluster_impl(1, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue, maxPeaks,
                        transientTime, values, nVal, pre, 0.0, kdeSampling, I1, I2, H, d_out);
}

int bifhip_LLE_device(int dimension, double tMax, double NT, int nPts, double h, double eps,
                      const double* ic, int nIC, const double* ranges, const int* mut, int wv,
                      double maxValue, double transientTime, const double* values, int nVal,
                      double* d_out)
{
    (void)wv;
    if (!d_out) return fail(BIFHIP_E_ARG, "null d_out");
    return lyap_impl(false, dimension, tMax, NT, nPts, h, eps, ic, nIC, ranges, mut, maxValue,
                     transientTime, values, nVal, d_out);
}

int bifhip_lle_peak_map_device(int nPts, double h, int nIC, const double* ic,
                               const double* ranges, const int* mut, const double* values,
                               int nVal, double peakTMax, double peakTransient, int writableVar,
                               int maxAmountOfPeaks, double dbscanEps, double lleTMax, double NT,
                               double lleTransient, double lleEps, int* d_peak, double* d_lle)
{
    RC_TRY(ensure_device());
    if (!d_peak || !d_lle) return fail(BIFHIP_E_ARG, "null device output");
    PeakMapSpec S;
    RC_TRY(peak_spec(S, 2, peakTMax, nPts, h, nIC, ic, ranges, mut, writableVar, 0.0,
                     peakTransient, values, nVal, 1));
    LyapSpec L;
    RC_TRY(lyap_spec(L, 2, lleTMax, NT, nPts, h, lleEps, ic, nIC, ranges, mut, 0.0, lleTransient,
                     values, nVal));
    progress("LLE2D+BIFURCATION2D", S.cells);
    RC_TRY(fused_cells_device(S, L, maxAmountOfPeaks, dbscanEps, nullptr, nullptr, 0,
                              (int)S.cells, d_lle, d_peak));
    return finish();
}

// ------------------------------------------------ kernel-level entry points

int bifhip_peak_finder_device(const double* d_data, int sizeOfBlock, int nBlocks, int* d_counts,
                              double* d_outPeaks, double* d_timeOfPeaks, double h)
{
    RC_TRY(ensure_device());
    if (!d_data || !d_counts || sizeOfBlock < 0 || nBlocks < 0)
        return fail(BIFHIP_E_ARG, "bad peak finder arguments");
    {
        Timed t("peak_finder");
        HIP_TRY(launch_peak_finder(d_data, sizeOfBlock, nBlocks, d_counts, d_outPeaks,
                                   d_timeOfPeaks, h, g.stream));
    }
    return finish();
}

int bifhip_dbscan_device(const double* d_peaks, const double* d_intervals, int sizeOfBlock,
                         int nBlocks, const int* d_counts, int maxPeaks, double eps, int* d_out)
{
    RC_TRY(ensure_device());
    if (!d_peaks || !d_intervals || !d_counts || !d_out || sizeOfBlock < 1 || nBlocks < 0)
        return fail(BIFHIP_E_ARG, "bad dbscan arguments");
    const double T = dbscan_threshold(eps);
    int *d_ovf, *d_novf;
    RC_TRY(pool_get("ovf", (size_t)std::max(1, nBlocks), &d_ovf));
    RC_TRY(pool_get("novf", 1, &d_novf));
    HIP_TRY(launch_fill(1, d_novf, 0, g.stream));
    {
        Timed t("dbscan");
        HIP_TRY(launch_dbscan_wave(nBlocks, d_counts, d_peaks, d_intervals, sizeOfBlock,
                                   sizeOfBlock, maxPeaks, T, d_out, d_ovf, d_novf, 0, g.stream));
    }
    int novf = 0;
    HIP_TRY(hipMemcpyAsync(&novf, d_novf, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (novf > 0)
        RC_TRY(dbscan_full_rows(novf, d_ovf, nBlocks, d_counts, d_peaks, d_intervals,
                                sizeOfBlock, T, d_out, nullptr));
    return finish();
}

int bifhip_kde_device(const double* d_peaks, int sizeOfBlock, int nBlocks, const int* d_counts,
                      int maxPeaks, int kdeSampling, double I1, double I2, double H, int* d_out)
{
    RC_TRY(ensure_device());
    if (!d_peaks || !d_counts || !d_out || sizeOfBlock < 1 || nBlocks < 0)
        return fail(BIFHIP_E_ARG, "bad kde arguments");
    int *d_ovf, *d_novf;
    RC_TRY(pool_get("ovf", (size_t)std::max(1, nBlocks), &d_ovf));
    RC_TRY(pool_get("novf", 1, &d_novf));
    HIP_TRY(launch_fill(1, d_novf, 0, g.stream));
    {
        Timed t("kde");
        // rows hold every peak (cap = sizeOfBlock), so nothing overflows
        HIP_TRY(launch_kde_wave(nBlocks, d_counts, d_peaks, sizeOfBlock, sizeOfBlock, maxPeaks,
                                kdeSampling, I1, I2, H, d_out, nullptr, d_ovf, d_novf, 0,
                                g.stream));
    }
    return finish();
}

int bifhip_gram_schmidt_device(const double* d_a, double* d_b, double* d_den, int n, int count)
{
    RC_TRY(ensure_device());
    if (!d_a || !d_b || n < 1 || n > 8 || count < 0)
        return fail(BIFHIP_E_ARG, "bad Gram-Schmidt arguments (n must be 1..8)");
    {
        Timed t("gram_schmidt");
        HIP_TRY(launch_gs(d_a, d_b, d_den, n, count, g.stream));
    }
    return finish();
}

int bifhip_series_device(int dimension, double tMax, int nPts, double h, int nIC,
                         const double* ic, const double* ranges, const int* mut, int wv,
                         double maxValue, double transientTime, const double* values, int nVal,
                         int pre, double* d_data, int* d_flags)
{
    RC_TRY(ensure_device());
    if (!d_data || !d_flags) return fail(BIFHIP_E_ARG, "null device buffer");
    PeakMapSpec S;
    RC_TRY(peak_spec(S, dimension, tMax, nPts, h, nIC, ic, ranges, mut, wv, maxValue,
                     transientTime, values, nVal, pre));
    BifArgs A = bif_args(S, 0, nullptr, (int)S.cells, S.N);
    A.rowStride = S.N;
    {
        Timed t("series");
        HIP_TRY(launch_series(g.model, A, wv, d_data, d_flags, g.stream));
    }
    return finish();
}

// ------------------------------------------------------------ writers

int bifhip_write_map_int(const char* path, int nPts, const double* ranges, int hdr, int prec,
                         const int* values, long long count)
{
    if (!path || !values || (hdr && !ranges) || nPts < 1) return fail(BIFHIP_E_ARG, "bad writer args");
    return write_map_int(path, nPts, ranges, hdr, prec, values, count);
}

int bifhip_write_map_double(const char* path, int nPts, const double* ranges, int hdr, int prec,
                            const double* values, long long count, int stride, int offset)
{
    if (!path || !values || (hdr && !ranges) || nPts < 1 || stride < 1)
        return fail(BIFHIP_E_ARG, "bad writer args");
    return write_map_double(path, nPts, ranges, hdr, prec, values, count, stride, offset);
}

int bifhip_write_bif1d(const char* path, int nPts, const double* ranges, const int* counts,
                       const double* peaks, int stride)
{
    if (!path || !ranges || !counts || !peaks || nPts < 1) return fail(BIFHIP_E_ARG, "bad writer args");
    return write_bif1d(path, nPts, ranges, counts, peaks, stride);
}

int bifhip_write_1d_values(const char* path, int nPts, const double* ranges,
                           const double* values, int perCell)
{
    if (!path || !ranges || !values || nPts < 1 || perCell < 1) return fail(BIFHIP_E_ARG, "bad writer args");
    return write_1d_values(path, nPts, ranges, values, perCell);
}

}  // extern "C"
