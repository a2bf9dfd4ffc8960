// host_mlp.cpp -- surrogate-refined LLE + peak-count map.
//
// 1. coarse map: the fused FP64 kernel on an nCoarse x nCoarse lattice;
// 2. training: 64 Fourier features -> 3 x 128 ReLU -> head (LLE + 16
//    peak-count classes + chaos logit), BF16 MFMA products with FP32
//    accumulation; a step is one fused forward/backward launch
//    (k_train_fused, gradient partials per 64 rows) and one Adam launch on
//    FP32 master weights; one epoch is captured as a HIP graph and replayed;
// 3. inference of the nFine x nFine lattice in chunks;
// 4. cells with low confidence or a neighbour of another class / LLE regime
//    are re-solved with the plain FP64 fused kernel, which also scores the
//    surrogate on them; a seeded sample of the remaining cells is re-solved
//    as validation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "bifhip.h"
#include "host_internal.h"

namespace bifhip {
typedef __bf16 bf16;
hipError_t launch_sum_splits(int n, int splits, const float* part, float* out, hipStream_t st);
hipError_t launch_infer_fused(long long rows, long long base, const int* list, int nLat,
                              const float* Bff, const bf16* wbf, const float* p, float lleMean,
                              float lleStd, float chaosThr, unsigned char* cls,
                              unsigned char* chaos, float* conf, float* confChaos, float* lle,
                              float* rawOut, hipStream_t st);
int train_fused_blocks(int rows);
hipError_t launch_train_fused(int rows, int nLat, const int* list, const int* tlist,
                              const float* Bff, const bf16* wbf, const bf16* wbfT, const float* p,
                              const int* cls, const float* lle, float wLle, float chaosStd,
                              float* gpart, long long stride, float* loss, int* step,
                              hipStream_t st);
hipError_t launch_adam(int n, float* p, float* m, float* v, const float* gpart, int splits,
                       long long stride, const unsigned char* keep, const int* bfIndex, bf16* wbf,
                       const int* bfIndexT, bf16* wbfT, float lr, float b1, float b2, float eps,
                       const int* step, hipStream_t st);
hipError_t launch_fill_f32(int n, float* p, float v, hipStream_t st);
hipError_t launch_fill(long long n, int* v, int value, hipStream_t st);
hipError_t launch_fill_f64(int n, double* p, double v, hipStream_t st);
hipError_t launch_flag(int n, const unsigned char* cls, const unsigned char* chaos,
                       const float* conf, const float* confChaos, float confThr,
                       float chaosConfThr, bool partial, int reading,
                       const unsigned char* coarseChaos, int nCoarse, float quadConf,
                       int* listFull, int* countFull, int* listPeak, int* countPeak, int* listLle, int* countLle,
                       unsigned long long vseed, unsigned int vthr, int* listVal, int* countVal,
                       unsigned char* flag, hipStream_t st);
hipError_t launch_scatter(int n, const int* list, const int* peak, const double* lle,
                          int* peakMap, double* lleMap, hipStream_t st);
hipError_t launch_fill_pred(long long n, const unsigned char* cls, const float* lle,
                            int* peakMap, double* lleMap, hipStream_t st);
hipError_t launch_regime_changed(int n, const int* list, const double* lle,
                                 const unsigned char* pch, double chaosThr, int* out, int* count,
                                 unsigned char* flag, hipStream_t st);
hipError_t launch_lle32_split(int n, const int* list, const double* lle32, double chaosThr,
                              double margin, int* cells, int* pos, int* count,
                              unsigned char* flag, hipStream_t st);
hipError_t launch_score(int nFull, int nTot, const int* list, const int* rpeak, const double* rlle,
                        int nPeak, const int* plist, const int* ppeak, int nLle,
                        const int* llist, const double* llle, const unsigned char* pcls,
                        const unsigned char* pch, const float* plle, double chaosThr,
                        double* acc, hipStream_t st);
namespace host {
struct PeakMapSpec;
struct LyapSpec;
int surrogate_resolve(int nPts, double h, int nIC, const double* ic, const double* ranges,
                      const int* mut, const double* values, int nVal, double peakTMax,
                      double peakTransient, int wv, int maxPeaks, double dbEps, double lleTMax,
                      double NT, double lleTransient, double lleEps, const int* h_list,
                      const int* d_list, int n, double* d_lle, int* d_peak);
int surrogate_resolve_peaks(int nPts, double h, int nIC, const double* ic, const double* ranges,
                            const int* mut, const double* values, int nVal, double peakTMax,
                            double peakTransient, int wv, int maxPeaks, double dbEps,
                            const int* h_list, const int* d_list, int n, int* d_peak);
int surrogate_resolve_peaks_begin(int nPts, double h, int nIC, const double* ic,
                                  const double* ranges, const int* mut, const double* values,
                                  int nVal, double peakTMax, double peakTransient, int wv,
                                  int maxPeaks, double dbEps, const int* d_list, int n,
                                  int* d_peak);
int surrogate_resolve_peaks_finish();
int surrogate_resolve_lle(int nPts, double h, int nIC, const double* ic, const double* ranges,
                          const int* mut, const double* values, int nVal, double lleTMax,
                          double NT, double lleTransient, double lleEps, const int* d_list, int n,
                          double* d_lle, int precision);
}  // namespace host
}  // namespace bifhip

using namespace bifhip;
using namespace bifhip::host;

namespace {

constexpr int kFeat = 64, kHid = 128, kHead = 32, kClasses = 16;
constexpr int kLiveHead = 2 + kClasses;   // LLE, 16 classes, chaos logit
// gradient partials: one per 64 rows of a batch (k_train_fused), so batches
// of up to 64 * kMaxSplits rows
constexpr int kMaxSplits = 64;

// flat FP32 parameter layout (row-major [out][in] matrices)
constexpr int oW1 = 0, oB1 = oW1 + kHid * kFeat;
constexpr int oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid;
constexpr int oWh = oB3 + kHid, oBh = oWh + kHead * kHid;
constexpr int kParams = oBh + kHead;
// BF16 weight copies, and transposed copies of W2, W3 and Wh for the
// backward products
constexpr int bW1 = 0, bW2 = bW1 + kHid * kFeat, bW3 = bW2 + kHid * kHid,
              bWh = bW3 + kHid * kHid, kBf = bWh + kHead * kHid;
constexpr int tW2 = 0, tW3 = tW2 + kHid * kHid, tWh = tW3 + kHid * kHid,
              kBfT = tWh + kHid * kHead;

struct Net {
    // gpart[z * kParams + i]: partial z of the gradient of parameter i
    float *p, *m, *v, *g, *bff, *loss, *gpart;
    bf16 *wbf, *wbfT;
    unsigned char* keep;
    int *bfIndex, *bfIndexT, *step;
};

int net_alloc(Net& N)
{
    RC_TRY(pool_get("mlp_p", kParams, &N.p));
    RC_TRY(pool_get("mlp_m", kParams, &N.m));
    RC_TRY(pool_get("mlp_v", kParams, &N.v));
    RC_TRY(pool_get("mlp_g", kParams, &N.g));
    RC_TRY(pool_get("mlp_gpart", (size_t)kMaxSplits * kParams, &N.gpart));
    RC_TRY(pool_get("mlp_bff", kFeat, &N.bff));
    RC_TRY(pool_get("mlp_loss", 3, &N.loss));
    RC_TRY(pool_get("mlp_wbf", kBf, &N.wbf));
    RC_TRY(pool_get("mlp_wbft", kBfT, &N.wbfT));
    RC_TRY(pool_get("mlp_keep", kParams, &N.keep));
    RC_TRY(pool_get("mlp_bfidx", kParams, &N.bfIndex));
    RC_TRY(pool_get("mlp_bfidxt", kParams, &N.bfIndexT));
    RC_TRY(pool_get("mlp_step", 1, &N.step));
    return BIFHIP_OK;
}

// Position of each parameter in the BF16 weight copies (-1: a bias) and in
// the transposed copies (-1: none).
void bf16_maps(std::vector<int>& bfi, std::vector<int>& bfiT)
{
    bfi.assign(kParams, -1);
    bfiT.assign(kParams, -1);
    auto map = [&](int off, int rows, int cols, int bfoff, int toff) {
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                bfi[off + r * cols + c] = bfoff + r * cols + c;
                if (toff >= 0) bfiT[off + r * cols + c] = toff + c * rows + r;
            }
    };
    map(oW1, kHid, kFeat, bW1, -1);
    map(oW2, kHid, kHid, bW2, tW2);
    map(oW3, kHid, kHid, bW3, tW3);
    map(oWh, kHead, kHid, bWh, tWh);
}

// BF16 copies (plain and transposed) of the FP32 parameters p, with the
// index maps that k_adam refreshes them through
int upload_weights(const Net& N, const float* p)
{
    std::vector<int> bfi, bfiT;
    bf16_maps(bfi, bfiT);
    std::vector<bf16> wbf(kBf), wbfT(kBfT);
    for (int i = 0; i < kParams; ++i) {
        if (bfi[i] >= 0) wbf[bfi[i]] = (bf16)p[i];
        if (bfiT[i] >= 0) wbfT[bfiT[i]] = (bf16)p[i];
    }
    HIP_TRY(hipMemcpy(N.wbf, wbf.data(), sizeof(bf16) * kBf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(N.wbfT, wbfT.data(), sizeof(bf16) * kBfT, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(N.bfIndex, bfi.data(), sizeof(int) * kParams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(N.bfIndexT, bfiT.data(), sizeof(int) * kParams, hipMemcpyHostToDevice));
    return BIFHIP_OK;
}

// seeded initialisation: He-uniform weights, zero biases, Fourier matrix
// N(0, sigma^2); padded head rows (17..31) stay zero
int net_init(Net& N, unsigned long long seed, float sigma)
{
    std::mt19937_64 rng(seed);
    auto u01 = [&]() { return (double)(rng() >> 11) * (1.0 / 9007199254740992.0); };
    std::vector<float> p(kParams, 0.0f);
    std::vector<unsigned char> keep(kParams, 1);
    auto fill = [&](int off, int rows, int cols, double scale, int liveRows) {
        const double lim = scale * std::sqrt(6.0 / cols);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                const int i = off + r * cols + c;
                p[i] = r < liveRows ? (float)((2 * u01() - 1) * lim) : 0.0f;
                keep[i] = r < liveRows;
            }
    };
    fill(oW1, kHid, kFeat, 1.0, kHid);
    fill(oW2, kHid, kHid, 1.0, kHid);
    fill(oW3, kHid, kHid, 1.0, kHid);
    fill(oWh, kHead, kHid, 0.5, kLiveHead);
    for (int i = kLiveHead; i < kHead; ++i) keep[oBh + i] = 0;
    std::vector<float> bff(kFeat);
    for (int i = 0; i < kFeat; i += 2) {
        const double r = std::sqrt(-2.0 * std::log(1.0 - u01()));
        const double th = 2.0 * M_PI * u01();
        bff[i] = (float)(sigma * r * std::cos(th));
        bff[i + 1] = (float)(sigma * r * std::sin(th));
    }
    HIP_TRY(hipMemcpy(N.p, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(N.m, 0, sizeof(float) * kParams));
    HIP_TRY(hipMemset(N.v, 0, sizeof(float) * kParams));
    HIP_TRY(hipMemcpy(N.bff, bff.data(), sizeof(float) * kFeat, hipMemcpyHostToDevice));
    RC_TRY(upload_weights(N, p.data()));
    HIP_TRY(hipMemcpy(N.keep, keep.data(), kParams, hipMemcpyHostToDevice));
    const int zero = 0;
    HIP_TRY(hipMemcpy(N.step, &zero, sizeof(int), hipMemcpyHostToDevice));
    return BIFHIP_OK;
}

struct TrainCfg {
    int epochs = 20, batch = 1024;
    float lr = 1e-3f, wLle = 1.0f, chaosStd = 0.0f;
    unsigned long long seed = 0;
    bool graph = true;
};

// one optimisation step on rows of the permuted sample list: the fused
// forward/backward launch (gradient partial per 64 rows, step counter
// advanced) and Adam over the partials in block order
int train_step(const Net& N, const TrainCfg& T, int nLat, const int* list, const int* cls,
               const float* lle, hipStream_t s)
{
    HIP_TRY(launch_train_fused(T.batch, nLat, list, nullptr, N.bff, N.wbf, N.wbfT, N.p, cls, lle,
                               T.wLle, T.chaosStd, N.gpart, kParams, N.loss, N.step, s));
    HIP_TRY(launch_adam(kParams, N.p, N.m, N.v, N.gpart, train_fused_blocks(T.batch), kParams,
                        N.keep, N.bfIndex, N.wbf, N.bfIndexT, N.wbfT, T.lr, 0.9f, 0.999f, 1e-8f,
                        N.step, s));
    return BIFHIP_OK;
}

// Inference of the whole nLat x nLat lattice (one fused launch); confChaos
// as in k_infer_fused (null: conf = min of class and chaos confidence).
int infer_impl(int nLat, double lleMean, double lleSd, double chaosThr, unsigned char* d_cls,
               unsigned char* d_chaos, float* d_conf, float* d_confChaos, float* d_lle)
{
    RC_TRY(ensure_device());
    if (nLat < 2 || !d_cls || !d_chaos || !d_conf || !d_lle)
        return fail(BIFHIP_E_ARG, "bad inference arguments");
    Net N;
    RC_TRY(net_alloc(N));
    const long long n = (long long)nLat * nLat;
    Timed t("mlp_infer");
    HIP_TRY(launch_infer_fused(n, 0, nullptr, nLat, N.bff, N.wbf, N.p, (float)lleMean,
                               (float)lleSd, (float)chaosThr, d_cls, d_chaos, d_conf, d_confChaos,
                               d_lle, nullptr, g.stream));
    return BIFHIP_OK;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

// Training uploads one shuffled sample order per epoch.  It is built in one
// of two pinned host buffers and copied asynchronously; a buffer is written
// again only after the copy that read it has completed (its event).  The
// training graph is captured and replayed on one long-lived non-blocking
// stream.
struct TrainStaging {
    hipStream_t cs = nullptr;
    int* perm[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    size_t cap = 0;                   // ints per buffer
};
TrainStaging g_train;

int train_staging(size_t n)
{
    TrainStaging& T = g_train;
    if (!T.cs) HIP_TRY(hipStreamCreateWithFlags(&T.cs, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b)
        if (!T.copied[b]) HIP_TRY(hipEventCreateWithFlags(&T.copied[b], hipEventDisableTiming));
    if (T.cap >= n) return BIFHIP_OK;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipEventSynchronize(T.copied[b]));
        if (T.perm[b]) HIP_TRY(hipHostFree(T.perm[b]));
        T.perm[b] = nullptr;
    }
    T.cap = 0;
    for (int b = 0; b < 2; ++b)
        HIP_TRY(hipHostMalloc((void**)&T.perm[b], sizeof(int) * n, hipHostMallocDefault));
    T.cap = n;
    return BIFHIP_OK;
}

}  // namespace

namespace bifhip {
namespace host {
// Frees the training staging buffers once their copies have completed.
int release_training_staging()
{
    TrainStaging& T = g_train;
    if (T.cs) HIP_TRY(hipStreamSynchronize(T.cs));
    for (int b = 0; b < 2; ++b) {
        if (T.copied[b]) HIP_TRY(hipEventSynchronize(T.copied[b]));
        if (T.perm[b]) HIP_TRY(hipHostFree(T.perm[b]));
        T.perm[b] = nullptr;
    }
    T.cap = 0;
    return BIFHIP_OK;
}
}  // namespace host
}  // namespace bifhip

extern "C" {

// Train the surrogate on a labelled lattice already in device memory:
// d_cls[nLat^2] class 0..15, d_lle[nLat^2] standardised LLE (NaN = no
// target); chaosStd = the chaos threshold in standardised units.  Mean
// losses (CE, MSE, chaos BCE) of the last epoch to loss_out[3] when not NULL.
int bifhip_mlp_train_device(int nLat, const int* d_cls, const float* d_lle, int epochs,
                            int batch, double lr, double wLle, unsigned long long seed,
                            double fourierSigma, double chaosStd, int useGraph, float* loss_out)
{
    RC_TRY(ensure_device());
    const int n = nLat * nLat;
    if (nLat < 2 || batch < 16 || batch > n || epochs < 0 || !d_cls || !d_lle)
        return fail(BIFHIP_E_ARG, "bad training arguments");
    if (train_fused_blocks(batch) > kMaxSplits)
        return fail(BIFHIP_E_ARG, "batch %d above %d rows", batch, 64 * kMaxSplits);
    Net N;
    RC_TRY(net_alloc(N));
    RC_TRY(net_init(N, seed, (float)fourierSigma));
    int* d_perm;
    const int steps = n / batch;
    RC_TRY(pool_get("tr_perm", (size_t)steps * batch, &d_perm));
    TrainCfg T;
    T.epochs = epochs;
    T.batch = batch;
    T.lr = (float)lr;
    T.wLle = (float)wLle;
    T.chaosStd = (float)chaosStd;
    T.seed = seed;
    std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ull);
    RC_TRY(train_staging((size_t)n));
    // graph replay on the capture stream, direct launches on the library's
    hipStream_t s = useGraph ? g_train.cs : g.stream;
    if (useGraph) {
        // the capture stream starts after the library stream's work (labels)
        hipEvent_t ready;
        HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        hipError_t e1 = hipEventRecord(ready, g.stream);
        hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(s, ready, 0) : e1;
        (void)hipEventDestroy(ready);
        HIP_TRY(e2);
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = BIFHIP_OK;
    for (int e = 0; e < epochs && rc == BIFHIP_OK; ++e) {
        // epoch e's order in staging buffer e % 2, once the copy that read
        // it two epochs ago has completed
        const int b = e & 1;
        int* perm = g_train.perm[b];
        if (hipEventSynchronize(g_train.copied[b]) != hipSuccess) {
            rc = fail(1, "staging wait");
            break;
        }
        std::iota(perm, perm + n, 0);
        for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng() % (uint64_t)(i + 1)]);
        if (hipMemcpyAsync(d_perm, perm, sizeof(int) * steps * batch, hipMemcpyHostToDevice,
                           s) != hipSuccess ||
            hipEventRecord(g_train.copied[b], s) != hipSuccess) {
            rc = fail(1, "order upload");
            break;
        }
        if (useGraph) {
            if (!exec) {
                if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                    rc = fail(1, "graph capture failed");
                    break;
                }
                int crc = BIFHIP_OK;
                if (launch_fill_f32(3, N.loss, 0.0f, s) != hipSuccess) crc = fail(1, "fill");
                for (int k = 0; k < steps && crc == BIFHIP_OK; ++k)
                    crc = train_step(N, T, nLat, d_perm + (size_t)k * batch, d_cls, d_lle, s);
                hipError_t ce = hipStreamEndCapture(s, &graph);
                if (crc != BIFHIP_OK) { rc = crc; break; }
                if (ce != hipSuccess || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) !=
                                            hipSuccess) {
                    rc = fail(1, "graph instantiation failed");
                    break;
                }
            }
            Timed t("mlp_train", s);
            if (hipGraphLaunch(exec, s) != hipSuccess) rc = fail(1, "graph launch failed");
        } else {
            Timed t("mlp_train", s);
            if (launch_fill_f32(3, N.loss, 0.0f, s) != hipSuccess) rc = fail(1, "fill");
            for (int k = 0; k < steps && rc == BIFHIP_OK; ++k)
                rc = train_step(N, T, nLat, d_perm + (size_t)k * batch, d_cls, d_lle, s);
        }
    }
    // the graph and its executable outlive their last replay
    hipError_t se = hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    RC_TRY(rc);
    HIP_TRY(se);
    if (loss_out) {
        HIP_TRY(hipMemcpy(loss_out, N.loss, sizeof(float) * 3, hipMemcpyDeviceToHost));
        if (steps > 0) {
            loss_out[0] /= steps;
            loss_out[1] /= steps;
            loss_out[2] /= steps;
        }
    }
    resolve_timers();
    return BIFHIP_OK;
}

// Surrogate inference of a whole nLat x nLat lattice: class, chaos regime,
// confidence and LLE (de-standardised with mean/sd) per cell.
int bifhip_mlp_infer_device(int nLat, double lleMean, double lleSd, double chaosThr,
                            unsigned char* d_cls, unsigned char* d_chaos, float* d_conf,
                            float* d_confChaos, float* d_lle)
{
    RC_TRY(infer_impl(nLat, lleMean, lleSd, chaosThr, d_cls, d_chaos, d_conf, d_confChaos,
                      d_lle));
    return finish();
}

// Parameters of the current network (FP32, layout W1 b1 W2 b2 W3 b3 Wh bh)
// and the Fourier matrix, for checks against a CPU reference.
int bifhip_mlp_get_params(float* params, int nParams, float* fourier, int nFourier)
{
    if (nParams != kParams || nFourier != kFeat) return fail(BIFHIP_E_ARG, "size mismatch");
    Net N;
    RC_TRY(net_alloc(N));
    HIP_TRY(hipMemcpy(params, N.p, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(fourier, N.bff, sizeof(float) * kFeat, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

// One forward + backward pass of given parameters on given lattice cells;
// gradients (same layout as the parameters) and the three losses back to
// the host.  Used to check the MFMA kernels against a CPU reference: the
// gradient and losses come from the fused training kernel (targets by row),
// the head outputs from the fused inference kernel.
int bifhip_mlp_gradient(int nLat, const float* params, const float* fourier, const int* cells,
                        const int* cls, const float* lle, int rows, double wLle, double chaosStd,
                        float* grads, float* losses, float* outputs)
{
    RC_TRY(ensure_device());
    if (rows < 1 || !params || !fourier || !cells || !cls || !lle || !grads)
        return fail(BIFHIP_E_ARG, "bad gradient arguments");
    if (train_fused_blocks(rows) > kMaxSplits)
        return fail(BIFHIP_E_ARG, "rows %d above %d", rows, 64 * kMaxSplits);
    Net N;
    RC_TRY(net_alloc(N));
    HIP_TRY(hipMemcpy(N.p, params, sizeof(float) * kParams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(N.bff, fourier, sizeof(float) * kFeat, hipMemcpyHostToDevice));
    RC_TRY(upload_weights(N, params));
    DBuf dc, dcls, dl, seq;
    HIP_TRY(dc.alloc(sizeof(int) * rows));
    HIP_TRY(dcls.alloc(sizeof(int) * rows));
    HIP_TRY(dl.alloc(sizeof(float) * rows));
    HIP_TRY(seq.alloc(sizeof(int) * rows));
    std::vector<int> iota(rows);
    std::iota(iota.begin(), iota.end(), 0);
    HIP_TRY(hipMemcpy(dc.p, cells, sizeof(int) * rows, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dcls.p, cls, sizeof(int) * rows, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dl.p, lle, sizeof(float) * rows, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(seq.p, iota.data(), sizeof(int) * rows, hipMemcpyHostToDevice));
    hipStream_t s = g.stream;
    HIP_TRY(launch_fill_f32(3, N.loss, 0.0f, s));
    HIP_TRY(launch_train_fused(rows, nLat, dc.as<int>(), seq.as<int>(), N.bff, N.wbf, N.wbfT, N.p,
                               dcls.as<int>(), dl.as<float>(), (float)wLle, (float)chaosStd,
                               N.gpart, kParams, N.loss, nullptr, s));
    HIP_TRY(launch_sum_splits(kParams, train_fused_blocks(rows), N.gpart, N.g, s));
    if (outputs) {
        float* d_out;
        RC_TRY(pool_get("gr_out", (size_t)rows * kHead, &d_out));
        HIP_TRY(launch_infer_fused(rows, 0, dc.as<int>(), nLat, N.bff, N.wbf, N.p, 0.0f, 1.0f,
                                   0.0f, nullptr, nullptr, nullptr, nullptr, nullptr, d_out, s));
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(outputs, d_out, sizeof(float) * rows * kHead, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(grads, N.g, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    if (losses) HIP_TRY(hipMemcpy(losses, N.loss, sizeof(float) * 3, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

// k_adam on caller-given state: `steps` successive Adam updates of n FP32
// parameters (host arrays params, m, v updated in place), the gradient of
// step k given as `splits` partials grads[(k * splits + z) * n + i] that the
// kernel adds in order z = 0..splits-1; keep[i] == 0 holds parameter i at
// zero.  The BF16 copy of every parameter after the last step goes to wbf
// (bit patterns) when it is not NULL.  For checks against a CPU Adam.
int bifhip_mlp_adam(int n, float* params, float* m, float* v, const float* grads, int splits,
                    int steps, const unsigned char* keep, double lr, double b1, double b2,
                    double eps, unsigned short* wbf)
{
    RC_TRY(ensure_device());
    if (n < 1 || splits < 1 || steps < 0 || !params || !m || !v || !keep || (steps && !grads))
        return fail(BIFHIP_E_ARG, "bad Adam arguments");
    DBuf dp, dm, dv, dg, dk, di, dw, ds;
    HIP_TRY(dp.alloc(sizeof(float) * n));
    HIP_TRY(dm.alloc(sizeof(float) * n));
    HIP_TRY(dv.alloc(sizeof(float) * n));
    HIP_TRY(dg.alloc(sizeof(float) * (size_t)n * splits));
    HIP_TRY(dk.alloc((size_t)n));
    HIP_TRY(di.alloc(sizeof(int) * n));
    HIP_TRY(dw.alloc(sizeof(bf16) * n));
    HIP_TRY(ds.alloc(sizeof(int)));
    std::vector<int> ident(n);
    std::iota(ident.begin(), ident.end(), 0);
    HIP_TRY(hipMemcpy(dp.p, params, sizeof(float) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dm.p, m, sizeof(float) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dv.p, v, sizeof(float) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dk.p, keep, (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di.p, ident.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    for (int k = 0; k < steps; ++k) {
        const int t = k + 1;
        HIP_TRY(hipMemcpy(dg.p, grads + (size_t)k * splits * n, sizeof(float) * (size_t)n * splits,
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ds.p, &t, sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(launch_adam(n, dp.as<float>(), dm.as<float>(), dv.as<float>(), dg.as<float>(),
                            splits, n, dk.as<unsigned char>(), di.as<int>(), dw.as<bf16>(),
                            nullptr, nullptr, (float)lr, (float)b1, (float)b2, (float)eps,
                            ds.as<int>(), g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
    }
    HIP_TRY(hipMemcpy(params, dp.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(m, dm.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(v, dv.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (wbf && steps > 0) HIP_TRY(hipMemcpy(wbf, dw.p, sizeof(bf16) * n, hipMemcpyDeviceToHost));
    return BIFHIP_OK;
}

int bifhip_mlp_param_count(void) { return kParams; }

int bifhip_surrogate_map_device(int nFine, int nCoarse, double h, int nIC, const double* ic,
                                const double* ranges, const int* mut, const double* values,
                                int nVal, double peakTMax, double peakTransient, int wv,
                                int maxPeaks, double dbEps, double lleTMax, double NT,
                                double lleTransient, double lleEps, int epochs, int batch,
                                double lr, unsigned long long seed, double confThreshold,
                                double chaosThreshold, double validationFraction, int* d_peak,
                                double* d_lle, unsigned char* d_resolved, double* stats)
{
    RC_TRY(ensure_device());
    if (nFine < 2 || nCoarse < 2 || !d_peak || !d_lle) return fail(BIFHIP_E_ARG, "bad grid");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = g.stream;
    double st[BIFHIP_SSTAT_COUNT];
    std::fill(st, st + BIFHIP_SSTAT_COUNT, 0.0);

    // 1. coarse map with the plain FP64 fused kernel
    const int nc = nCoarse * nCoarse;
    int* d_cpeak;
    double* d_clle;
    RC_TRY(pool_get("sg_cpeak", nc, &d_cpeak));
    RC_TRY(pool_get("sg_clle", nc, &d_clle));
    RC_TRY(surrogate_resolve(nCoarse, h, nIC, ic, ranges, mut, values, nVal, peakTMax,
                             peakTransient, wv, maxPeaks, dbEps, lleTMax, NT, lleTransient,
                             lleEps, nullptr, nullptr, nc, d_clle, d_cpeak));
    std::vector<int> cpeak(nc);
    std::vector<double> clle(nc);
    HIP_TRY(hipMemcpy(cpeak.data(), d_cpeak, sizeof(int) * nc, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(clle.data(), d_clle, sizeof(double) * nc, hipMemcpyDeviceToHost));
    st[BIFHIP_SSTAT_MS_COARSE] = ms_since(t0);

    // labels: class = clamp(count, 1, 16) - 1; LLE standardised over finite cells
    double sum = 0, sum2 = 0;
    long long nfin = 0;
    for (int i = 0; i < nc; ++i)
        if (std::isfinite(clle[i])) {
            sum += clle[i];
            sum2 += clle[i] * clle[i];
            ++nfin;
        }
    const double mean = nfin ? sum / nfin : 0.0;
    const double sd = nfin ? std::max(1e-6, std::sqrt(std::max(0.0, sum2 / nfin - mean * mean)))
                           : 1.0;
    std::vector<int> cls(nc);
    std::vector<float> lstd(nc);
    std::vector<unsigned char> cchaos(nc);
    for (int i = 0; i < nc; ++i) {
        cls[i] = std::min(kClasses, std::max(1, cpeak[i])) - 1;
        lstd[i] = std::isfinite(clle[i]) ? (float)((clle[i] - mean) / sd) : NAN;
        cchaos[i] = clle[i] > chaosThreshold ? 1 : 0;     // NaN: not chaotic
    }
    unsigned char* d_cchaos;
    RC_TRY(pool_get("sg_cchaos", nc, &d_cchaos));
    HIP_TRY(hipMemcpy(d_cchaos, cchaos.data(), (size_t)nc, hipMemcpyHostToDevice));
    int* d_cls;
    float* d_lstd;
    RC_TRY(pool_get("sg_cls", nc, &d_cls));
    RC_TRY(pool_get("sg_lstd", nc, &d_lstd));
    HIP_TRY(hipMemcpy(d_cls, cls.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_lstd, lstd.data(), sizeof(float) * nc, hipMemcpyHostToDevice));

    // 2. training
    const auto t1 = std::chrono::steady_clock::now();
    float losses[3] = {0, 0, 0};
    const double chaosStd = (chaosThreshold - mean) / sd;
    RC_TRY(bifhip_mlp_train_device(nCoarse, d_cls, d_lstd, epochs, std::min(batch, nc), lr, 1.0,
                                   seed, 6.0, chaosStd, g.useGraph, losses));
    st[BIFHIP_SSTAT_MS_TRAIN] = ms_since(t1);
    st[BIFHIP_SSTAT_LOSS_CE] = losses[0];
    st[BIFHIP_SSTAT_LOSS_MSE] = losses[1];

    // 3. inference
    const auto t2 = std::chrono::steady_clock::now();
    const long long nf = (long long)nFine * nFine;
    const bool partial = g.partialResolve != 0;
    unsigned char *d_pcls, *d_pch, *d_flag;
    float *d_conf, *d_cconf, *d_plle;
    RC_TRY(pool_get("sg_pcls", (size_t)nf, &d_pcls));
    RC_TRY(pool_get("sg_flag", (size_t)nf, &d_flag));
    RC_TRY(pool_get("sg_pch", (size_t)nf, &d_pch));
    RC_TRY(pool_get("sg_conf", (size_t)nf, &d_conf));
    RC_TRY(pool_get("sg_cconf", (size_t)nf, &d_cconf));
    RC_TRY(pool_get("sg_plle", (size_t)nf, &d_plle));
    RC_TRY(infer_impl(nFine, mean, sd, chaosThreshold, d_pcls, d_pch, d_conf, d_cconf, d_plle));
    HIP_TRY(hipStreamSynchronize(s));
    st[BIFHIP_SSTAT_MS_INFER] = ms_since(t2);

    // 4. flag and re-solve
    const auto t3 = std::chrono::steady_clock::now();
    // full re-solves first, then the validation sample (cells drawn on the
    // device among the unflagged ones, a seeded hash per cell), in one list
    int *d_all, *d_vlist, *d_plist, *d_llist, *d_count;
    RC_TRY(pool_get("sg_all", (size_t)nf, &d_all));
    RC_TRY(pool_get("sg_vlist", (size_t)(validationFraction > 0 ? nf : 1), &d_vlist));
    RC_TRY(pool_get("sg_plist", (size_t)(partial ? nf : 1), &d_plist));
    RC_TRY(pool_get("sg_llist", (size_t)(partial ? nf : 1), &d_llist));
    RC_TRY(pool_get("sg_count", 6, &d_count));
    HIP_TRY(launch_fill(6, d_count, 0, s));
    const double vf = std::min(1.0, std::max(0.0, validationFraction));
    const unsigned int vthr =
        vf > 0 ? (unsigned int)std::min(4294967295.0, std::floor(vf * 4294967296.0)) : 0u;
    HIP_TRY(launch_flag(nFine, d_pcls, d_pch, d_conf, d_cconf, (float)confThreshold,
                        (float)g.chaosConfidence, partial, g.reading,
                        g.quadConfidence > 0 ? d_cchaos : nullptr, nCoarse,
                        (float)g.quadConfidence, d_all, d_count, d_plist,
                        d_count + 1, d_llist, d_count + 3, seed + 17, vthr, d_vlist, d_count + 2,
                        d_flag, s));
    HIP_TRY(launch_fill_pred(nf, d_pcls, d_plle, d_peak, d_lle, s));
    int counts[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(counts, d_count, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int nflag = counts[0], npeak = counts[1], nval = counts[2], nlle = counts[3];
    const int ntot = nflag + nval;
    if (nval)
        HIP_TRY(hipMemcpyAsync(d_all + nflag, d_vlist, sizeof(int) * nval,
                               hipMemcpyDeviceToDevice, s));
    int *d_rpeak, *d_ppeak;
    double *d_rlle, *d_llle;
    RC_TRY(pool_get("sg_rpeak", (size_t)std::max(1, ntot), &d_rpeak));
    RC_TRY(pool_get("sg_rlle", (size_t)std::max(1, ntot), &d_rlle));
    RC_TRY(pool_get("sg_ppeak", (size_t)std::max(1, npeak), &d_ppeak));
    RC_TRY(pool_get("sg_llle", (size_t)std::max(1, nlle), &d_llle));
    HIP_TRY(hipStreamSynchronize(s));
    st[BIFHIP_SSTAT_MS_FLAG] = ms_since(t3);
    const auto t4 = std::chrono::steady_clock::now();
    if (ntot > 0)
        RC_TRY(surrogate_resolve(nFine, h, nIC, ic, ranges, mut, values, nVal, peakTMax,
                                 peakTransient, wv, maxPeaks, dbEps, lleTMax, NT, lleTransient,
                                 lleEps, nullptr, d_all, ntot, d_rlle, d_rpeak));
    // the class-only peak counts: their last clustering pass overlaps the
    // regime-only LLE integration enqueued next
    if (npeak > 0)
        RC_TRY(surrogate_resolve_peaks_begin(nFine, h, nIC, ic, ranges, mut, values, nVal,
                                             peakTMax, peakTransient, wv, maxPeaks, dbEps,
                                             d_plist, npeak, d_ppeak));
    int nchg = 0, nnear = nlle;
    int *d_clist = nullptr, *d_cpeak2 = nullptr;
    const bool lle32 = g.lle32Margin >= 0 && g.model != BIFHIP_MODEL_MEMRISTIVE;
    if (nlle > 0 && lle32) {
        // FP32 bulk LLE of every regime-only cell; the cells it leaves within
        // the margin of the threshold are re-solved in FP64 and put back in
        // their list positions
        RC_TRY(surrogate_resolve_lle(nFine, h, nIC, ic, ranges, mut, values, nVal, lleTMax, NT,
                                     lleTransient, lleEps, d_llist, nlle, d_llle,
                                     BIFHIP_PREC_FP32));
        int *d_ncell, *d_npos;
        double* d_nlle;
        RC_TRY(pool_get("sg_ncell", (size_t)nlle, &d_ncell));
        RC_TRY(pool_get("sg_npos", (size_t)nlle, &d_npos));
        HIP_TRY(launch_lle32_split(nlle, d_llist, d_llle, chaosThreshold, g.lle32Margin, d_ncell,
                                   d_npos, d_count + 5, d_flag, s));
        HIP_TRY(hipMemcpyAsync(&nnear, d_count + 5, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (nnear > 0) {
            RC_TRY(pool_get("sg_nlle", (size_t)nnear, &d_nlle));
            RC_TRY(surrogate_resolve_lle(nFine, h, nIC, ic, ranges, mut, values, nVal, lleTMax,
                                         NT, lleTransient, lleEps, d_ncell, nnear, d_nlle,
                                         BIFHIP_PREC_FP64));
            HIP_TRY(launch_scatter(nnear, d_npos, nullptr, d_nlle, nullptr, d_llle, s));
        }
    } else if (nlle > 0) {
        RC_TRY(surrogate_resolve_lle(nFine, h, nIC, ic, ranges, mut, values, nVal, lleTMax, NT,
                                     lleTransient, lleEps, d_llist, nlle, d_llle,
                                     BIFHIP_PREC_FP64));
    }
    RC_TRY(surrogate_resolve_peaks_finish());
    if (nlle > 0 && g.reading != BIFHIP_READING_CHAOS) {
        // a regime that changed under the re-solve: the peak count as well
        // (not under the chaos-only reading, whose peak counts stay predicted)
        RC_TRY(pool_get("sg_clist", (size_t)nlle, &d_clist));
        HIP_TRY(launch_regime_changed(nlle, d_llist, d_llle, d_pch, chaosThreshold, d_clist,
                                      d_count + 4, d_flag, s));
        HIP_TRY(hipMemcpyAsync(&nchg, d_count + 4, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (nchg > 0) {
            RC_TRY(pool_get("sg_cpeak2", (size_t)nchg, &d_cpeak2));
            RC_TRY(surrogate_resolve_peaks(nFine, h, nIC, ic, ranges, mut, values, nVal,
                                           peakTMax, peakTransient, wv, maxPeaks, dbEps, nullptr,
                                           d_clist, nchg, d_cpeak2));
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    st[BIFHIP_SSTAT_MS_SOLVE] = ms_since(t4);
    const auto t5 = std::chrono::steady_clock::now();
    // score the surrogate on every re-solved cell (on the device) before
    // the re-solved values overwrite the predictions
    double* d_acc;
    RC_TRY(pool_get("sg_acc", 20, &d_acc));
    HIP_TRY(launch_fill_f64(20, d_acc, 0.0, s));
    HIP_TRY(launch_score(nflag, ntot, d_all, d_rpeak, d_rlle, npeak, d_plist, d_ppeak, nlle,
                         d_llist, d_llle, d_pcls, d_pch, d_plle, chaosThreshold, d_acc, s));
    double acc[20];
    HIP_TRY(hipMemcpyAsync(acc, d_acc, sizeof(acc), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    auto score = [&](int gi, double* clsAcc, double* regAcc, double* mae) {
        const double* a = acc + gi * 5;
        *clsAcc = a[0] > 0 ? a[1] / a[0] : 1.0;
        *regAcc = a[0] > 0 ? a[2] / a[0] : 1.0;
        *mae = a[3] > 0 ? a[4] / a[3] : 0.0;
    };
    score(0, &st[BIFHIP_SSTAT_FLAGGED_CLASS_ACC], &st[BIFHIP_SSTAT_FLAGGED_REGIME_ACC],
          &st[BIFHIP_SSTAT_FLAGGED_LLE_MAE]);
    score(1, &st[BIFHIP_SSTAT_VALID_CLASS_ACC], &st[BIFHIP_SSTAT_VALID_REGIME_ACC],
          &st[BIFHIP_SSTAT_VALID_LLE_MAE]);
    st[BIFHIP_SSTAT_PEAK_ONLY_CLASS_ACC] = acc[10] > 0 ? acc[11] / acc[10] : 1.0;
    st[BIFHIP_SSTAT_LLE_ONLY_REGIME_ACC] = acc[15] > 0 ? acc[17] / acc[15] : 1.0;
    HIP_TRY(launch_scatter(ntot, d_all, d_rpeak, d_rlle, d_peak, d_lle, s));
    HIP_TRY(launch_scatter(npeak, d_plist, d_ppeak, nullptr, d_peak, d_lle, s));
    HIP_TRY(launch_scatter(nlle, d_llist, nullptr, d_llle, nullptr, d_lle, s));
    HIP_TRY(launch_scatter(nchg, d_clist, d_cpeak2, nullptr, d_peak, nullptr, s));
    // 1 re-solved in full, 2 peak count only, 3 validation sample, 4 LLE
    // only, 5 LLE then peak count (regime changed), 6 and 7 as 4 and 5 with
    // the FP32 bulk LLE, 0 predicted
    if (d_resolved)
        HIP_TRY(hipMemcpyAsync(d_resolved, d_flag, (size_t)nf, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    st[BIFHIP_SSTAT_MS_RESOLVE] = ms_since(t3);
    st[BIFHIP_SSTAT_MS_SCORE] = ms_since(t5);
    st[BIFHIP_SSTAT_RESOLVED] = nflag;
    st[BIFHIP_SSTAT_VALIDATION] = nval;
    st[BIFHIP_SSTAT_RESOLVED_FRAC] = (double)nflag / (double)nf;
    st[BIFHIP_SSTAT_PEAK_ONLY] = npeak;
    st[BIFHIP_SSTAT_LLE_ONLY] = nlle;
    st[BIFHIP_SSTAT_LLE_ONLY_RECOUNTED] = nchg;
    st[BIFHIP_SSTAT_LLE_ONLY_FP32] = nlle - nnear;
    st[BIFHIP_SSTAT_MS_TOTAL] = ms_since(t0);
    st[BIFHIP_SSTAT_LLE_MEAN] = mean;
    st[BIFHIP_SSTAT_LLE_SD] = sd;
    if (stats) std::memcpy(stats, st, sizeof(st));
    resolve_timers();
    return BIFHIP_OK;
}

}  // extern "C"
