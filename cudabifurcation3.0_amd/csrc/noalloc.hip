// noalloc.hip -- dbscanCUDA over device peak rows without any allocation.
//
// bifhip_dbscan_device_noalloc is bifhip_dbscan_device with every scratch
// buffer it needs -- the overflow list of cells the wave kernel leaves to
// the workgroup kernel, the list of cells beyond the workgroup kernel's
// 16 384 peaks, their counters and the lane kernel's label rows -- taken
// from one file-scope device array instead of the library's workspace
// pool.  It is for callers that cannot call hipMalloc (tight loops, code
// that forbids device allocation).  The kernels are the ones
// bifhip_dbscan_device launches, so the results are identical.
//
// Each GPU holds its own copy of `scratch` until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (the array's name in host code is a host-side
// stand-in, not a device address).  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bifhip.h"
#include "host_internal.h"

namespace bifhip {
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
hipError_t launch_fill(long long n, int* v, int value, hipStream_t st);
namespace host {
double dbscan_threshold(double eps);
}
}  // namespace bifhip

namespace {

// Layout of the array, in ints:
//   [0]            overflow count        [1]  count of cells beyond kBlkMaxP
//   [kHdr, +kMax)  overflow list (cells the wave kernel does not hold)
//   [.., +kMax)    cells of the overflow list beyond the workgroup kernel
//   [.., end)      lane-kernel label rows, 2 * sizeOfBlock ints per cell
constexpr long long kHdr = 64;
constexpr long long kMax = BIFHIP_NOALLOC_MAX_BLOCKS;
constexpr long long kLaneInts = 2LL * BIFHIP_NOALLOC_MAX_ROW;
constexpr long long kScratchInts = kHdr + 2 * kMax + kLaneInts;
static_assert(kScratchInts * 4 == BIFHIP_NOALLOC_SCRATCH_BYTES, "scratch size in bifhip.h");

}  // namespace

static __device__ __align__(256) int scratch[kScratchInts];

namespace {

constexpr int kMaxDevices = 64;
int* g_scratchAddr[kMaxDevices];     // per device, from hipGetSymbolAddress

// The address of the current device's array, looked up on the first call
// made with that device current.
int scratch_addr(int** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return bifhip::host::fail(BIFHIP_E_LIMIT, "device %d beyond the %d devices supported", dev,
                                  kMaxDevices);
    if (!g_scratchAddr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchAddr[dev] = static_cast<int*>(p);
    }
    *out = g_scratchAddr[dev];
    return BIFHIP_OK;
}

// Entries of list[0..n) whose cells hold more than minP peaks, appended to
// sel (order irrelevant: each entry writes its own cell).
__global__ void k_select_over(int n, const int* __restrict__ list, const int* __restrict__ counts,
                              int minP, int* __restrict__ sel, int* __restrict__ nSel)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int cell = list[t];
    if (counts[cell] > minP) sel[atomicAdd(nSel, 1)] = cell;
}

}  // namespace

using namespace bifhip;
using namespace bifhip::host;

extern "C" int bifhip_dbscan_device_noalloc(const double* d_peaks, const double* d_intervals,
                                            int sizeOfBlock, int nBlocks, const int* d_counts,
                                            int maxPeaks, double eps, int* d_out)
{
    RC_TRY(ensure_device());
    if (!d_peaks || !d_intervals || !d_counts || !d_out || sizeOfBlock < 1 || nBlocks < 0)
        return fail(BIFHIP_E_ARG, "bad dbscan arguments");
    if (nBlocks > kMax)
        return fail(BIFHIP_E_LIMIT,
                    "bifhip_dbscan_device_noalloc: %d blocks exceed its limit of %lld per call",
                    nBlocks, kMax);
    const int blkMax = dbscan_block_max_peaks();
    const bool laneCells = sizeOfBlock > blkMax;   // rows that can hold more than blkMax
    if (sizeOfBlock > BIFHIP_NOALLOC_MAX_ROW)
        return fail(BIFHIP_E_LIMIT,
                    "bifhip_dbscan_device_noalloc: rows of %d peaks exceed its limit of %d",
                    sizeOfBlock, BIFHIP_NOALLOC_MAX_ROW);
    if (nBlocks == 0) return finish();
    int* base = nullptr;
    RC_TRY(scratch_addr(&base));
    int* d_novf = base;
    int* d_nbig = base + 1;
    int* d_ovf = base + kHdr;
    int* d_big = d_ovf + kMax;
    int* d_lane = d_big + kMax;
    const double T = dbscan_threshold(eps);
    // the array keeps the previous call's contents: clear both counters
    HIP_TRY(launch_fill(2, d_novf, 0, g.stream));
    HIP_TRY(launch_dbscan_wave(nBlocks, d_counts, d_peaks, d_intervals, sizeOfBlock, sizeOfBlock,
                               maxPeaks, T, d_out, d_ovf, d_novf, 0, g.stream));
    int novf = 0;
    HIP_TRY(hipMemcpyAsync(&novf, d_novf, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (novf > 0) {
        // cells up to blkMax peaks; larger ones get -1 here and are
        // overwritten by the lane kernel below
        HIP_TRY(launch_dbscan_block(novf, d_ovf, d_counts, d_peaks, d_intervals, sizeOfBlock, T,
                                    d_out, nullptr, g.stream));
        if (laneCells) {
            hipLaunchKernelGGL(k_select_over, dim3((novf + 255) / 256), dim3(256), 0, g.stream,
                               novf, d_ovf, d_counts, blkMax, d_big, d_nbig);
            HIP_TRY(hipGetLastError());
            int nbig = 0;
            HIP_TRY(hipMemcpyAsync(&nbig, d_nbig, sizeof(int), hipMemcpyDeviceToHost, g.stream));
            HIP_TRY(hipStreamSynchronize(g.stream));
            // as many cells per launch as the label rows hold; the launches
            // run in stream order, so each reuses the rows
            const long long stride = 2LL * sizeOfBlock;
            const int per = (int)std::max(1LL, kLaneInts / stride);
            for (int off = 0; off < nbig; off += per) {
                const int m = std::min(per, nbig - off);
                HIP_TRY(launch_dbscan_lane(m, d_big + off, d_counts, d_peaks, d_intervals,
                                           sizeOfBlock, T, d_lane, stride, d_out, nullptr,
                                           g.stream));
            }
        }
    }
    return finish();
}
