// sortArrayNoAlloc: sortArray for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// No counterpart in the reference.  The sort runs exactly the kernels and
// passes sortArray runs, on the engine's stream; only the scratch differs.
// Instead of the engine's grow-only buffers it uses `scratch`, a fixed array
// in this library's own device code: digit histograms, digit starts, tile
// counters, per-segment counts and the lookback words of up to
// SORT_NOALLOC_MAX_ELEMENTS keys.  Each GPU the library runs on holds its own
// copy until the process ends; every call on a GPU shares it, so calls that
// overlap on one GPU race (one at a time per device).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>

#include "engine_internal.h"

namespace {

// Layout of `scratch`, in 32-bit words: histograms and digit starts of up to
// MGPU_MAX_SEGS segments, the tile counters (padded to 256 words), the
// per-segment counts of each pass, then the lookback words: 4 passes of
// 256 words per tile for SORT_NOALLOC_MAX_ELEMENTS keys in tiles of 8192 (the
// smaller of the engine's two tile shapes).
constexpr size_t kHistOff = 0;
constexpr size_t kStartOff = kHistOff + kLaneHistWords;
constexpr size_t kCounterOff = kStartOff + kLaneHistWords;
constexpr size_t kSegCountOff = kCounterOff + 256;
constexpr size_t kLookbackOff = kSegCountOff + kLaneSegCountWords;
constexpr size_t kLookbackWords =
    (size_t)MGPU_MAX_PASSES * ((size_t)SORT_NOALLOC_MAX_ELEMENTS / 8192) * MGPU_RADIX;
constexpr size_t kScratchWords = kLookbackOff + kLookbackWords;
static_assert(MGPU_LANE_COUNTERS <= 256, "tile counters exceed their slot");
static_assert(kLookbackOff % 64 == 0, "lookback words start on a 256-byte boundary");

static __device__ __align__(256) uint32_t scratch[kScratchWords];

// Zeroes the words the count kernel accumulates into (histograms and
// counters) after a call that may have stopped part way.
__global__ void __launch_bounds__(256) mgpu_noalloc_clear(uint32_t* p, uint32_t words) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < words; i += gridDim.x * 256) p[i] = 0;
}

constexpr int kMaxDevices = 64;
std::mutex g_setupMutex;
bool g_setupDone[kMaxDevices] = {};
uint32_t* g_scratchPtr[kMaxDevices] = {};       // device address of `scratch` per GPU
std::atomic<bool> g_dirty[kMaxDevices];          // histograms may hold counts

}  // namespace

// Called by sortCreateEngine with `device` current: looks up this GPU's copy
// of `scratch` once (hipGetSymbolAddress; the array's name is only a host-side
// stand-in, never a device address).  Returns false if the lookup failed, in
// which case sortArrayNoAlloc reports SORT_STATUS_DEVICE_ERROR on this GPU.
bool mgpu_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> g(g_setupMutex);
    if (!g_setupDone[device]) {
        g_setupDone[device] = true;
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess && ptr)
            g_scratchPtr[device] = static_cast<uint32_t*>(ptr);
        else
            (void)hipGetLastError();
        g_dirty[device].store(true);  // cleared before the first sort
    }
    return g_scratchPtr[device] != nullptr;
}

extern "C" {

sortStatus_t SORTAPI sortArrayNoAlloc(sortEngine_t engine, sortData_t data) {
    if (!engine || !data) return SORT_STATUS_INVALID_VALUE;
    if (mgpu_lane_take_error(engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    const int numBits = data->endBit - data->firstBit;
    if (!numBits) return SORT_STATUS_SUCCESS;
    sortStatus_t st = mgpu_validate_data(engine, data);
    if (st) return st;
    if (data->numElements > SORT_NOALLOC_MAX_ELEMENTS) return SORT_STATUS_CONFIG_NOT_SUPPORTED;
    const int dev = engine->device;
    uint32_t* const ptr = dev >= 0 && dev < kMaxDevices ? g_scratchPtr[dev] : nullptr;
    if (!ptr) return SORT_STATUS_DEVICE_ERROR;

    // A view of `scratch` shaped like the engine's own scratch.
    SortLane v;
    v.stream = engine->lane.stream;
    v.hist = ptr + kHistOff;
    v.digitStart = ptr + kStartOff;
    v.tileCounters = ptr + kCounterOff;
    v.segCount = ptr + kSegCountOff;
    v.lookback = ptr + kLookbackOff;
    v.lookbackWords = kLookbackWords;
    v.errorWord = engine->lane.errorWord;
    v.errorDev = engine->lane.errorDev;
    v.noAlloc = true;
    if (g_dirty[dev].load()) {
        // A fill kernel, not a memset: callers may capture this in a graph.
        const uint32_t words = (uint32_t)(kSegCountOff - kHistOff);
        hipLaunchKernelGGL(mgpu_noalloc_clear, dim3(64), dim3(256), 0, v.stream, v.hist, words);
        if (hipGetLastError() != hipSuccess) return SORT_STATUS_LAUNCH_ERROR;
        g_dirty[dev].store(false);
    }
    int widths[MGPU_MAX_PASSES];
    const int passes = mgpu_plan_passes(numBits, MGPU_RADIX_BITS, widths);
    st = mgpu_sort_data(engine, v, data, passes, widths);
    // The count kernel leaves histograms and counters for the scan to clear;
    // a sort that stopped in between leaves them dirty for the next call.
    if (v.histDirty) g_dirty[dev].store(true);
    return st;
}

}  // extern "C"
