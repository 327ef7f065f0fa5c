// libmgpusort engine: the C ABI of include/mgpusort.h over the HIP module API.
//
// Reference: sort/src/mgpusort/mgpusort.cpp (engine, pass loop, data
// management) and sortdll.h (engine state).  What is kept from the reference
// and why is listed beside each function; what differs:
//   * passes are 8-bit digits (4 passes for 32-bit keys) chosen by
//     sortPlanPasses instead of the Fermi tables of sorttables.cpp;
//   * every pass is one rank_scatter launch after a single digit_count +
//     digit_scan per sort (the reference runs count + 3 hist + sort per pass,
//     mgpusort.cpp:424-516);
//   * keys past numElements are never read or written, so there is no end-key
//     save/set/restore (mgpusort.cpp:402-419,591-598);
//   * pass errors are returned instead of swallowed (mgpusort.cpp:583,600).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mgpusort.h"
#include "kernel_args.h"
#include "engine_internal.h"

namespace {

const char* const kStatusStrings[] = {
    "SORT_STATUS_SUCCESS",          "SORT_STATUS_NOT_INITIALIZED",
    "SORT_STATUS_DEVICE_ALLOC_FAILED", "SORT_STATUS_HOST_ALLOC_FAILED",
    "SORT_STATUS_CONFIG_NOT_SUPPORTED", "SORT_STATUS_CONTEXT_MISMATCH",
    "SORT_STATUS_INVALID_CONTEXT",  "SORT_STATUS_KERNEL_NOT_FOUND",
    "SORT_STATUS_KERNEL_ERROR",     "SORT_STATUS_LAUNCH_ERROR",
    "SORT_STATUS_INVALID_VALUE",    "SORT_STATUS_DEVICE_ERROR",
    "SORT_STATUS_INTERNAL_ERROR",   "SORT_STATUS_UNSUPPORTED_DEVICE"};

constexpr int kAllocRound = 2048;  // buffer rounding of mgpusort.cpp:663,696

inline int round_up(int x, int r) { return ((x + r - 1) / r) * r; }

std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&sortCreateEngine), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.rfind('/');
        if (s != std::string::npos) return p.substr(0, s + 1);
    }
    return std::string("./");
}

// Scratch that grows only (AllocSortResources, mgpusort.cpp:319-343).
sortStatus_t ensure_lookback(SortLane& L, uint64_t words) {
    if (words <= L.lookbackWords) return SORT_STATUS_SUCCESS;
    if (L.lookback) {
        if (hipStreamSynchronize(L.stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
        (void)hipFree(L.lookback);
        L.lookback = nullptr;
        L.lookbackWords = 0;
    }
    uint64_t alloc = words + (words >> 3) + 1024;  // headroom against regrowth
    if (hipMalloc(&L.lookback, alloc * sizeof(uint32_t)) != hipSuccess) {
        L.lookback = nullptr;
        return SORT_STATUS_DEVICE_ALLOC_FAILED;
    }
    L.lookbackWords = alloc & ~uint64_t(3);
    return SORT_STATUS_SUCCESS;
}

}  // namespace

const PassVariant& mgpu_pick_variant(const sortEngine_d* e, uint64_t keys) {
    const uint64_t largeTiles = (keys + e->large.tile() - 1) / e->large.tile();
    return largeTiles * kCUsPerLargeTile >= (uint64_t)e->numCUs ? e->large : e->small;
}

SortPlanChoice mgpu_pick_plan(const sortEngine_d* e, uint32_t n, int valueCount) {
    const int cls = valueCount == 0    ? MGPU_SORT_CLASS_KEYS
                    : valueCount == -1 ? MGPU_SORT_CLASS_INDEX
                    : valueCount == 1  ? MGPU_SORT_CLASS_KV
                                       : MGPU_SORT_CLASS_MULTI;
    int lg = 0;
    while (lg < 31 && (1u << lg) < n) ++lg;
    const sortEngine_d::SortTuneEntry* entry = nullptr;
    for (const auto& en : e->sortTable[cls])
        if (lg <= en.maxLog2) {
            entry = &en;
            break;
        }
    SortPlanChoice c;
    c.variant = entry ? (entry->large ? &e->large : &e->small) : &mgpu_pick_variant(e, n);
    const uint32_t numTiles = (n + c.variant->tile() - 1) / c.variant->tile();
    // Segments need a tile per XCD; from kXsegMinTilesPerSeg tiles per XCD
    // they win (measured), unless the table or MGPUSORT_XCD=1 says otherwise.
    bool xcd;
    if (e->xcdForce) xcd = true;
    else if (entry) xcd = entry->xcd;
    else xcd = numTiles >= kXsegMinTilesPerSeg * e->xcdSegs;
    c.segs = e->xcdSegs > 1 && xcd && numTiles >= e->xcdSegs ? e->xcdSegs : 1u;
    return c;
}

void mgpu_load_sort_table(sortEngine_d* e) {
    for (auto& t : e->sortTable) t.clear();
    std::ifstream f(e->kernelPath + "mgpusort_sort_tuning.txt");
    std::string line;
    while (f && std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        int maxLog2 = 0;
        std::string values, tile, chains;
        if (!(is >> maxLog2 >> values >> tile >> chains)) continue;
        const int cls = values == "keys" ? MGPU_SORT_CLASS_KEYS
                        : values == "index" ? MGPU_SORT_CLASS_INDEX
                        : values == "kv" ? MGPU_SORT_CLASS_KV
                        : values == "multi" ? MGPU_SORT_CLASS_MULTI
                                            : -1;
        if (cls < 0 || maxLog2 < 0 || (tile != "small" && tile != "large") ||
            (chains != "one" && chains != "xcd"))
            continue;
        e->sortTable[cls].push_back({maxLog2, tile == "large", chains == "xcd"});
    }
    for (auto& t : e->sortTable) {
        std::stable_sort(t.begin(), t.end(),
                         [](const sortEngine_d::SortTuneEntry& a,
                            const sortEngine_d::SortTuneEntry& b) { return a.maxLog2 < b.maxLog2; });
        if (!t.empty()) t.back().maxLog2 = 31;
    }
}

sortStatus_t mgpu_ensure_lookback(sortEngine_d* e, SortLane& lane, uint32_t n, int passes) {
    // Enough for either tile shape.
    const uint32_t tile = std::min(e->small.tile(), e->large.tile());
    const uint64_t tiles = (n + tile - 1) / tile;
    return ensure_lookback(lane, (uint64_t)passes * tiles * MGPU_RADIX);
}

sortStatus_t mgpu_ensure_lookback_words(SortLane& lane, uint64_t words) {
    return ensure_lookback(lane, words);
}

sortStatus_t mgpu_ensure_multi(SortLane& L, uint32_t arrays) {
    if (arrays <= L.multiArrays && !L.multiDirty) return SORT_STATUS_SUCCESS;
    if (hipStreamSynchronize(L.stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
    if (arrays > L.multiArrays) {
        if (L.multi) (void)hipFree(L.multi);
        L.multi = nullptr;
        L.multiArrays = 0;
        const uint32_t cap = arrays + arrays / 2 + 16;
        if (hipMalloc(&L.multi, (size_t)cap * 2 * MGPU_MAX_PASSES * MGPU_RADIX * sizeof(uint32_t)) !=
            hipSuccess) {
            L.multi = nullptr;
            return SORT_STATUS_DEVICE_ALLOC_FAILED;
        }
        L.multiArrays = cap;
    }
    // Stream-ordered: a plain hipMemset runs on the null stream, which a
    // non-blocking lane stream does not wait for.
    if (hipMemsetAsync(L.multi, 0,
                       (size_t)L.multiArrays * MGPU_MAX_PASSES * MGPU_RADIX * sizeof(uint32_t),
                       L.stream) != hipSuccess)
        return SORT_STATUS_DEVICE_ERROR;
    L.multiDirty = false;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t mgpu_lane_init(SortLane* l) {
    // The error word: coherent pinned memory, so a kernel's store reaches the
    // host without a copy and nothing on the normal path pays for it.
    void* word = nullptr;
    if (hipHostMalloc(&word, kLanePinnedBytes, hipHostMallocMapped | hipHostMallocCoherent) !=
        hipSuccess) {
        l->errorWord = nullptr;
        return SORT_STATUS_HOST_ALLOC_FAILED;
    }
    l->errorWord = static_cast<uint32_t*>(word);
    std::memset(word, 0, kLanePinnedBytes);
    void* dev = nullptr;
    if (hipHostGetDevicePointer(&dev, word, 0) != hipSuccess || !dev)
        return SORT_STATUS_DEVICE_ERROR;
    l->errorDev = static_cast<uint32_t*>(dev);
    l->reportWord = l->errorWord + kLaneReportOffset;
    l->reportDev = l->errorDev + kLaneReportOffset;
    const size_t words = 2 * kLaneHistWords + MGPU_LANE_COUNTERS + kLaneSegCountWords;
    if (hipMalloc(&l->fixed, words * sizeof(uint32_t)) != hipSuccess) {
        l->fixed = nullptr;
        return SORT_STATUS_DEVICE_ALLOC_FAILED;
    }
    // On the lane's own stream, finished before the lane is handed out (the
    // null stream would not order it before a non-blocking lane stream).
    if (hipMemsetAsync(l->fixed, 0, words * sizeof(uint32_t), l->stream) != hipSuccess ||
        hipStreamSynchronize(l->stream) != hipSuccess)
        return SORT_STATUS_DEVICE_ERROR;
    l->hist = l->fixed;
    l->digitStart = l->hist + kLaneHistWords;
    l->tileCounters = l->digitStart + kLaneHistWords;
    l->segCount = l->tileCounters + MGPU_LANE_COUNTERS;
    static_assert(MGPU_LANE_COUNTERS % 2 == 0 && kLaneHistWords % 2 == 0,
                  "segCount takes 64-bit atomics: it starts on an 8-byte boundary");
    l->histDirty = false;
    return SORT_STATUS_SUCCESS;
}

void mgpu_lane_free(SortLane* l) {
    if (l->side) (void)hipStreamSynchronize(l->side);
    if (l->fork) (void)hipEventDestroy(l->fork);
    if (l->join) (void)hipEventDestroy(l->join);
    if (l->side) (void)hipStreamDestroy(l->side);
    l->fork = l->join = nullptr;
    l->side = nullptr;
    if (l->streamSwitch) (void)hipEventDestroy(l->streamSwitch);
    l->streamSwitch = nullptr;
    if (l->multi) (void)hipFree(l->multi);
    l->multi = nullptr;
    l->multiArrays = 0;
    if (l->lookback) (void)hipFree(l->lookback);
    if (l->fixed) (void)hipFree(l->fixed);
    l->lookback = nullptr;
    l->lookbackWords = 0;
    l->fixed = nullptr;
    if (l->errorWord) (void)hipHostFree(l->errorWord);
    l->errorWord = l->errorDev = l->reportWord = l->reportDev = nullptr;
}

uint32_t mgpu_lane_peek_error(const SortLane& l) {
    return l.errorWord ? __atomic_load_n(l.errorWord, __ATOMIC_ACQUIRE) : 0u;
}

sortStatus_t mgpu_lane_take_error(SortLane& l) {
    if (!l.errorWord || __atomic_load_n(l.errorWord, __ATOMIC_ACQUIRE) == 0)
        return SORT_STATUS_SUCCESS;
    return __atomic_exchange_n(l.errorWord, 0u, __ATOMIC_ACQ_REL) ? SORT_STATUS_KERNEL_ERROR
                                                                  : SORT_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Internal helpers shared with the segmented path.

int mgpu_plan_passes(int numBits, int maxWidth, int* widths) {
    if (numBits <= 0) return 0;
    if (maxWidth < 1) maxWidth = 1;
    if (maxWidth > MGPU_RADIX_BITS) maxWidth = MGPU_RADIX_BITS;
    int passes = (numBits + maxWidth - 1) / maxWidth;
    int base = numBits / passes, extra = numBits % passes;
    // Narrower passes first, as sortArrayEx orders them (mgpusort.cpp:638-642).
    for (int p = 0; p < passes; ++p)
        if (widths) widths[p] = base + (p >= passes - extra ? 1 : 0);
    return passes;
}

namespace {

hipEvent_t pooled_event(sortEngine_d* e) {
    if (!e->eventPool.empty()) {
        hipEvent_t ev = e->eventPool.back();
        e->eventPool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

}  // namespace

sortStatus_t mgpu_launch(sortEngine_d* e, hipStream_t stream, hipFunction_t f, unsigned grid,
                         unsigned block, void* args, bool timed) {
    void* params[] = {args};
    hipEvent_t start = nullptr, stop = nullptr;
    timed = timed && e->timing;
    if (timed) {
        start = pooled_event(e);
        stop = pooled_event(e);
        if (!start || !stop || hipEventRecord(start, stream) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
    }
    hipError_t r = hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, stream, params, nullptr);
    if (r != hipSuccess) return SORT_STATUS_LAUNCH_ERROR;
    if (timed) {
        if (hipEventRecord(stop, stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
        e->timed.push_back({f, start, stop});
    }
    return SORT_STATUS_SUCCESS;
}

// key_transform of n keys in place on `stream` (type MGPU_KEYS_*).
static sortStatus_t mgpu_key_transform_launch(sortEngine_d* e, hipStream_t stream,
                                              uint32_t* keys, uint32_t n, uint32_t type,
                                              bool toRadix) {
    mgpuKeyTransformArgs a;
    std::memset(&a, 0, sizeof(a));
    a.keys = keys;
    a.n = n;
    a.type = type;
    a.to_radix = toRadix ? 1u : 0u;
    const uint32_t grid = std::min<uint32_t>((n + 1023) / 1024, (uint32_t)e->numCUs * 8);
    return mgpu_launch(e, stream, e->fnKeyTransform, grid, 256, &a);
}

// One sort of n keys by [firstBit, firstBit + sum(widths)) from in-buffers to
// out-buffers alternating per pass: pass p reads keys[p&1]-side pointers.
// Pointers are passed per pass by the caller through `bufs`.
sortStatus_t mgpu_sort_passes(sortEngine_d* e, SortLane& L, uint32_t n, int firstBit,
                              int numPasses, const int* widths, int valueCount,
                              uint32_t* const keyBuf[2],
                              uint32_t* const valBuf[MGPU_MAX_VALUES][2], bool earlyExit,
                              int* executed, int keyType) {
    int done = 0;
    if (executed) *executed = 0;
    if (n == 0 || numPasses == 0) return SORT_STATUS_SUCCESS;
    if (numPasses > MGPU_MAX_PASSES) return SORT_STATUS_INTERNAL_ERROR;
    // Signed and float keys: the kernels' key type, and the count's form of
    // the mapping to the unsigned order (mgpu_key_map's two words).
    uint32_t kernelKeyType = 0, toLo = 0, toHi = 0;
    if (keyType == SORT_KEYS_INT32) {
        kernelKeyType = MGPU_KEYS_INT32;
        toLo = toHi = 0x80000000u;
    } else if (keyType == SORT_KEYS_FLOAT32) {
        kernelKeyType = MGPU_KEYS_FLOAT32;
        toLo = 0x80000000u;
        toHi = 0xFFFFFFFFu;
    } else if (keyType != SORT_KEYS_UINT32) {
        return SORT_STATUS_INVALID_VALUE;
    }
    const SortPlanChoice choice = mgpu_pick_plan(e, n, valueCount);
    const PassVariant& var = *choice.variant;
    const uint32_t tile = var.tile();
    const uint32_t numTiles = (n + tile - 1) / tile;
    const uint64_t lbWords = (uint64_t)numPasses * numTiles * MGPU_RADIX;
    sortStatus_t st = SORT_STATUS_SUCCESS;
    const bool timed = !L.noAlloc;
    // Typed keys are mapped inside the count and the first and last executed
    // passes where the tile shape has those forms (atomic-add ranking), else
    // in place by key_transform before the count and after the last pass.
    const uint32_t typedKeys = kernelKeyType;
    const bool fused = typedKeys != 0 && var.passKM[0] != nullptr && var.xsegKM[0] != nullptr &&
                       e->fnCountKM != nullptr && e->fnCountColsKM != nullptr;
    if (typedKeys != 0 && !fused) {
        if (L.noAlloc) return SORT_STATUS_CONFIG_NOT_SUPPORTED;
        st = mgpu_key_transform_launch(e, L.stream, keyBuf[0], n, typedKeys, true);
        if (st) return st;
        kernelKeyType = 0;
    }
    if (L.noAlloc) {
        // Scratch of a fixed size, filled by nothing that allocates or waits.
        if (lbWords > L.lookbackWords) return SORT_STATUS_CONFIG_NOT_SUPPORTED;
        if (L.histDirty) return SORT_STATUS_INTERNAL_ERROR;
        earlyExit = false;
    } else {
        st = ensure_lookback(L, lbWords);
        if (st) return st;
    }
    // One segment per XCD (each with a lookback chain in its XCD's L2) for
    // large sorts; else one chain across the device.  digit_count counts the
    // segments of the input order, which only the first executed pass reads;
    // each pass counts the next one's segments as it writes.
    const uint32_t segs = choice.segs;

    if (L.histDirty) {
        if (hipMemsetAsync(L.hist, 0, kLaneHistWords * sizeof(uint32_t), L.stream) !=
                hipSuccess ||
            hipMemsetAsync(L.tileCounters, 0, MGPU_LANE_COUNTERS * sizeof(uint32_t), L.stream) !=
                hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        L.histDirty = false;
    }

    // digit_count
    mgpuCountArgs ca;
    std::memset(&ca, 0, sizeof(ca));
    ca.keys = keyBuf[0];
    ca.hist = L.hist;
    ca.clear = L.lookback;
    ca.clear_words = (lbWords + 3) & ~uint64_t(3);
    ca.n = n;
    ca.num_passes = (uint32_t)numPasses;
    ca.vec_ok = (reinterpret_cast<uintptr_t>(keyBuf[0]) & 15) == 0;
    int totalBits = 0;
    {
        int bit = firstBit;
        for (int p = 0; p < numPasses; ++p) {
            ca.shift[p] = (uint32_t)bit;
            ca.mask[p] = (1u << widths[p]) - 1u;
            bit += widths[p];
            totalBits += widths[p];
        }
    }
    if (earlyExit) {
        ca.unsorted = L.tileCounters + MGPU_UNSORTED_SLOT;
        ca.check_shift = (uint32_t)firstBit;
        ca.check_mask = totalBits >= 32 ? 0xffffffffu : ((1u << totalBits) - 1u);
    }
    if (fused) {
        ca.map_lo = toLo;
        ca.map_hi = toHi;
    }
    // Chunks cut per segment: of about 8192 keys for the 256-thread kernel (at
    // most 8 workgroups per CU), 32768 for the lane-column one (one per CU),
    // 16384 below 2^22 keys (per launch at 2^20: 10.6 against 11.6 us; at
    // 2^22 32768 is faster, 13.7 against 15.0 us).
    ca.num_segs = segs;
    ca.tile = tile;
    ca.num_tiles = numTiles;
    const bool cols = n >= kCountColsMinKeys;
    const uint32_t countThreads = cols ? 1024u : 256u;
    const uint32_t chunkTarget = cols ? (n < (1u << 22) ? 16384u : 32768u) : 8192u;
    const uint32_t maxGrid = (uint32_t)e->numCUs * (cols ? 1u : 8u);
    uint64_t segKeys = 0;  // keys of the largest segment
    for (uint32_t x = 0; x < segs; ++x) {
        const uint64_t b = (uint64_t)mgpu_seg_first_tile(x, numTiles, segs) * tile;
        const uint64_t en =
            x + 1 == segs ? n : (uint64_t)mgpu_seg_first_tile(x + 1, numTiles, segs) * tile;
        segKeys = std::max(segKeys, en - b);
    }
    uint32_t perSeg = (uint32_t)std::max<uint64_t>(1, (segKeys + chunkTarget - 1) / chunkTarget);
    perSeg = std::max(1u, std::min(perSeg, maxGrid / segs));
    ca.chunk = (uint32_t)((segKeys + perSeg - 1) / perSeg);
    ca.chunk = (ca.chunk + 1023) & ~1023u;
    ca.chunks_per_seg = (uint32_t)((segKeys + ca.chunk - 1) / ca.chunk);
    uint32_t grid = ca.chunks_per_seg * segs;
    // Enough workgroups to zero the lookback words quickly as well (the
    // extra ones find their chunk empty).
    uint32_t clearGrid =
        (uint32_t)((ca.clear_words / 4 + countThreads * 16 - 1) / (countThreads * 16));
    if (clearGrid > maxGrid) clearGrid = maxGrid;
    if (clearGrid > grid) {
        ca.chunks_per_seg = (clearGrid + segs - 1) / segs;
        grid = ca.chunks_per_seg * segs;
    }
    L.histDirty = true;
    hipFunction_t countFn = cols ? (fused ? e->fnCountColsKM : e->fnCountCols)
                                 : (fused ? e->fnCountKM : e->fnCount);
    st = mgpu_launch(e, L.stream, countFn, grid, countThreads, &ca, timed);
    if (st) return st;

    // digit_scan (also puts hist and the counters back to zero; with early
    // exit it also fills the lane's host-visible report).
    mgpuScanArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.hist = L.hist;
    sa.digit_start = L.digitStart;
    sa.tile_counters = L.tileCounters;
    sa.num_passes = (uint32_t)numPasses;
    sa.num_segs = segs;
    sa.seg_count = L.segCount;
    sa.seg_count_words = segs > 1 ? (uint32_t)kLaneSegCountWords : 0u;
    sa.report = earlyExit ? L.reportDev : nullptr;
    st = mgpu_launch(e, L.stream, e->fnScan, (unsigned)numPasses, 256, &sa, timed);
    if (st) return st;

    // Early exit (the role of the reference's sortDetectCounters readback and
    // its codes 1-3, mgpusort.cpp:433-451, for all passes at once), from the
    // scan's report, read once the stream is idle:
    //   * already-sorted input (by the whole bit field) skips every pass;
    //   * a pass whose digit is the same for every key is the identity;
    //   * so is the first pass when the input is already in its digit's
    //     order (digit_count counts that digit's descents, as the reference's
    //     count checks the current digit);
    //   * pass 0, when it runs, takes the early-exit kernel (waves already in
    //     digit order rank by position, without LDS atomics) when its digit
    //     has at most 3 descents per 1024-key wave: were they spread evenly,
    //     5% of the waves would be in order, and clustered disorder (nearly
    //     sorted input) leaves more.
    bool run[MGPU_MAX_PASSES];
    for (int p = 0; p < numPasses; ++p) run[p] = true;
    int eePass = -1;
    if (earlyExit) {
        if (hipStreamSynchronize(L.stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
        const volatile uint32_t* rep = L.reportWord;
        const bool sorted = rep[0] == 0;
        const uint32_t firstDescents = rep[1];
        bool leading = true;  // no pass has run yet: the input is as given
        for (int p = 0; p < numPasses; ++p) {
            const uint32_t most = rep[2 + p];
            // Descents are counted for pass 0's digit only: what a later
            // leading pass would find in the input is not known.
            const bool inOrder = p == 0 && firstDescents == 0;
            run[p] = !sorted && most != n && !(leading && inOrder);
            if (run[p] && leading) {
                leading = false;
                const uint64_t waves = (n + 1023) / 1024;
                const bool few = p == 0 && firstDescents <= 3 * waves;
                if (e->eePolicy == 1 || (e->eePolicy < 0 && few)) eePass = p;
            }
        }
    }
    L.histDirty = false;
    int lastRun = -1;
    for (int p = 0; p < numPasses; ++p)
        if (run[p]) lastRun = p;

    // rank_scatter per executed pass; executed pass k reads side k & 1.  The
    // first maps signed or float keys to their unsigned order as it reads
    // them, the last maps them back as it writes them.
    int bit = firstBit;
    for (int p = 0; p < numPasses; ++p) {
        const int shiftBits = bit;
        bit += widths[p];
        if (!run[p]) continue;
        mgpuPassArgs pa;
        std::memset(&pa, 0, sizeof(pa));
        const int in = done & 1, out = in ^ 1;
        pa.keys_in = keyBuf[in];
        pa.keys_out = keyBuf[out];
        int mode;
        if (valueCount == 0) mode = MGPU_VALUES_NONE;
        else if (valueCount == -1) mode = done == 0 ? MGPU_VALUES_INDEX : MGPU_VALUES_SINGLE;
        else if (valueCount == 1) mode = MGPU_VALUES_SINGLE;
        else mode = MGPU_VALUES_MULTI;
        int carried = valueCount == -1 ? 1 : valueCount;
        for (int v = 0; v < carried; ++v) {
            pa.vals_in[v] = valBuf[v][in];
            pa.vals_out[v] = valBuf[v][out];
        }
        pa.digit_start = L.digitStart + p * MGPU_RADIX;
        pa.lookback = L.lookback + (uint64_t)p * numTiles * MGPU_RADIX;
        pa.tile_counter = L.tileCounters + MGPU_TILE_COUNTER(p, 0);
        pa.finished = L.tileCounters + MGPU_FINISHED_SLOT + p;
        pa.error_flag = L.errorDev;
        pa.spin_limit = e->spinLimit;
        const bool maps = fused && (done == 0 || p == lastRun);
        if (maps && done == 0) pa.key_types |= kernelKeyType;
        if (maps && p == lastRun) pa.key_types |= kernelKeyType << 8;
        pa.n = n;
        pa.num_tiles = numTiles;
        pa.shift = (uint32_t)shiftBits;
        pa.mask = (1u << widths[p]) - 1u;
        pa.num_values = (uint32_t)carried;
        pa.num_segs = segs;
        if (segs > 1) {
            pa.seg_count_in = done == 0 ? nullptr
                                        : L.segCount + (size_t)p * MGPU_XSEG_COUNT_SEGS * MGPU_RADIX;
            int nextRun = p + 1;
            while (nextRun < numPasses && !run[nextRun]) ++nextRun;
            if (nextRun < numPasses) {
                pa.seg_count_out =
                    L.segCount + (size_t)nextRun * MGPU_XSEG_COUNT_SEGS * MGPU_RADIX;
                pa.next_mode = L.tileCounters + MGPU_MODE_SLOT + nextRun;
                pa.next_shift = ca.shift[nextRun];
                pa.next_mask = ca.mask[nextRun];
            }
            // Segment workgroups loop over their XCD's tiles: as many as fit
            // at once (at least one per XCD).
            const uint32_t g = std::max(segs, std::min(numTiles, var.xsegGrid));
            hipFunction_t f = maps                              ? var.xsegKM[mode]
                              : p == eePass && var.xsegEE[mode] ? var.xsegEE[mode]
                                                                : var.xseg[mode];
            st = mgpu_launch(e, L.stream, f, g, var.blockThreads, &pa, timed);
        } else {
            hipFunction_t f = maps                              ? var.passKM[mode]
                              : p == eePass && var.passEE[mode] ? var.passEE[mode]
                                                                : var.pass[mode];
            st = mgpu_launch(e, L.stream, f, numTiles, var.blockThreads, &pa, timed);
        }
        if (st) return st;
        ++done;
        if (executed) *executed = done;
    }
    if (typedKeys != 0 && !fused) {
        st = mgpu_key_transform_launch(e, L.stream, keyBuf[done & 1], n, typedKeys, false);
        if (st) return st;
    }
    // Index generation with no pass left to run: the indices are the identity.
    if (done == 0 && valueCount == -1) {
        mgpuIotaArgs ia;
        ia.out = valBuf[0][0];
        ia.n = n;
        ia.pad_ = 0;
        uint32_t g = std::min<uint32_t>((n + 255) / 256, (uint32_t)e->numCUs * 8);
        st = mgpu_launch(e, L.stream, e->fnIota, g, 256, &ia, timed);
        if (st) return st;
    }
    return SORT_STATUS_SUCCESS;
}

// Validation shared by sortArray and sortArrayEx (mgpusort.cpp:360-364,
// 609-611, 625).
sortStatus_t mgpu_validate_data(sortEngine_d* engine, const sortData_d* data) {
    if (!engine || !data) return SORT_STATUS_INVALID_VALUE;
    if (data->numElements < 0 || data->numElements > data->maxElements)
        return SORT_STATUS_INVALID_VALUE;
    if (data->firstBit < 0 || data->endBit > 32 || data->endBit < data->firstBit)
        return SORT_STATUS_INVALID_VALUE;
    if (data->valueCount < -1 || data->valueCount > 6) return SORT_STATUS_INVALID_VALUE;
    if (data->numElements > SORT_MAX_ELEMENTS) return SORT_STATUS_CONFIG_NOT_SUPPORTED;
    return SORT_STATUS_SUCCESS;
}

// sortArrayFromList (mgpusort.cpp:560-601) without the end-key handling:
// runs the passes, swaps every buffer pair once per executed pass and toggles
// parity, exactly as sortPass does (mgpusort.cpp:518-526).
sortStatus_t mgpu_sort_data(sortEngine_d* e, SortLane& lane, sortData_d* data, int numPasses,
                            const int* widths, int keyType) {
    const uint32_t n = (uint32_t)data->numElements;
    if (n == 0 || numPasses == 0) return SORT_STATUS_SUCCESS;
    const int carried = data->valueCount == -1 ? 1 : data->valueCount;
    if (!data->keys[0] || !data->keys[1]) return SORT_STATUS_INVALID_VALUE;
    // Index generation reads no values on its first pass, so values1[0] is
    // needed only as the target of a second pass.
    for (int v = 0; v < carried; ++v)
        if (!data->values[v][1] ||
            (!data->values[v][0] && (data->valueCount != -1 || numPasses > 1)))
            return SORT_STATUS_INVALID_VALUE;

    uint32_t* keyBuf[2] = {reinterpret_cast<uint32_t*>(data->keys[0]),
                           reinterpret_cast<uint32_t*>(data->keys[1])};
    uint32_t* valBuf[MGPU_MAX_VALUES][2] = {};
    for (int v = 0; v < MGPU_MAX_VALUES; ++v)
        for (int s = 0; s < 2; ++s) valBuf[v][s] = reinterpret_cast<uint32_t*>(data->values[v][s]);

    int executed = 0;
    sortStatus_t st = mgpu_sort_passes(e, lane, n, data->firstBit, numPasses, widths,
                                       data->valueCount, keyBuf, valBuf, data->earlyExit,
                                       &executed, keyType);
    if (st) return st;
    if (executed & 1) {
        std::swap(data->keys[0], data->keys[1]);
        for (int v = 0; v < MGPU_MAX_VALUES; ++v) std::swap(data->values[v][0], data->values[v][1]);
    }
    data->parity ^= (executed & 1);
    return SORT_STATUS_SUCCESS;
}

namespace {

sortStatus_t sort_with_widths(sortEngine_t e, sortData_t data, int numPasses, const int* widths) {
    return mgpu_sort_data(e, e->lane, data, numPasses, widths);
}

}  // namespace

// ---------------------------------------------------------------------------
// Exported C ABI.

extern "C" {

const char* SORTAPI sortStatusString(sortStatus_t status) {
    int code = (int)status;
    if (code < 0 || code >= (int)(sizeof(kStatusStrings) / sizeof(kStatusStrings[0]))) return 0;
    return kStatusStrings[code];
}

int SORTAPI sortPlanPasses(int numBits, int* widths) {
    if (numBits < 0 || numBits > 32) return 0;
    return mgpu_plan_passes(numBits, MGPU_RADIX_BITS, widths);
}

// sortCreateEngine (mgpusort.cpp:200-250): attach to the current device,
// check it, load the code objects, allocate the small fixed scratch.
sortStatus_t SORTAPI sortCreateEngine(const char* kernelPath, sortEngine_t* engine) {
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    *engine = nullptr;
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0) return SORT_STATUS_INVALID_CONTEXT;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SORT_STATUS_INVALID_CONTEXT;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SORT_STATUS_UNSUPPORTED_DEVICE;

    sortEngine_d* e = new (std::nothrow) sortEngine_d();
    if (!e) return SORT_STATUS_HOST_ALLOC_FAILED;
    e->device = device;
    e->numCUs = prop.multiProcessorCount;
    const char* env = std::getenv("MGPUSORT_KERNEL_PATH");
    if (env && *env) e->kernelPath = env;
    else if (kernelPath && *kernelPath) e->kernelPath = kernelPath;
    else e->kernelPath = library_dir();

    sortStatus_t st = mgpu_engine_load(e);
    if (st) {
        mgpu_engine_destroy(e);
        return st;
    }
    // sortArrayNoAlloc's scratch on this device (its absence only disables
    // that entry point).
    (void)mgpu_noalloc_setup(device);
    *engine = e;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortLoadKernel(sortEngine_t engine, int numSortThreads, int valuesPerThread,
                                    bool useTransList, int valueCount) {
    (void)useTransList;
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    // Same acceptance as LoadKernels (mgpusort.cpp:148-153,182-184).
    if (numSortThreads != 128 && numSortThreads != 256) return SORT_STATUS_INVALID_VALUE;
    if (valuesPerThread != 4 && valuesPerThread != 8) return SORT_STATUS_INVALID_VALUE;
    if (valueCount < -1 || valueCount > 6) return SORT_STATUS_INVALID_VALUE;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortIncEngine(sortEngine_t engine) {
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    ++engine->refcount;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortReleaseEngine(sortEngine_t engine) {
    if (engine && --engine->refcount == 0) mgpu_engine_destroy(engine);
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortDestroyEngine(sortEngine_t engine) { return sortReleaseEngine(engine); }

sortStatus_t SORTAPI sortSetStream(sortEngine_t engine, void* stream) {
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    hipStream_t next = reinterpret_cast<hipStream_t>(stream);
    SortLane& L = engine->lane;
    if (next != L.stream) {
        // Work still queued on the old stream uses the lane's scratch (digit
        // counts, lookback words, tile counters) that the next sort reuses:
        // the new stream waits for it.
        if (!L.streamSwitch &&
            hipEventCreateWithFlags(&L.streamSwitch, hipEventDisableTiming) != hipSuccess) {
            L.streamSwitch = nullptr;
            return SORT_STATUS_DEVICE_ERROR;
        }
        if (hipEventRecord(L.streamSwitch, L.stream) != hipSuccess ||
            hipStreamWaitEvent(next, L.streamSwitch, 0) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        L.stream = next;
    }
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortCheckDevice(sortEngine_t engine) {
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    if (hipStreamSynchronize(engine->lane.stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
    return mgpu_lane_take_error(engine->lane);
}

sortStatus_t SORTAPI sortTransformKeys(sortEngine_t engine, sortDevicePtr_t keys, int numElements,
                                       int keyType, bool toRadixOrder) {
    if (!engine || numElements < 0 || numElements > SORT_MAX_ELEMENTS ||
        (keyType != SORT_KEYS_INT32 && keyType != SORT_KEYS_FLOAT32) || (!keys && numElements))
        return SORT_STATUS_INVALID_VALUE;
    if (numElements == 0) return SORT_STATUS_SUCCESS;
    return mgpu_key_transform_launch(
        engine, engine->lane.stream, reinterpret_cast<uint32_t*>(keys), (uint32_t)numElements,
        keyType == SORT_KEYS_INT32 ? MGPU_KEYS_INT32 : MGPU_KEYS_FLOAT32, toRadixOrder);
}

sortStatus_t SORTAPI sortEngineInfo(sortEngine_t engine, sortEngineInfo_t* info) {
    if (!engine || !info) return SORT_STATUS_INVALID_VALUE;
    std::memset(info, 0, sizeof(*info));
    info->device = engine->device;
    info->numCUs = engine->numCUs;
    info->xcdSegments = (int)engine->xcdSegs;
    info->rankMethod = (int)engine->rankMethod;
    info->smallTileKeys = (int)engine->small.tile();
    info->largeTileKeys = (int)engine->large.tile();
    if (hipStreamSynchronize(engine->lane.stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
    info->deviceError = mgpu_lane_peek_error(engine->lane);
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortEnableTiming(sortEngine_t engine, bool enable) {
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    engine->timing = enable;
    return SORT_STATUS_SUCCESS;
}

int SORTAPI sortReadTiming(sortEngine_t engine, sortKernelTiming_t* out, int maxEntries) {
    if (!engine) return -1;
    if (hipStreamSynchronize(engine->lane.stream) != hipSuccess) return -1;
    std::vector<sortKernelTiming_t> agg;
    std::map<hipFunction_t, size_t> slot;
    for (auto& t : engine->timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) != hipSuccess) ms = -1.f;
        auto it = slot.find(t.fn);
        if (it == slot.end()) {
            sortKernelTiming_t k;
            std::memset(&k, 0, sizeof(k));
            std::strncpy(k.name, engine->kernelNames[t.fn].c_str(), sizeof(k.name) - 1);
            k.minMs = 1e30;
            it = slot.emplace(t.fn, agg.size()).first;
            agg.push_back(k);
        }
        sortKernelTiming_t& k = agg[it->second];
        k.launches += 1;
        k.totalMs += ms;
        if (ms < k.minMs) k.minMs = ms;
        if (ms > k.maxMs) k.maxMs = ms;
        engine->eventPool.push_back(t.start);
        engine->eventPool.push_back(t.stop);
    }
    engine->timed.clear();
    for (int i = 0; i < (int)agg.size() && i < maxEntries && out; ++i) out[i] = agg[i];
    return (int)agg.size();
}

sortStatus_t SORTAPI sortGetStream(sortEngine_t engine, void** stream) {
    if (!engine || !stream) return SORT_STATUS_INVALID_VALUE;
    *stream = reinterpret_cast<void*>(engine->lane.stream);
    return SORT_STATUS_SUCCESS;
}

// sortArray (mgpusort.cpp:607-616).  A self-check failure of earlier work on
// the engine's stream that has not been reported yet is returned here, before
// anything is queued (the sort itself is asynchronous, so its own failures
// surface at the next call, at sortCheckDevice, or in a synchronous wrapper).
sortStatus_t SORTAPI sortArray(sortEngine_t engine, sortData_t data) {
    if (!engine || !data) return SORT_STATUS_INVALID_VALUE;
    if (mgpu_lane_take_error(engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    int numBits = data->endBit - data->firstBit;
    if (!numBits) return SORT_STATUS_SUCCESS;  // as mgpusort.cpp:610
    sortStatus_t st = mgpu_validate_data(engine, data);
    if (st) return st;
    int widths[MGPU_MAX_PASSES];
    int passes = mgpu_plan_passes(numBits, MGPU_RADIX_BITS, widths);
    return sort_with_widths(engine, data, passes, widths);
}

// sortArray over int32 or float32 keys, with the order mapping inside the
// count and the first and last executed passes (no pass of its own).
sortStatus_t SORTAPI sortArrayTyped(sortEngine_t engine, sortData_t data, int keyType) {
    if (!engine || !data) return SORT_STATUS_INVALID_VALUE;
    if (keyType != SORT_KEYS_UINT32 && keyType != SORT_KEYS_INT32 &&
        keyType != SORT_KEYS_FLOAT32)
        return SORT_STATUS_INVALID_VALUE;
    if (mgpu_lane_take_error(engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    int numBits = data->endBit - data->firstBit;
    if (!numBits) return SORT_STATUS_SUCCESS;
    sortStatus_t st = mgpu_validate_data(engine, data);
    if (st) return st;
    int widths[MGPU_MAX_PASSES];
    int passes = mgpu_plan_passes(numBits, MGPU_RADIX_BITS, widths);
    return mgpu_sort_data(engine, engine->lane, data, passes, widths, keyType);
}

// sortArrayEx (mgpusort.cpp:618-648): the caller's digit width sets the pass
// count; the widths are split evenly with narrower passes first.
sortStatus_t SORTAPI sortArrayEx(sortEngine_t engine, sortData_t data, int numSortThreads,
                                 int valuesPerThread, int bitPass, bool useTransList) {
    (void)useTransList;
    if (!engine || !data) return SORT_STATUS_INVALID_VALUE;
    if ((numSortThreads != 128 && numSortThreads != 256) || valuesPerThread != 8)
        return SORT_STATUS_INVALID_VALUE;
    if (data->numElements > data->maxElements) return SORT_STATUS_INVALID_VALUE;
    if (mgpu_lane_take_error(engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    if (bitPass <= 0 || bitPass > 6) bitPass = 6;
    int numBits = data->endBit - data->firstBit;
    if (!numBits) return SORT_STATUS_SUCCESS;
    sortStatus_t st = mgpu_validate_data(engine, data);
    if (st) return st;
    if (bitPass > numBits) bitPass = numBits;
    int widths[32];
    int passes = mgpu_plan_passes(numBits, bitPass, widths);
    // More than MGPU_MAX_PASSES passes: run them in groups (each group one
    // digit_count + scan); parity and swaps stay per pass.
    int done = 0, bit = data->firstBit;
    while (done < passes) {
        int group = passes - done < MGPU_MAX_PASSES ? passes - done : MGPU_MAX_PASSES;
        int savedFirst = data->firstBit;
        int savedValueCount = data->valueCount;
        data->firstBit = bit;
        // Index generation only on the first executed pass (GetSortCode,
        // mgpusort.cpp:553); later groups carry the indices as one value.
        if (done > 0 && savedValueCount == -1) data->valueCount = 1;
        st = sort_with_widths(engine, data, group, widths + done);
        data->firstBit = savedFirst;
        data->valueCount = savedValueCount;
        if (st) return st;
        for (int p = 0; p < group; ++p) bit += widths[done + p];
        done += group;
    }
    return SORT_STATUS_SUCCESS;
}

// sortCreateData (mgpusort.cpp:655-682).  Unlike the reference, a failed
// allocation is returned (and nothing is leaked).
sortStatus_t SORTAPI sortCreateData(sortEngine_t engine, int maxElements, int valueCount,
                                    sortData_t* data) {
    if (!data) return SORT_STATUS_INVALID_VALUE;
    *data = nullptr;
    if (valueCount > 6 || valueCount < -1 || maxElements < 0) return SORT_STATUS_INVALID_VALUE;
    sortData_d* d = new (std::nothrow) sortData_d;
    if (!d) return SORT_STATUS_HOST_ALLOC_FAILED;
    std::memset(d, 0, sizeof(*d));
    d->maxElements = round_up(maxElements, kAllocRound);
    d->numElements = maxElements;
    d->valueCount = valueCount;
    d->firstBit = 0;
    d->endBit = 32;
    sortStatus_t st = sortAllocData(engine, d);
    if (st) {
        sortFreeData(d);
        delete d;
        return st;
    }
    *data = d;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortDestroyData(sortData_t data) {
    if (!data) return SORT_STATUS_SUCCESS;
    sortFreeData(data);
    delete data;
    return SORT_STATUS_SUCCESS;
}

// sortAllocData (mgpusort.cpp:692-719).
sortStatus_t SORTAPI sortAllocData(sortEngine_t engine, sortData_t data) {
    if (!data) return SORT_STATUS_INVALID_VALUE;
    (void)engine;
    if (data->valueCount > 6 || data->valueCount < -1 || data->maxElements < 0)
        return SORT_STATUS_INVALID_VALUE;
    size_t elems = (size_t)round_up(data->maxElements, kAllocRound);
    if (!elems) elems = kAllocRound;
    int count = data->valueCount == -1 ? 1 : data->valueCount;
    CUdeviceptr* ptrs = data->keys;  // keys[2] then values[6][2], contiguous
    CUdeviceptr fresh[14] = {};
    for (int i = 0; i < 2 + 2 * count; ++i) {
        if (ptrs[i]) continue;
        void* p = nullptr;
        if (hipMalloc(&p, elems * sizeof(uint32_t)) != hipSuccess) {
            for (int j = 0; j < i; ++j)
                if (fresh[j]) (void)hipFree(reinterpret_cast<void*>(fresh[j]));
            return SORT_STATUS_DEVICE_ALLOC_FAILED;
        }
        fresh[i] = reinterpret_cast<CUdeviceptr>(p);
    }
    for (int i = 0; i < 14; ++i)
        if (fresh[i]) ptrs[i] = fresh[i];
    data->parity = 0;
    return SORT_STATUS_SUCCESS;
}

// sortFreeData (mgpusort.cpp:721-734): frees every non-null buffer.
sortStatus_t SORTAPI sortFreeData(sortData_d* data) {
    if (!data) return SORT_STATUS_INVALID_VALUE;
    CUdeviceptr* ptrs = data->keys;
    for (int i = 0; i < 14; ++i)
        if (ptrs[i]) (void)hipFree(reinterpret_cast<void*>(ptrs[i]));
    std::memset(data, 0, sizeof(sortData_d));
    return SORT_STATUS_SUCCESS;
}

// sortHost (mgpusort.cpp:740-776), copying 4 bytes per element (the reference
// copies sizeof(double) per element, past the end of both buffers).
sortStatus_t SORTAPI sortHost(sortEngine_t engine, unsigned int* keys, unsigned int* values,
                              int numElements, int numBits) {
    if (!engine || !keys || numElements < 0 || numBits < 0 || numBits > 32)
        return SORT_STATUS_INVALID_VALUE;
    sortData_d d;
    std::memset(&d, 0, sizeof(d));
    d.maxElements = numElements;
    d.numElements = numElements;
    d.valueCount = values ? 1 : 0;
    sortStatus_t st = sortAllocData(engine, &d);
    if (st) return st;
    d.endBit = numBits;
    size_t bytes = sizeof(uint32_t) * (size_t)numElements;
    hipStream_t s = engine->lane.stream;
    do {
        if (hipMemcpyAsync(reinterpret_cast<void*>(d.keys[0]), keys, bytes, hipMemcpyHostToDevice,
                           s) != hipSuccess) { st = SORT_STATUS_DEVICE_ERROR; break; }
        if (values && hipMemcpyAsync(reinterpret_cast<void*>(d.values1[0]), values, bytes,
                                     hipMemcpyHostToDevice, s) != hipSuccess) {
            st = SORT_STATUS_DEVICE_ERROR; break;
        }
        st = sortArray(engine, &d);
        if (st) break;
        if (hipMemcpyAsync(keys, reinterpret_cast<void*>(d.keys[0]), bytes, hipMemcpyDeviceToHost,
                           s) != hipSuccess) { st = SORT_STATUS_DEVICE_ERROR; break; }
        if (values && hipMemcpyAsync(values, reinterpret_cast<void*>(d.values1[0]), bytes,
                                     hipMemcpyDeviceToHost, s) != hipSuccess) {
            st = SORT_STATUS_DEVICE_ERROR; break;
        }
        if (hipStreamSynchronize(s) != hipSuccess) st = SORT_STATUS_DEVICE_ERROR;
        else st = mgpu_lane_take_error(engine->lane);
    } while (0);
    (void)hipStreamSynchronize(s);
    sortFreeData(&d);
    return st;
}

// sortDevice (mgpusort.cpp:782-806).  When the pass count is odd the result
// is copied from keys[0] (the library buffer) back into the caller's buffer;
// the reference copies from keys[1], which by then is the caller's own
// unsorted buffer.
sortStatus_t SORTAPI sortDevice(sortEngine_t engine, CUdeviceptr keys, CUdeviceptr values,
                                int numElements, int numBits) {
    if (!engine || !keys || numElements < 0 || numBits < 0 || numBits > 32)
        return SORT_STATUS_INVALID_VALUE;
    sortData_d d;
    std::memset(&d, 0, sizeof(d));
    d.keys[0] = keys;
    if (values) d.values1[0] = values;
    d.maxElements = numElements;
    d.numElements = numElements;
    d.valueCount = values ? 1 : 0;
    sortStatus_t st = sortAllocData(engine, &d);
    if (st) return st;
    d.endBit = numBits;
    st = sortArray(engine, &d);
    size_t bytes = sizeof(uint32_t) * (size_t)numElements;
    if (!st && d.parity) {
        if (hipMemcpyAsync(reinterpret_cast<void*>(keys), reinterpret_cast<void*>(d.keys[0]), bytes,
                           hipMemcpyDeviceToDevice, engine->lane.stream) != hipSuccess)
            st = SORT_STATUS_DEVICE_ERROR;
        if (!st && values &&
            hipMemcpyAsync(reinterpret_cast<void*>(values), reinterpret_cast<void*>(d.values1[0]),
                           bytes, hipMemcpyDeviceToDevice, engine->lane.stream) != hipSuccess)
            st = SORT_STATUS_DEVICE_ERROR;
    }
    if (hipStreamSynchronize(engine->lane.stream) != hipSuccess && !st) st = SORT_STATUS_DEVICE_ERROR;
    if (!st) st = mgpu_lane_take_error(engine->lane);
    // Free only what sortAllocData added: the caller's buffers sit at
    // [parity] after the sort.
    int p = d.parity & 1;
    d.keys[p] = 0;
    if (values) d.values1[p] = 0;
    sortFreeData(&d);
    return st;
}

}  // extern "C"
