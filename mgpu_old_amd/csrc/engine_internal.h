// Engine state shared by the translation units of libmgpusort (the
// counterpart of sort/src/mgpusort/sortdll.h:7-74).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/mgpusort.h"
#include "kernel_args.h"
#include "segmented_internal.h"

// A rank_scatter tile shape: blockThreads * itemsPerThread keys per tile, and
// the kernels of that shape (by value mode, MGPU_VALUES_*).
struct PassVariant {
    uint32_t blockThreads = 0;
    uint32_t itemsPerThread = 0;
    hipFunction_t pass[4] = {};   // mgpu_rank_scatter_b*_i*_r*_v<mode>: one device-wide chain
    hipFunction_t xseg[4] = {};   // mgpu_rank_scatter_b*_i*_r*_x_v<mode>: a chain per XCD segment
    hipFunction_t multi[4] = {};  // mgpu_multi_scatter_b*_i*_r*_v<mode>
    hipFunction_t multiX[4] = {}; // mgpu_multi_scatter_b*_i*_r*_x_v<mode>: chains per XCD
    uint32_t multiXGrid = 0;      // workgroups of a multiX launch: all that fit
    // Packed tiles of several arrays (atomic-add ranking only; null when not
    // built for the shape): mgpu_multi_packed_b*_i*_v<mode> and ..._x_v<mode>.
    hipFunction_t packed[4] = {};
    hipFunction_t packedX[4] = {};
    uint32_t packedXGrid = 0;
    // Early-exit forms (atomic-add ranking only; null when not built for the
    // shape): mgpu_rank_scatter_b*_i*_r1_ee_v<mode> and ..._r1_x_ee_v<mode>.
    hipFunction_t passEE[4] = {};
    hipFunction_t xsegEE[4] = {};
    // Forms that map signed or float keys in a typed sort's first and last
    // passes (atomic-add ranking only; null when not built for the shape):
    // mgpu_rank_scatter_b*_i*_r1_km_v<mode> and ..._r1_x_km_v<mode>.
    hipFunction_t passKM[4] = {};
    hipFunction_t xsegKM[4] = {};
    uint32_t xsegGrid = 0;        // workgroups of an xseg launch: all that fit on the device
    uint32_t tile() const { return blockThreads * itemsPerThread; }
};

// Words of a lane's histogram (and of its digit starts): one set of pass
// histograms per segment.
constexpr size_t kLaneHistWords = (size_t)MGPU_MAX_SEGS * MGPU_MAX_PASSES * MGPU_RADIX;
// Per-segment digit counts of each pass's input that the pass before it
// fills in (segments per XCD).
constexpr size_t kLaneSegCountWords = (size_t)MGPU_MAX_PASSES * MGPU_XSEG_COUNT_SEGS * MGPU_RADIX;

// One in-order queue of sorts: a stream plus the device scratch of the
// device-wide passes.  The engine owns one (its default lane, on the stream
// chosen with sortSetStream); the sort server owns one per in-flight batch.
struct SortLane {
    hipStream_t stream = nullptr;
    uint32_t* fixed = nullptr;  // hist, digitStart ([segments][4][256] each), counters, segCount
    uint32_t* hist = nullptr;
    uint32_t* digitStart = nullptr;
    uint32_t* tileCounters = nullptr;
    uint32_t* segCount = nullptr;  // [MGPU_MAX_PASSES][MGPU_XSEG_COUNT_SEGS][MGPU_RADIX]
    bool histDirty = false;  // hist/tileCounters may hold a partial sort's counts
    uint32_t* lookback = nullptr;  // grow-only
    uint64_t lookbackWords = 0;
    // Device-wide passes over several arrays: hist and digitStart per array
    // ([arrays][4][256] each, grow-only, hist zero between sorts).
    uint32_t* multi = nullptr;
    uint32_t multiArrays = 0;
    bool multiDirty = false;
    // Segmented sorts: a second stream on which the persistent kernels run
    // beside the device-wide passes (forked from and joined back into
    // `stream` by the two events); created on first use.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool noSide = false;  // never fork: everything on `stream` (sort server slots)
    // Error word of the lane's kernels: pinned host memory mapped into the
    // device, which a kernel sets (system-scope store) when a self-check
    // fails: 1 a lookback wait gave up, 2 a tile counter left dirty, 4 a
    // scatter offset out of range, 5 a segment left without a workgroup on
    // its XCD.  The host reads it without a copy.
    uint32_t* errorWord = nullptr;  // host address
    uint32_t* errorDev = nullptr;   // the same word as the kernels address it
    // The early-exit report digit_scan writes in the same pinned block (its
    // own cache line): descents and the largest digit count of each pass.
    uint32_t* reportWord = nullptr;
    uint32_t* reportDev = nullptr;
    hipEvent_t streamSwitch = nullptr;  // orders a new stream after the old (sortSetStream)
    // A view over caller-provided scratch (sortArrayNoAlloc): the passes
    // never allocate, clear with a memset, read back or time anything; a sort
    // whose lookback words do not fit lookbackWords is refused.
    bool noAlloc = false;
};

// The lane's pinned, device-mapped block: the error word at 0, the early-exit
// report from word kLaneReportOffset.
constexpr size_t kLanePinnedBytes = 256;
constexpr size_t kLaneReportOffset = 32;

sortStatus_t mgpu_lane_init(SortLane* l);
void mgpu_lane_free(SortLane* l);
// The lane's error code (0: none) without clearing it.
uint32_t mgpu_lane_peek_error(const SortLane& l);
// SORT_STATUS_KERNEL_ERROR, and the word cleared, when a kernel of the lane
// has set it since the last take; SORT_STATUS_SUCCESS otherwise.  Only work
// that has finished is seen: callers synchronize first when they need all of
// it.
sortStatus_t mgpu_lane_take_error(SortLane& l);

struct sortEngine_d {
    int refcount = 1;  // sortIncEngine / sortReleaseEngine (not atomic, as util/util.h:84-98)
    int device = -1;
    int numCUs = 0;
    std::string kernelPath;
    SortLane lane;  // default lane: sortArray and friends run here

    hipModule_t radixModule = nullptr;
    hipFunction_t fnCount = nullptr;      // 256 threads, per-wave histograms
    hipFunction_t fnCountCols = nullptr;  // 1024 threads, lane-column histograms (large sorts)
    hipFunction_t fnCountKM = nullptr;    // the two over signed or float keys
    hipFunction_t fnCountColsKM = nullptr;
    hipFunction_t fnScan = nullptr;
    hipFunction_t fnIota = nullptr;
    hipFunction_t fnKeyTransform = nullptr;
    hipFunction_t fnMultiCount = nullptr;      // 256 threads, per-wave histograms
    hipFunction_t fnMultiCountCols = nullptr;  // 1024 threads, lane-column histograms
    hipFunction_t fnMultiScan = nullptr;
    // Tile shapes: `large` once a pass has a large tile for every
    // kCUsPerLargeTile CUs, `small` below that (more, smaller tiles keep the
    // CUs busy).
    PassVariant small, large;
    uint32_t rankMethod = 0;  // 1: atomic-add ranking (checked at creation), 0: match masks
    // Segments per XCD for single-array passes: the number of XCDs when the
    // creation-time check passed (and MGPUSORT_XCD is not 0), else 1.
    // xcdForce (MGPUSORT_XCD=1): use them at every size where each XCD gets
    // a tile, not only from kXsegMinTilesPerSeg tiles per XCD.
    uint32_t xcdSegs = 1;
    bool xcdForce = false;

    // Polls of a lookback word before a pass gives up (MGPUSORT_SPIN_LIMIT
    // at creation; 0 makes every lookback give up at once, for tests of the
    // error path).
    uint32_t spinLimit = 1u << 22;
    // Early-exit kernel for the first pass that runs (MGPUSORT_EE at
    // creation): -1 by the descent counts, 0 never, 1 always.
    int eePolicy = -1;

    // sortArray tuning table (kernelPath + "mgpusort_sort_tuning.txt", written
    // by the autotuner): per value class (MGPU_SORT_CLASS_*), entries sorted
    // by maxLog2; a class without entries uses the built-in rules.
    struct SortTuneEntry {
        int maxLog2;
        bool large;  // tile shape: large (else small)
        bool xcd;    // a lookback chain per XCD (else one across the device)
    };
    std::vector<SortTuneEntry> sortTable[4];

    // Segmented sort: kernel variants [seg_variant_index] (null where not
    // built), workgroups per CU of each, the size-bucket tuning table and the
    // device-wide passes' settings.
    hipModule_t segModule = nullptr;
    hipFunction_t segFn[kSegVariants] = {};
    int segOccupancy[kSegVariants] = {};
    // The same variants without the routine for segments larger than a
    // workgroup's tile (mgpu_segsort_*_ng): fewer registers, more waves.
    hipFunction_t segFnTile[kSegVariants] = {};
    int segOccupancyTile[kSegVariants] = {};
    std::vector<SegTuneEntry> segTable;
    SegWideTune segWide;

    // Per-kernel timing (sortEnableTiming / sortReadTiming).
    struct TimedLaunch {
        hipFunction_t fn;
        hipEvent_t start, stop;
    };
    bool timing = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> eventPool;
    std::map<hipFunction_t, std::string> kernelNames;
};

// Looks up the current device's copy of sortArrayNoAlloc's scratch array
// once per device (false: not available on this device).
bool mgpu_noalloc_setup(int device);

// Loads the code objects named from e->kernelPath and allocates the fixed
// scratch.  On failure the engine is left for mgpu_engine_destroy.
sortStatus_t mgpu_engine_load(sortEngine_d* e);
void mgpu_engine_destroy(sortEngine_d* e);

// Pass split of numBits into passes of at most maxWidth bits; returns the
// pass count and fills widths (narrower passes first).
int mgpu_plan_passes(int numBits, int maxWidth, int* widths);

// timed = false: never event-timed, even with sortEnableTiming on.
sortStatus_t mgpu_launch(sortEngine_d* e, hipStream_t stream, hipFunction_t f, unsigned grid,
                         unsigned block, void* args, bool timed = true);

// A pass over `keys` keys in all uses the large tile shape once it has a
// large tile for every this many CUs (2^21 keys on 256 CUs), else the small
// one (measured: 1024 x 16 is ahead from 2^21 keys up, 512 x 16 below).
constexpr uint32_t kCUsPerLargeTile = 2;
// Segments per XCD pay a fixed cost per pass (each workgroup's counts for the
// next pass are flushed once); they win from about this many tiles per XCD
// (2^24 keys in 16384-key tiles on 8 XCDs).
constexpr uint32_t kXsegMinTilesPerSeg = 128;
// digit_count with lane-column histograms (one 1024-thread workgroup per CU)
// from this many keys up; the per-wave form (256 threads, up to 8 per CU)
// below.  Measured per launch: 2^20 keys 11.7 against 12.8 us, 2^21 12.2
// against 15.2, 2^22 13.8 against 20.0; equal at 2^16-2^19; the per-wave
// form faster at 2^12 and below (7.6 against 8.2 us).
constexpr uint32_t kCountColsMinKeys = 1u << 20;
const PassVariant& mgpu_pick_variant(const sortEngine_d* e, uint64_t keys);

// Value classes of the sortArray tuning table.
enum { MGPU_SORT_CLASS_KEYS = 0, MGPU_SORT_CLASS_INDEX = 1, MGPU_SORT_CLASS_KV = 2,
       MGPU_SORT_CLASS_MULTI = 3 };
// Tile shape and lookback segments (1, or one per XCD) of a single-array sort
// of n keys: the tuning table's entry for its size and value class, else the
// built-in rules; MGPUSORT_TILE and MGPUSORT_XCD override both.
struct SortPlanChoice {
    const PassVariant* variant;
    uint32_t segs;
};
SortPlanChoice mgpu_pick_plan(const sortEngine_d* e, uint32_t n, int valueCount);
// Reads kernelPath + "mgpusort_sort_tuning.txt" into e->sortTable (missing
// file or no valid line: every class keeps the built-in rules).
void mgpu_load_sort_table(sortEngine_d* e);

// Grows the lane's lookback words to cover a device-wide sort of n keys in
// `passes` passes (a reallocation synchronizes the lane's stream).
sortStatus_t mgpu_ensure_lookback(sortEngine_d* e, SortLane& lane, uint32_t n, int passes);
sortStatus_t mgpu_ensure_lookback_words(SortLane& lane, uint64_t words);
// Grows the per-array scratch of the multi-array passes to `arrays` arrays.
sortStatus_t mgpu_ensure_multi(SortLane& lane, uint32_t arrays);

// Runs numPasses digit passes over n keys starting at firstBit.  Executed pass
// k reads buffer side (k & 1) and writes side (k & 1) ^ 1 of keyBuf / valBuf,
// so the result is on side (executed & 1).  Without earlyExit every pass
// runs; with it, passes whose digit is constant are skipped, and all of them
// when the keys are already in order (one blocking readback).  valueCount as
// in sortData_d.  keyType SORT_KEYS_INT32 / SORT_KEYS_FLOAT32: the keys are
// read and checked in their unsigned order (sortTransformKeys' mapping) by
// digit_count and the first executed pass, and mapped back by the last one.
sortStatus_t mgpu_sort_passes(sortEngine_d* e, SortLane& lane, uint32_t n, int firstBit,
                              int numPasses, const int* widths, int valueCount,
                              uint32_t* const keyBuf[2],
                              uint32_t* const valBuf[MGPU_MAX_VALUES][2],
                              bool earlyExit = false, int* executed = nullptr,
                              int keyType = SORT_KEYS_UINT32);

// Argument checks of sortArray / sortArrayEx (mgpusort.cpp:360-364,609-611).
sortStatus_t mgpu_validate_data(sortEngine_d* e, const sortData_d* data);
// sortArrayFromList (mgpusort.cpp:560-601) on a lane: runs the passes over
// data's buffers, swaps every buffer pair once per executed pass and toggles
// parity (mgpusort.cpp:518-526).
sortStatus_t mgpu_sort_data(sortEngine_d* e, SortLane& lane, sortData_d* data, int numPasses,
                            const int* widths, int keyType = SORT_KEYS_UINT32);
