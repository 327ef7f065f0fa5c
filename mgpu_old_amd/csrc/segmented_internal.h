// Segmented-sort state shared between the engine and segmented.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/mgpusort.h"
#include "kernel_args.h"

struct sortEngine_d;
struct SortLane;
struct PassVariant;

// Kernel variants of mgpusort_seg.hsaco: radix bits 4..8 x items per thread
// {8, 16}, named mgpu_segsort_r<bits>_i<items>.
constexpr int kSegMinBits = 4;
constexpr int kSegMaxBits = 8;
constexpr int kSegVariants = (kSegMaxBits - kSegMinBits + 1) * 2;

inline int seg_variant_index(int rb, int ipt) { return (rb - kSegMinBits) * 2 + (ipt == 16); }
inline int seg_variant_bits(int v) { return kSegMinBits + v / 2; }
inline int seg_variant_ipt(int v) { return (v & 1) ? 16 : 8; }

// One row of the tuning table: segments with ceil(log2 n) <= maxLog2 use
// this variant (rows sorted by maxLog2).
struct SegTuneEntry {
    int maxLog2;
    int radixBits;
    int itemsPerThread;
};

// One persistent launch of one kernel variant.
struct SegLaunch {
    int variant = 0;
    hipFunction_t fn = nullptr;
    uint32_t grid = 0;
    uint32_t bigOffset = 0, smallOffset = 0;
    uint32_t numBig = 0, numSmall = 0;
    int passes = 0;
    int widths[MGPU_SEG_MAX_PASSES] = {};
};

// Reusable home for the device arrays of successive builds on one lane
// (descriptors, work lists, queue counters), filled with async copies from
// pinned staging so that a build neither allocates nor synchronizes.
struct SegArena {
    void* dev = nullptr;
    size_t devBytes = 0;
    void* host = nullptr;  // pinned
    size_t hostBytes = 0;
    size_t counterWords = 0;  // zeroed counter words at the end of dev
};

// Everything one segmented sort launches, plus its captured graph.
struct SegPlanData {
    sortEngine_d* engine = nullptr;
    int valueCount = 0;
    int firstBit = 0;
    int numBits = 0;
    int sortPasses = 0;  // passes sortArray would run: fixes the result side
    int sortWidths[MGPU_MAX_PASSES] = {};
    int resultSide = 0;
    std::vector<mgpuSegDesc> hostSegs;
    std::vector<SegLaunch> launches;
    std::vector<uint32_t> wide;  // segments sorted by the device-wide path
    const PassVariant* wideVariant = nullptr;  // tile shape of the device-wide passes
    uint32_t wideTiles = 0;      // tiles of all device-wide segments
    uint32_t wideChunks = 0;     // digit_count chunks of them
    mgpuSegDesc* devWide = nullptr;
    uint32_t* devTileBegin = nullptr;
    uint32_t* devTileSeg = nullptr;  // [wideTiles]: the device-wide segment of each tile
    uint32_t* devChunks = nullptr;
    void* devBlock = nullptr;  // owned allocation (null when an arena holds the arrays)
    mgpuSegDesc* devSegs = nullptr;
    uint32_t* devLists = nullptr;
    uint32_t* devCounters = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
};

sortStatus_t mgpu_seg_load(sortEngine_d* e);
void mgpu_seg_unload(sortEngine_d* e);
// Plans one segmented sort.  Without an arena the device arrays get an
// allocation of their own (synchronous copies); with one they are copied
// asynchronously on lane.stream into the arena, which must not be in use by
// earlier work still running on another stream.
sortStatus_t mgpu_seg_build(sortEngine_d* e, SortLane& lane, const mgpuSegDesc* segs,
                            int numSegments, int valueCount, int firstBit, int endBit,
                            SegPlanData* out, SegArena* arena = nullptr);
void mgpu_arena_free(SegArena* a);
// Grows an arena (synchronizing `stream` first, whose work may still read the
// old one) to hold `payload` bytes of plan arrays and `counterWords` queue
// counters; does nothing when it already does.
sortStatus_t mgpu_arena_reserve(SegArena* arena, size_t payload, size_t counterWords,
                                hipStream_t stream);
// Upper bound on the plan arrays of a build over `numSegments` segments
// holding `keys` keys in total, in tiles of at least `minTile` keys.
size_t mgpu_seg_payload_bound(uint32_t numSegments, uint64_t keys, uint32_t minTile);
constexpr size_t kSegCounterWords = 4 * kSegVariants;
// The lane's side stream and fork/join events, on which the persistent
// kernels of a plan run beside its device-wide passes.
sortStatus_t mgpu_lane_ensure_side(SortLane& lane);
sortStatus_t mgpu_seg_enqueue(SegPlanData& P, SortLane& lane);
void mgpu_seg_release(SegPlanData& P);
