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
// {8, 16} x workgroup size {128 (up to 7 bits), 256, 512}, named
// mgpu_segsort_r<bits>_i<items>_b<threads>.
constexpr int kSegMinBits = 4;
constexpr int kSegMaxBits = 8;
constexpr int kSegBlocks = 3;
constexpr int kSegVariants = (kSegMaxBits - kSegMinBits + 1) * 2 * kSegBlocks;

inline int seg_block_index(int block) { return block == 128 ? 0 : block == 512 ? 2 : 1; }
inline int seg_block_threads(int b) { return b == 0 ? 128 : b == 2 ? 512 : 256; }
inline bool seg_variant_valid(int rb, int ipt, int block) {
    return rb >= kSegMinBits && rb <= kSegMaxBits && (ipt == 8 || ipt == 16) &&
           (block == 128 || block == 256 || block == 512) && block >= (1 << rb);
}
inline int seg_variant_index(int rb, int ipt, int block) {
    return ((rb - kSegMinBits) * 2 + (ipt == 16)) * kSegBlocks + seg_block_index(block);
}
inline int seg_variant_bits(int v) { return kSegMinBits + v / (2 * kSegBlocks); }
inline int seg_variant_ipt(int v) { return ((v / kSegBlocks) & 1) ? 16 : 8; }
inline int seg_variant_block(int v) { return seg_block_threads(v % kSegBlocks); }

// One row of the tuning table: segments with ceil(log2 n) <= maxLog2 use
// this variant (rows sorted by maxLog2).
struct SegTuneEntry {
    int maxLog2;
    int radixBits;
    int itemsPerThread;
    int blockThreads;
};

// The device-wide passes of a segmented sort: segments of at least minKeys
// keys take them; digit_count reads pieces of at most chunk keys; tile shape
// (0: auto, 1 small, 2 large); packed tiles (-1: when the segments have one
// chain each, 0 never, 1 always).
struct SegWideTune {
    uint32_t minKeys = 4096 + 1;
    uint32_t chunk = 1u << 14;
    int tile = 0;
    int pack = -1;
};

// One persistent launch of one kernel variant.
struct SegLaunch {
    int variant = 0;
    hipFunction_t fn = nullptr;
    uint32_t grid = 0;
    uint32_t block = 256;
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
    // Packed tiles (mgpu_multi_packed_*): the segments back to back in wave
    // chunks; devTileBegin then holds each segment's first chunk and
    // devTileSeg two words per tile (its first segment; part starts, chunks
    // holding keys and whether it continues a segment).
    bool widePacked = false;
    // The count kernel reads every device-wide segment as one piece and
    // writes its digit starts (no global histogram, no scan launch).
    bool wideDirectCount = false;
    mgpuSegDesc* devWide = nullptr;
    uint32_t* devTileBegin = nullptr;  // [segments + 1]: first tile of each (packed: first chunk)
    uint32_t* devTileSeg = nullptr;    // [wideTiles]: the segment of each tile (packed: tile words)
    uint32_t* devChunks = nullptr;
    // Lookback chains per XCD (wideXcds > 1): the device-wide segments are
    // grouped by XCD, XCD x holding global tiles [xcdTileBegin[x],
    // xcdTileBegin[x + 1]).
    uint32_t wideXcds = 1;
    // One chain per segment (wideXcds == 1): groups of segments whose tiles
    // have a counter each (wideGroups > 1), laid out the same way.
    uint32_t wideGroups = 1;
    std::vector<uint32_t> xcdTileBegin;
    uint32_t* devXcdTileBegin = nullptr;
    void* devBlock = nullptr;  // owned allocation (null when an arena holds the arrays)
    mgpuSegDesc* devSegs = nullptr;
    uint32_t* devLists = nullptr;
    uint32_t* devCounters = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    // Scratch of the device-wide passes: the lane's (grow-only, so a later
    // call on the lane may free and replace it) unless the plan owns its
    // own, which a captured plan does: its graph's kernels address it.
    uint32_t* ownLookback = nullptr;  // [passes][wideTiles][MGPU_RADIX]
    uint32_t* ownMulti = nullptr;     // hist then digit starts, [wide][4][MGPU_RADIX] each
};

sortStatus_t mgpu_seg_load(sortEngine_d* e);
void mgpu_seg_unload(sortEngine_d* e);
// Plans one segmented sort.  Without an arena the device arrays get an
// allocation of their own (synchronous copies); with one they are copied
// asynchronously on lane.stream into the arena, which must not be in use by
// earlier work still running on another stream.
// ownScratch: the plan allocates the scratch of its device-wide passes
// instead of using the lane's (plans that are captured and replayed).
sortStatus_t mgpu_seg_build(sortEngine_d* e, SortLane& lane, const mgpuSegDesc* segs,
                            int numSegments, int valueCount, int firstBit, int endBit,
                            SegPlanData* out, SegArena* arena = nullptr,
                            bool ownScratch = false);
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
