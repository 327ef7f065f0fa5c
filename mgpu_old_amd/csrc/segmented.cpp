// Segmented sort host side: work lists, size-bucket tuning table, one
// persistent launch per kernel variant, hipGraph capture for replay.
//
// No reference counterpart (the reference sorts one array per call,
// sort/src/mgpusort/mgpusort.cpp:607-616); each segment's result is defined
// as what sortArray would leave for it: stable by the bit field, values
// following, on side (passes & 1) of its buffers with passes =
// sortPlanPasses(endBit - firstBit).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "engine_internal.h"
#include "segmented_internal.h"

namespace {

// Segments whose device-wide tiles average at least this many take the
// lookback chains per XCD (measured, key + value, replays: 64 segments
// log-uniform in [2^18, 2^20] 0.734 -> 0.640 ms, 512 in [2^15, 2^17] 0.847 ->
// 0.764 ms; 4096 in [1, 2^16], 2.7 tiles each, 0.673 -> 0.729 ms).
constexpr uint64_t kXcdMinTiles = 4;

// Tile counters of a one-chain-per-segment launch (groups of segments).
constexpr long kTileCounterGroups = 8;

// Longest segment the device-wide count reads with one workgroup, writing
// its digit starts directly (a 2^18-key segment: 1 MB, a few microseconds).
constexpr uint32_t kDirectCountMax = 1u << 18;

// Segments of at least wide.minKeys keys go to the device-wide passes (all
// of them together, each spread over many workgroups with a lookback chain)
// instead of one workgroup each.  Above 4096 keys (the largest a 256-thread
// workgroup sorts in LDS) one workgroup would stream the segment through
// global memory tile by tile, bound by memory latency: a 64K-key segment took
// up to 1.3 ms, and a batch of a few such segments left the GPU idle.
// Measured: 4096 segments log-uniform in [1, 2^16] in 2.04 ms with the
// threshold above 2^16, 0.74 ms above 4096.  The tuning table's "wide" row
// sets it; MGPUSORT_SEG_WIDE_MIN overrides both (measurements).
uint32_t device_wide_min(const sortEngine_d* e) {
    const char* v = std::getenv("MGPUSORT_SEG_WIDE_MIN");
    const long x = v ? std::strtol(v, nullptr, 10) : 0;
    return x > 0 ? (uint32_t)x : e->segWide.minKeys;
}
// Keys per digit_count workgroup on that path (a multiple of 1024 up to 2^20,
// where the 16-bit lane-column counts of the 1024-thread count kernel would
// start to overflow; MGPUSORT_SEG_WIDE_CHUNK overrides the table).
uint32_t clamp_chunk(long x) { return (uint32_t)std::min<long>(std::max<long>(x, 1024), 1L << 20) & ~1023u; }
uint32_t wide_chunk(const sortEngine_d* e) {
    const char* v = std::getenv("MGPUSORT_SEG_WIDE_CHUNK");
    const long x = v ? std::strtol(v, nullptr, 10) : 0;
    return x >= 1024 ? clamp_chunk(x) : e->segWide.chunk;
}
// Tile shape of the device-wide passes: the table's, unless
// MGPUSORT_SEG_WIDE_TILE=auto|small|large overrides it (measurements).
int wide_tile(const sortEngine_d* e) {
    const char* v = std::getenv("MGPUSORT_SEG_WIDE_TILE");
    if (v && std::strcmp(v, "auto") == 0) return 0;
    if (v && std::strcmp(v, "small") == 0) return 1;
    if (v && std::strcmp(v, "large") == 0) return 2;
    return e->segWide.tile;
}
// Packed tiles for the device-wide passes (segments back to back in wave
// chunks): the table's choice, by default when the segments have one
// lookback chain each (short chains, where a tile per segment leaves most
// last tiles part empty); MGPUSORT_SEG_PACK=1 packs always (chains per XCD
// too), =0 never.
int pack_mode(const sortEngine_d* e) {
    const char* v = std::getenv("MGPUSORT_SEG_PACK");
    if (v && v[0] == '0' && v[1] == 0) return 0;
    if (v && v[0] == '1' && v[1] == 0) return 1;
    return e->segWide.pack;
}
// Count kernel of the device-wide passes: lane-column histograms in
// 1024-thread workgroups (default), or per-wave histograms in 256-thread
// workgroups with MGPUSORT_SEG_COUNT=wave.  Measured by plan replays, key +
// value, per-wave -> lane-column: 4096 segments log-uniform in [1, 2^16]
// 0.674 -> 0.666 ms; 512 in [2^15, 2^17] 0.841 -> 0.781 ms with one chain per
// segment and 0.755 -> 0.695 ms with chains per XCD; 64 in [2^18, 2^20]
// 0.639 -> 0.645 ms (chains per XCD).
bool count_cols() {
    const char* v = std::getenv("MGPUSORT_SEG_COUNT");
    return !(v && std::strcmp(v, "wave") == 0);
}

// Workgroups per CU of a persistent kernel that runs beside device-wide
// passes: one (at most its occupancy), so that every CU keeps room for the
// count and pass workgroups, which are on the critical path, instead of the
// persistent grid filling the device first.  Measured by plan replays, key +
// value, occupancy -> one per CU: 4096 segments log-uniform in [1, 2^16]
// 0.585 -> 0.577 ms; 2000 in [2^10, 2^14] 0.284 -> 0.266 ms; 8000 in [2^8,
// 2^15] 1.21 -> 1.17 ms; 16000 in [1, 2^13] 0.415 -> 0.346 ms; 500 in [2^4,
// 2^17] 0.255 ms either way (two per CU: as occupancy).
// MGPUSORT_SEG_MIX_WGS=<n> overrides it (measurements).
int mixed_workgroups_per_cu() {
    const char* v = std::getenv("MGPUSORT_SEG_MIX_WGS");
    const long x = v ? std::strtol(v, nullptr, 10) : 0;
    return x > 0 ? (int)std::min<long>(x, 64) : 1;
}

const SegTuneEntry kDefaultTable[] = {
    // max log2(n), radix bits, items per thread, workgroup size
    {6, 8, 8, 256}, {10, 8, 16, 256}, {12, 8, 16, 256}, {31, 8, 16, 256},
};

// "r<bits>i<items>" (256 threads) or "r<bits>i<items>b<threads>".
bool parse_variant(const char* s, int* rb, int* ipt, int* block) {
    if (!s) return false;
    *block = 256;
    const int got = std::sscanf(s, "r%di%db%d", rb, ipt, block);
    return got >= 2 && seg_variant_valid(*rb, *ipt, *block);
}

}  // namespace

sortStatus_t mgpu_seg_load(sortEngine_d* e) {
    std::string path = e->kernelPath + "mgpusort_seg.hsaco";
    if (hipModuleLoad(&e->segModule, path.c_str()) != hipSuccess) {
        e->segModule = nullptr;
        return SORT_STATUS_KERNEL_NOT_FOUND;
    }
    for (int v = 0; v < kSegVariants; ++v) {
        const int rb = seg_variant_bits(v), ipt = seg_variant_ipt(v), block = seg_variant_block(v);
        if (!seg_variant_valid(rb, ipt, block)) continue;
        for (int tileOnly = 0; tileOnly < 2; ++tileOnly) {
            char name[64];
            std::snprintf(name, sizeof(name), "mgpu_segsort_r%d_i%d_b%d%s", rb, ipt, block,
                          tileOnly ? "_ng" : "");
            hipFunction_t f = nullptr;
            if (hipModuleGetFunction(&f, e->segModule, name) != hipSuccess)
                return SORT_STATUS_KERNEL_ERROR;
            e->kernelNames[f] = name;
            int blocks = 0;
            if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, block, 0) !=
                    hipSuccess || blocks < 1)
                blocks = 1;
            (tileOnly ? e->segFnTile : e->segFn)[v] = f;
            (tileOnly ? e->segOccupancyTile : e->segOccupancy)[v] = blocks;
        }
    }

    // Tuning table: kernelPath + "mgpusort_seg_tuning.txt" when present
    // (written by the autotuner), else the built-in default.  Rows
    // "<max ceil(log2 n)> <radix bits> <keys per thread> [<threads>]" pick the
    // persistent kernel per size bucket; a row "wide <min keys> <chunk>
    // <auto|small|large>" sets the device-wide passes.
    e->segTable.assign(std::begin(kDefaultTable), std::end(kDefaultTable));
    e->segWide = SegWideTune();
    std::ifstream f(e->kernelPath + "mgpusort_seg_tuning.txt");
    if (f) {
        std::vector<SegTuneEntry> t;
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream is(line);
            if (line.compare(0, 5, "wide ") == 0) {
                std::string w, tile, pack;
                long minKeys = 0, chunk = 0;
                if (is >> w >> minKeys >> chunk >> tile && minKeys > 0 && chunk >= 1024 &&
                    (tile == "auto" || tile == "small" || tile == "large")) {
                    if (!(is >> pack)) pack = "auto";
                    if (pack == "auto" || pack == "packed" || pack == "unpacked") {
                        e->segWide.minKeys = (uint32_t)minKeys;
                        e->segWide.chunk = clamp_chunk(chunk);
                        e->segWide.tile = tile == "small" ? 1 : tile == "large" ? 2 : 0;
                        e->segWide.pack = pack == "packed" ? 1 : pack == "unpacked" ? 0 : -1;
                    }
                }
                continue;
            }
            SegTuneEntry en;
            if (is >> en.maxLog2 >> en.radixBits >> en.itemsPerThread) {
                if (!(is >> en.blockThreads)) en.blockThreads = 256;
                if (!seg_variant_valid(en.radixBits, en.itemsPerThread, en.blockThreads)) continue;
                t.push_back(en);
            }
        }
        if (!t.empty()) {
            std::sort(t.begin(), t.end(), [](const SegTuneEntry& a, const SegTuneEntry& b) {
                return a.maxLog2 < b.maxLog2;
            });
            t.back().maxLog2 = 31;
            e->segTable = t;
        }
    }
    return SORT_STATUS_SUCCESS;
}

void mgpu_seg_unload(sortEngine_d* e) {
    if (e->segModule) (void)hipModuleUnload(e->segModule);
    e->segModule = nullptr;
}

namespace {

int ceil_log2(uint32_t n) {
    int l = 0;
    while ((1ull << l) < n) ++l;
    return l;
}

void pick_variant(const sortEngine_d* e, uint32_t n, int* rb, int* ipt, int* block) {
    const char* forced = std::getenv("MGPUSORT_SEG_VARIANT");
    if (parse_variant(forced, rb, ipt, block)) return;
    int l = ceil_log2(n);
    for (const SegTuneEntry& t : e->segTable)
        if (l <= t.maxLog2) {
            *rb = t.radixBits;
            *ipt = t.itemsPerThread;
            *block = t.blockThreads;
            return;
        }
    *rb = 8;
    *ipt = 16;
    *block = 256;
}

}  // namespace

// Builds the launches of one segmented sort.  Device-side arrays are owned by
// the plan; nothing is launched here.
sortStatus_t mgpu_lane_ensure_side(SortLane& lane) {
    if (lane.side) return SORT_STATUS_SUCCESS;
    // Default priority, as the lane's own stream.
    if (hipStreamCreateWithFlags(&lane.side, hipStreamNonBlocking) != hipSuccess) {
        lane.side = nullptr;
        return SORT_STATUS_DEVICE_ERROR;
    }
    if (hipEventCreateWithFlags(&lane.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&lane.join, hipEventDisableTiming) != hipSuccess)
        return SORT_STATUS_DEVICE_ERROR;
    return SORT_STATUS_SUCCESS;
}

size_t mgpu_seg_payload_bound(uint32_t numSegments, uint64_t keys, uint32_t minTile) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t segs = std::max<size_t>(1, numSegments);
    const size_t chunks = keys / 1024 + segs;  // pieces of at least 1024 keys
    // Tiles: one per minTile keys plus a partial one per segment (packed:
    // plus padding to a tile boundary per XCD), two words each when packed.
    const size_t tiles = keys / std::max(1u, minTile) + segs + MGPU_MAX_SEGS;
    return 2 * al(sizeof(mgpuSegDesc) * segs) + al(sizeof(uint32_t) * segs) +
           al(sizeof(uint32_t) * (segs + 1)) + al(4 * sizeof(uint32_t) * chunks) +
           al(2 * sizeof(uint32_t) * tiles) + al(sizeof(uint32_t) * (MGPU_MAX_SEGS + 1));
}

sortStatus_t mgpu_arena_reserve(SegArena* arena, size_t payload, size_t counterWords,
                                hipStream_t stream) {
    // Layout: [payload capacity = hostBytes][counterWords counters].  The
    // counters are only ever left zero by the kernels, so they are cleared
    // once, here.
    if (arena->hostBytes >= payload && arena->counterWords >= counterWords)
        return SORT_STATUS_SUCCESS;
    if (hipStreamSynchronize(stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
    mgpu_arena_free(arena);
    const size_t payloadCap = 2 * payload + 4096;
    const size_t words = 2 * std::max<size_t>(counterWords, 64);
    const size_t cap = payloadCap + words * sizeof(uint32_t);
    if (hipMalloc(&arena->dev, cap) != hipSuccess) {
        arena->dev = nullptr;
        return SORT_STATUS_DEVICE_ALLOC_FAILED;
    }
    if (hipHostMalloc(&arena->host, payloadCap, hipHostMallocDefault) != hipSuccess) {
        arena->host = nullptr;
        mgpu_arena_free(arena);
        return SORT_STATUS_HOST_ALLOC_FAILED;
    }
    arena->devBytes = cap;
    arena->hostBytes = payloadCap;
    arena->counterWords = words;
    // On the stream, ahead of the descriptor copies that follow: a
    // null-stream fill could land after such a copy on a non-blocking stream
    // and zero the descriptors the kernels then follow.
    if (hipMemsetAsync(arena->dev, 0, cap, stream) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t mgpu_seg_build(sortEngine_d* e, SortLane& lane, const mgpuSegDesc* segs,
                            int numSegments, int valueCount, int firstBit, int endBit,
                            SegPlanData* out, SegArena* arena, bool ownScratch) {
    SegPlanData& P = *out;
    P.engine = e;
    P.valueCount = valueCount;
    P.firstBit = firstBit;
    P.numBits = endBit - firstBit;
    P.sortPasses = mgpu_plan_passes(P.numBits, MGPU_RADIX_BITS, P.sortWidths);
    P.resultSide = P.sortPasses & 1;
    P.hostSegs.assign(segs, segs + numSegments);

    // Group segments by kernel variant; very large ones go device-wide.
    const uint32_t wideMin = device_wide_min(e), chunkKeys = wide_chunk(e);
    struct Group {
        std::vector<uint32_t> big, small;
    };
    Group groups[kSegVariants];
    for (int i = 0; i < numSegments; ++i) {
        const uint32_t n = segs[i].n;
        if (n == 0 || P.numBits == 0) continue;
        if (n >= wideMin) {
            P.wide.push_back((uint32_t)i);
            continue;
        }
        int rb, ipt, block;
        pick_variant(e, n, &rb, &ipt, &block);
        Group& g = groups[seg_variant_index(rb, ipt, block)];
        if (n <= 64u * (uint32_t)ipt) g.small.push_back((uint32_t)i);
        else g.big.push_back((uint32_t)i);
    }

    // Persistent kernels forked beside device-wide passes take at most
    // mixed_workgroups_per_cu() workgroups per CU.
    const bool mixed = !P.wide.empty() && !lane.noSide;
    const int mixedWgs = mixed_workgroups_per_cu();

    // Device arrays: descriptors, then per group [big list | small list],
    // then 4 counters per group.
    std::vector<uint32_t> lists;
    for (int v = 0; v < kSegVariants; ++v) {
        Group& g = groups[v];
        if (g.big.empty() && g.small.empty()) continue;
        std::stable_sort(g.big.begin(), g.big.end(),
                         [&](uint32_t a, uint32_t b) { return segs[a].n > segs[b].n; });
        SegLaunch L;
        L.variant = v;
        L.bigOffset = (uint32_t)lists.size();
        lists.insert(lists.end(), g.big.begin(), g.big.end());
        L.smallOffset = (uint32_t)lists.size();
        lists.insert(lists.end(), g.small.begin(), g.small.end());
        L.numBig = (uint32_t)g.big.size();
        L.numSmall = (uint32_t)g.small.size();
        const int rb = seg_variant_bits(v);
        // The kernel without the global-memory routine when every block-list
        // segment fits a workgroup's tile (the list is largest first).
        const uint32_t tileKeys = (uint32_t)(seg_variant_block(v) * seg_variant_ipt(v));
        const bool tileOnly = g.big.empty() || segs[g.big.front()].n <= tileKeys;
        L.fn = tileOnly ? e->segFnTile[v] : e->segFn[v];
        L.block = (uint32_t)seg_variant_block(v);
        const uint32_t wavesPerGroup = L.block / 64;
        const uint32_t work = L.numBig + (L.numSmall + wavesPerGroup - 1) / wavesPerGroup;
        int occ = tileOnly ? e->segOccupancyTile[v] : e->segOccupancy[v];
        if (mixed) occ = std::min(occ, mixedWgs);
        const uint32_t maxGrid = (uint32_t)(occ * e->numCUs);
        L.grid = std::max(1u, std::min(work, maxGrid));
        L.passes = mgpu_plan_passes(P.numBits, rb, L.widths);
        P.launches.push_back(L);
    }

    // Device-wide segments: their descriptors, their tile layout, and the
    // digit_count chunks.
    std::vector<mgpuSegDesc> wideDescs;
    std::vector<uint32_t> tileBegin, chunks, tileSeg;
    P.widePacked = false;
    if (!P.wide.empty()) {
        uint64_t wideKeys = 0;
        for (uint32_t k : P.wide) wideKeys += segs[k].n;
        // Packed tiles leave no partial tile per segment, so they take the
        // shape sortArray would take for all of the keys.  Without packing,
        // each segment's last tile is partial, so segments of a few tiles
        // take the small tile shape (4096 segments log-uniform in [1, 2^16]:
        // 0.74 -> 0.69 ms); the large shape once they average 4 large tiles.
        const int packMode = pack_mode(e);
        const bool packBuilt = e->large.packed[0] != nullptr && e->small.packed[0] != nullptr;
        const uint64_t avgKeys = wideKeys / P.wide.size();
        const int tileChoice = wide_tile(e);
        if (tileChoice == 1) P.wideVariant = &e->small;
        else if (tileChoice == 2) P.wideVariant = &e->large;
        else if (packBuilt && packMode != 0) P.wideVariant = &mgpu_pick_variant(e, wideKeys);
        else P.wideVariant = avgKeys >= 4ull * e->large.tile() ? &mgpu_pick_variant(e, wideKeys) : &e->small;
        const uint32_t tile = P.wideVariant->tile();
        const uint32_t waves = P.wideVariant->blockThreads / 64;
        const uint32_t chunkLen = 64 * P.wideVariant->itemsPerThread;  // packed: keys per wave chunk
        const uint32_t unit = tile;  // work of a segment for the XCD choice: its tiles
        // Lookback chains per XCD: each segment's tiles on one XCD, whose L2
        // alone holds its chain (~0.45 us a round trip instead of ~0.85 us
        // through the device-coherent level).  Segments go to the XCD with
        // the least work so far, largest first, and are then laid out XCD
        // by XCD.  Only for segments of longer chains (at least kXcdMinTiles
        // tiles on average: short chains wait little, and the workgroups
        // that loop over an XCD's tiles pay a counter round trip and a
        // barrier per tile), and when no segment outweighs a quarter of an
        // XCD's share (else one XCD would do most of the work).
        // MGPUSORT_SEG_XCD=0|1 forces the choice.
        P.wideXcds = 1;
        P.xcdTileBegin.clear();
        std::vector<size_t> groupEnd(1, P.wide.size());  // end of each XCD's segments
        // Segments into G groups, largest first to the group with the least
        // work, laid out group by group (P.wide reordered, groupEnd and the
        // groups' first tiles filled in).
        auto group_by_load = [&](uint32_t G) {
            std::vector<uint32_t> order(P.wide);
            std::stable_sort(order.begin(), order.end(),
                             [&](uint32_t a, uint32_t b) { return segs[a].n > segs[b].n; });
            std::vector<std::vector<uint32_t>> group(G);
            std::vector<uint64_t> load(G, 0);
            for (uint32_t k : order) {
                const uint32_t x =
                    (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
                group[x].push_back(k);
                load[x] += (segs[k].n + unit - 1) / unit;
            }
            P.wide.clear();
            groupEnd.clear();
            P.xcdTileBegin.assign(1, 0);
            for (uint32_t x = 0; x < G; ++x) {
                P.wide.insert(P.wide.end(), group[x].begin(), group[x].end());
                groupEnd.push_back(P.wide.size());
                P.xcdTileBegin.push_back(P.xcdTileBegin.back() + (uint32_t)load[x]);
            }
        };
        const uint32_t X = e->xcdSegs;
        const bool xcdKernels = P.wideVariant->multiX[MGPU_VALUES_NONE] != nullptr &&
                                (packMode != 1 || P.wideVariant->packedX[MGPU_VALUES_NONE] != nullptr);
        if (X > 1 && xcdKernels) {
            uint64_t total = 0, most = 0;
            for (uint32_t k : P.wide) {
                const uint64_t t = (segs[k].n + unit - 1) / unit;
                total += t;
                most = std::max(most, t);
            }
            const char* force = std::getenv("MGPUSORT_SEG_XCD");
            bool use = most * 4 * X <= total && total >= kXcdMinTiles * P.wide.size();
            if (force && (force[0] == '0' || force[0] == '1') && force[1] == 0)
                use = force[0] == '1' && most * 4 * X <= total;
            if (use) {
                group_by_load(X);
                P.wideXcds = X;
            }
        }
        // One chain per segment: the segments still go into groups (largest
        // first, to the least loaded), each group's tiles handed out by a
        // counter of its own (MGPUSORT_SEG_GROUPS, default 8; 1 = one
        // counter for every tile).
        P.wideGroups = 1;
        if (P.wideXcds == 1 && P.wide.size() > 1) {
            const char* v = std::getenv("MGPUSORT_SEG_GROUPS");
            const long want = v ? std::strtol(v, nullptr, 10) : kTileCounterGroups;
            const uint32_t G = (uint32_t)std::max<long>(
                1, std::min<long>({want, (long)MGPU_MAX_SEGS, (long)P.wide.size()}));
            if (G > 1) {
                group_by_load(G);
                P.wideGroups = G;
            }
        }
        P.widePacked = P.wideVariant->packed[MGPU_VALUES_NONE] != nullptr &&
                       (packMode == 1 || (packMode == -1 && P.wideXcds == 1));
        // The count reads each segment as one piece and writes its digit
        // starts itself when no segment is longer than kDirectCountMax keys
        // (one workgroup per segment stays short) and no piece size was
        // asked for; else pieces of chunkKeys with a histogram in global
        // memory and a scan launch after them.
        uint32_t longest = 0;
        for (uint32_t k : P.wide) longest = std::max(longest, segs[k].n);
        P.wideDirectCount = count_cols() && longest <= kDirectCountMax &&
                            std::getenv("MGPUSORT_SEG_WIDE_CHUNK") == nullptr;
        const uint32_t pieceKeys = P.wideDirectCount ? longest : chunkKeys;
        // Pieces longest first: workgroups start in about launch order, so
        // the longest counts start first and the short ones fill in behind
        // them (greedy longest-processing-time order).  4096 segments
        // log-uniform in [1, 2^16]: count of the 1,007 device-wide ones
        // 61 -> 51 us alone, replay 0.598 -> 0.586 ms.
        struct Piece { uint32_t seg, begin, end; };
        std::vector<Piece> pieces;
        for (size_t k = 0; k < P.wide.size(); ++k) {
            const mgpuSegDesc& d = segs[P.wide[k]];
            wideDescs.push_back(d);
            for (uint32_t b = 0; b < d.n; b += pieceKeys)
                pieces.push_back({(uint32_t)k, b, std::min(d.n, b + pieceKeys)});
        }
        std::stable_sort(pieces.begin(), pieces.end(), [](const Piece& x, const Piece& y) {
            return x.end - x.begin > y.end - y.begin;
        });
        for (const Piece& c : pieces) {
            chunks.push_back(c.seg);
            chunks.push_back(c.begin);
            chunks.push_back(c.end);
            chunks.push_back(0);
        }
        if (P.widePacked) {
            // Wave chunks back to back, `waves` to a tile.  A segment that
            // would be the (max parts + 1)-th part of a tile starts the next
            // tile instead, and each XCD's segments end on a tile boundary:
            // padding chunks only ever end a tile.
            const uint32_t maxParts = MGPU_PACK_MAX_PARTS(waves);
            std::vector<uint64_t> begin(P.wide.size());
            auto chunksOf = [&](size_t k) { return ((uint64_t)segs[P.wide[k]].n + chunkLen - 1) / chunkLen; };
            uint64_t cpos = 0;
            uint32_t parts = 0;  // parts of the tile holding chunk cpos
            std::vector<uint32_t> xb(1, 0);
            size_t k = 0;
            for (size_t g = 0; g < groupEnd.size(); ++g) {
                for (; k < groupEnd[g]; ++k) {
                    if (cpos % waves != 0 && parts >= maxParts) cpos += waves - cpos % waves;
                    if (cpos % waves == 0) parts = 0;
                    begin[k] = cpos;
                    const uint64_t end = cpos + chunksOf(k);
                    if (end / waves == cpos / waves) ++parts;
                    else parts = end % waves != 0 ? 1u : 0u;
                    cpos = end;
                }
                if (cpos % waves != 0) cpos += waves - cpos % waves;
                parts = 0;
                xb.push_back((uint32_t)(cpos / waves));
            }
            const uint64_t tiles = cpos / waves;
            if (cpos >= (1ull << 32)) return SORT_STATUS_INVALID_VALUE;
            tileBegin.resize(P.wide.size());
            for (size_t i = 0; i < P.wide.size(); ++i) tileBegin[i] = (uint32_t)begin[i];
            tileSeg.assign(2 * tiles, 0);
            size_t j = 0;
            for (uint64_t t = 0; t < tiles; ++t) {
                const uint64_t c0 = t * waves;
                while (j + 1 < P.wide.size() && begin[j + 1] <= c0) ++j;
                uint32_t startMask = 0, live = 0;
                for (size_t i = j; i < P.wide.size() && begin[i] < c0 + waves; ++i) {
                    if (begin[i] > c0) startMask |= 1u << (uint32_t)(begin[i] - c0);
                    live = (uint32_t)std::min<uint64_t>(waves, begin[i] + chunksOf(i) - c0);
                }
                tileSeg[2 * t] = (uint32_t)j;
                tileSeg[2 * t + 1] = startMask | (live << 16) | ((begin[j] < c0 ? 1u : 0u) << 24);
            }
            P.wideTiles = (uint32_t)tiles;
            if (P.wideXcds > 1 || P.wideGroups > 1) P.xcdTileBegin = xb;
        } else {
            tileBegin.push_back(0);
            for (size_t k = 0; k < P.wide.size(); ++k) {
                tileBegin.push_back(tileBegin.back() + (segs[P.wide[k]].n + tile - 1) / tile);
                tileSeg.resize(tileBegin.back(), (uint32_t)k);
            }
            P.wideTiles = tileBegin.back();
        }
        P.wideChunks = (uint32_t)(chunks.size() / 4);
    }
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t descBytes = al(sizeof(mgpuSegDesc) * (size_t)std::max(1, numSegments));
    const size_t listBytes = al(sizeof(uint32_t) * std::max<size_t>(1, lists.size()));
    const size_t wideBytes = al(sizeof(mgpuSegDesc) * wideDescs.size());
    const size_t tbBytes = al(sizeof(uint32_t) * tileBegin.size());
    const size_t chBytes = al(sizeof(uint32_t) * chunks.size());
    const size_t tsBytes = al(sizeof(uint32_t) * tileSeg.size());
    const size_t xbBytes = al(sizeof(uint32_t) * P.xcdTileBegin.size());
    const size_t ctrWords = 4 * std::max<size_t>(1, P.launches.size());
    const size_t payload = descBytes + listBytes + wideBytes + tbBytes + chBytes + tsBytes + xbBytes;
    const size_t total = payload + ctrWords * sizeof(uint32_t);
    char* base = nullptr;
    if (!arena) {
        if (hipMalloc(&P.devBlock, total) != hipSuccess) {
            P.devBlock = nullptr;
            return SORT_STATUS_DEVICE_ALLOC_FAILED;
        }
        base = static_cast<char*>(P.devBlock);
        if (numSegments && hipMemcpy(base, segs, sizeof(mgpuSegDesc) * numSegments,
                                     hipMemcpyHostToDevice) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        if (!lists.empty() && hipMemcpy(base + descBytes, lists.data(),
                                        sizeof(uint32_t) * lists.size(),
                                        hipMemcpyHostToDevice) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        size_t off = descBytes + listBytes;
        if (!wideDescs.empty() &&
            (hipMemcpy(base + off, wideDescs.data(), sizeof(mgpuSegDesc) * wideDescs.size(),
                       hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(base + off + wideBytes, tileBegin.data(), sizeof(uint32_t) * tileBegin.size(),
                       hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(base + off + wideBytes + tbBytes, chunks.data(),
                       sizeof(uint32_t) * chunks.size(), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(base + off + wideBytes + tbBytes + chBytes, tileSeg.data(),
                       sizeof(uint32_t) * tileSeg.size(), hipMemcpyHostToDevice) != hipSuccess))
            return SORT_STATUS_DEVICE_ERROR;
        if (!P.xcdTileBegin.empty() &&
            hipMemcpy(base + off + wideBytes + tbBytes + chBytes + tsBytes, P.xcdTileBegin.data(),
                      sizeof(uint32_t) * P.xcdTileBegin.size(), hipMemcpyHostToDevice) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        // Stream-ordered before the launches on the lane's stream (a plain
        // hipMemset would run on the null stream, unordered with a
        // non-blocking lane stream).
        if (hipMemsetAsync(base + payload, 0, ctrWords * sizeof(uint32_t), lane.stream) !=
            hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
    } else {
        // Queue counters at the end of the arena, zero between builds.
        sortStatus_t st = mgpu_arena_reserve(arena, payload, ctrWords, lane.stream);
        if (st) return st;
        base = static_cast<char*>(arena->dev);
        char* h = static_cast<char*>(arena->host);
        if (numSegments) std::memcpy(h, segs, sizeof(mgpuSegDesc) * numSegments);
        if (!lists.empty()) std::memcpy(h + descBytes, lists.data(), sizeof(uint32_t) * lists.size());
        if (!wideDescs.empty()) {
            const size_t off = descBytes + listBytes;
            std::memcpy(h + off, wideDescs.data(), sizeof(mgpuSegDesc) * wideDescs.size());
            std::memcpy(h + off + wideBytes, tileBegin.data(), sizeof(uint32_t) * tileBegin.size());
            std::memcpy(h + off + wideBytes + tbBytes, chunks.data(), sizeof(uint32_t) * chunks.size());
            std::memcpy(h + off + wideBytes + tbBytes + chBytes, tileSeg.data(),
                        sizeof(uint32_t) * tileSeg.size());
            if (!P.xcdTileBegin.empty())
                std::memcpy(h + off + wideBytes + tbBytes + chBytes + tsBytes,
                            P.xcdTileBegin.data(), sizeof(uint32_t) * P.xcdTileBegin.size());
        }
        if (hipMemcpyAsync(base, h, payload, hipMemcpyHostToDevice, lane.stream) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        P.devCounters = reinterpret_cast<uint32_t*>(base + arena->hostBytes);
    }
    P.devSegs = reinterpret_cast<mgpuSegDesc*>(base);
    P.devLists = reinterpret_cast<uint32_t*>(base + descBytes);
    if (!arena) P.devCounters = reinterpret_cast<uint32_t*>(base + payload);
    P.devWide = reinterpret_cast<mgpuSegDesc*>(base + descBytes + listBytes);
    P.devTileBegin = reinterpret_cast<uint32_t*>(base + descBytes + listBytes + wideBytes);
    P.devChunks = reinterpret_cast<uint32_t*>(base + descBytes + listBytes + wideBytes + tbBytes);
    P.devTileSeg =
        reinterpret_cast<uint32_t*>(base + descBytes + listBytes + wideBytes + tbBytes + chBytes);
    P.devXcdTileBegin = reinterpret_cast<uint32_t*>(base + descBytes + listBytes + wideBytes +
                                                    tbBytes + chBytes + tsBytes);

    // Device-wide segments: lookback words and per-array scratch before any
    // capture.
    if (!P.wide.empty()) {
        const uint64_t lbWords = (uint64_t)P.sortPasses * P.wideTiles * MGPU_RADIX;
        sortStatus_t st = SORT_STATUS_SUCCESS;
        if (ownScratch) {
            // Histograms must start at zero (the scan leaves them zero after
            // every run); lookback words are zeroed by the count kernel.
            const size_t multiWords = (size_t)2 * P.wide.size() * MGPU_MAX_PASSES * MGPU_RADIX;
            if (hipMalloc(&P.ownLookback, ((lbWords + 3) & ~uint64_t(3)) * sizeof(uint32_t)) !=
                    hipSuccess ||
                hipMalloc(&P.ownMulti, multiWords * sizeof(uint32_t)) != hipSuccess)
                return SORT_STATUS_DEVICE_ALLOC_FAILED;
            if (hipMemsetAsync(P.ownMulti, 0, multiWords * sizeof(uint32_t) / 2, lane.stream) !=
                    hipSuccess ||
                hipStreamSynchronize(lane.stream) != hipSuccess)
                return SORT_STATUS_DEVICE_ERROR;
        } else {
            st = mgpu_ensure_lookback_words(lane, lbWords);
            if (!st) st = mgpu_ensure_multi(lane, (uint32_t)P.wide.size());
        }
        if (st) return st;
    }
    // The side stream (before any capture) when there is something to fork.
    if (!lane.noSide && !P.launches.empty() && !P.wide.empty()) {
        sortStatus_t st = mgpu_lane_ensure_side(lane);
        if (st) return st;
    }
    return SORT_STATUS_SUCCESS;
}

namespace {

// The device-wide passes over every large segment of the plan at once:
// digit_count over their chunks, digit_scan per segment, then one
// rank_scatter launch per pass over the tiles of all of them.
sortStatus_t mgpu_enqueue_wide(SegPlanData& P, SortLane& lane) {
    sortEngine_d* e = P.engine;
    const uint32_t np = (uint32_t)P.sortPasses;
    const uint64_t lbWords = (uint64_t)np * P.wideTiles * MGPU_RADIX;
    const uint32_t numWide = (uint32_t)P.wide.size();
    uint32_t* const lookback = P.ownLookback ? P.ownLookback : lane.lookback;
    uint32_t* const hist = P.ownMulti ? P.ownMulti : lane.multi;
    uint32_t* const digitStart =
        P.ownMulti ? P.ownMulti + (size_t)numWide * MGPU_MAX_PASSES * MGPU_RADIX
                   : lane.multi + (size_t)lane.multiArrays * MGPU_MAX_PASSES * MGPU_RADIX;

    mgpuMultiCountArgs ca;
    std::memset(&ca, 0, sizeof(ca));
    ca.segs = P.devWide;
    ca.chunks = P.devChunks;
    ca.hist = hist;
    ca.clear = lookback;
    ca.clear_words = (lbWords + 3) & ~uint64_t(3);
    ca.num_chunks = P.wideChunks;
    ca.num_passes = np;
    int bit = P.firstBit;
    for (uint32_t p = 0; p < np; ++p) {
        ca.shift[p] = (uint32_t)bit;
        ca.mask[p] = (1u << P.sortWidths[p]) - 1u;
        bit += P.sortWidths[p];
    }
    // One workgroup per piece, enough of them to zero the lookback words
    // quickly as well.
    const bool cols = count_cols();
    const uint32_t threads = cols ? 1024u : 256u;
    uint32_t grid = P.wideChunks;
    const uint32_t clearGrid = (uint32_t)std::min<uint64_t>(
        (ca.clear_words / 4 + threads * 16 - 1) / (threads * 16),
        (uint64_t)e->numCUs * (cols ? 2 : 8));
    grid = std::max(grid, clearGrid);
    if (!P.ownMulti) lane.multiDirty = true;
    ca.digit_start = digitStart;
    ca.tile_counters = lane.tileCounters;
    ca.direct = P.wideDirectCount ? 1u : 0u;
    sortStatus_t st = mgpu_launch(e, lane.stream, cols ? e->fnMultiCountCols : e->fnMultiCount,
                                  grid, threads, &ca);
    if (st) return st;

    if (!P.wideDirectCount) {
        mgpuMultiScanArgs sa;
        sa.hist = hist;
        sa.digit_start = digitStart;
        sa.tile_counters = lane.tileCounters;
        sa.num_passes = np;
        sa.pad_ = 0;
        st = mgpu_launch(e, lane.stream, e->fnMultiScan, numWide, 256, &sa);
        if (st) return st;
    }
    if (!P.ownMulti) lane.multiDirty = false;

    const int carried = P.valueCount == -1 ? 1 : P.valueCount;
    for (uint32_t p = 0; p < np; ++p) {
        mgpuMultiPassArgs pa;
        std::memset(&pa, 0, sizeof(pa));
        pa.segs = P.devWide;
        pa.tile_begin = P.devTileBegin;
        pa.tile_seg = P.devTileSeg;
        pa.digit_start = digitStart;
        pa.lookback = lookback + (uint64_t)p * P.wideTiles * MGPU_RADIX;
        pa.tile_counter = lane.tileCounters + MGPU_TILE_COUNTER(p, 0);
        pa.error_flag = lane.errorDev;
        pa.spin_limit = e->spinLimit;
        pa.num_segs = numWide;
        pa.num_tiles = P.wideTiles;
        pa.pass = p;
        pa.in_side = p & 1;
        pa.shift = ca.shift[p];
        pa.mask = ca.mask[p];
        pa.num_values = (uint32_t)carried;
        int mode;
        if (P.valueCount == 0) mode = MGPU_VALUES_NONE;
        else if (P.valueCount == -1) mode = p == 0 ? MGPU_VALUES_INDEX : MGPU_VALUES_SINGLE;
        else if (P.valueCount == 1) mode = MGPU_VALUES_SINGLE;
        else mode = MGPU_VALUES_MULTI;
        if (P.widePacked) {
            pa.chunk_begin = P.devTileBegin;
            pa.tile_info = P.devTileSeg;
            pa.tile_begin = nullptr;
            pa.tile_seg = nullptr;
        }
        if (P.wideXcds > 1) {
            // Workgroups that loop over their XCD's tiles: as many as fit.
            pa.xcd_tile_begin = P.devXcdTileBegin;
            pa.num_xcds = P.wideXcds;
            pa.finished = lane.tileCounters + MGPU_FINISHED_SLOT + p;
            const uint32_t fit =
                P.widePacked ? P.wideVariant->packedXGrid : P.wideVariant->multiXGrid;
            const uint32_t g = std::max(P.wideXcds, std::min(P.wideTiles, fit));
            st = mgpu_launch(e, lane.stream,
                             P.widePacked ? P.wideVariant->packedX[mode] : P.wideVariant->multiX[mode],
                             g, P.wideVariant->blockThreads, &pa);
        } else {
            if (P.wideGroups > 1) {
                pa.xcd_tile_begin = P.devXcdTileBegin;
                pa.num_xcds = P.wideGroups;
            }
            st = mgpu_launch(e, lane.stream,
                             P.widePacked ? P.wideVariant->packed[mode] : P.wideVariant->multi[mode],
                             P.wideTiles, P.wideVariant->blockThreads, &pa);
        }
        if (st) return st;
    }
    return SORT_STATUS_SUCCESS;
}

}  // namespace

// Queues the launches of a built plan on the lane's stream.
sortStatus_t mgpu_seg_enqueue(SegPlanData& P, SortLane& lane) {
    sortEngine_d* e = P.engine;
    // The persistent kernels and the device-wide passes sort different
    // segments with separate scratch: with both, the former run on the
    // lane's side stream, forked after the work already queued and joined
    // before anything queued later (parallel branches in a captured graph).
    // (One side stream per persistent launch, so that those also overlap
    // each other, measured no faster: 4096 segments 0.614, 0.613, 0.619 ms
    // with one, two and three.  Forking after the device-wide count, so that
    // the count runs alone, was slower: 0.600 -> 0.639 ms.)
    const bool fork = !P.launches.empty() && !P.wide.empty() && lane.side;
    hipStream_t ls = lane.stream;
    if (fork) {
        if (hipEventRecord(lane.fork, lane.stream) != hipSuccess ||
            hipStreamWaitEvent(lane.side, lane.fork, 0) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        ls = lane.side;
    }
    for (size_t li = 0; li < P.launches.size(); ++li) {
        const SegLaunch& L = P.launches[li];
        mgpuSegArgs a;
        std::memset(&a, 0, sizeof(a));
        a.segs = P.devSegs;
        a.big_list = P.devLists + L.bigOffset;
        a.small_list = P.devLists + L.smallOffset;
        a.counters = P.devCounters + 4 * li;
        a.error_flag = lane.errorDev;
        a.num_big = L.numBig;
        a.num_small = L.numSmall;
        a.num_values = P.valueCount == -1 ? 1u : (uint32_t)P.valueCount;
        a.index_mode = P.valueCount == -1;
        a.result_side = (uint32_t)P.resultSide;
        a.num_passes = (uint32_t)L.passes;
        a.first_bit = (uint32_t)P.firstBit;
        a.num_bits = (uint32_t)P.numBits;
        a.rank1 = e->rankMethod == 1 ? 1u : 0u;
        sortStatus_t st = mgpu_launch(e, ls, L.fn, L.grid, L.block, &a);
        if (st) return st;
    }
    if (!P.wide.empty()) {
        sortStatus_t st = mgpu_enqueue_wide(P, lane);
        if (st) return st;
    }
    if (fork && (hipEventRecord(lane.join, lane.side) != hipSuccess ||
                 hipStreamWaitEvent(lane.stream, lane.join, 0) != hipSuccess))
        return SORT_STATUS_DEVICE_ERROR;
    return SORT_STATUS_SUCCESS;
}

void mgpu_arena_free(SegArena* a) {
    if (a->dev) (void)hipFree(a->dev);
    if (a->host) (void)hipHostFree(a->host);
    *a = SegArena();
}

void mgpu_seg_release(SegPlanData& P) {
    if (P.graphExec) (void)hipGraphExecDestroy(P.graphExec);
    if (P.graph) (void)hipGraphDestroy(P.graph);
    if (P.devBlock) (void)hipFree(P.devBlock);
    if (P.ownLookback) (void)hipFree(P.ownLookback);
    if (P.ownMulti) (void)hipFree(P.ownMulti);
    P.graphExec = nullptr;
    P.graph = nullptr;
    P.devBlock = nullptr;
    P.ownLookback = P.ownMulti = nullptr;
}

// ---------------------------------------------------------------------------
// Public entry points.

namespace {

sortStatus_t validate_segments(const sortData_d* d, int numSegments, const int* starts) {
    if (!d || numSegments < 0 || (numSegments > 0 && !starts)) return SORT_STATUS_INVALID_VALUE;
    if (d->firstBit < 0 || d->endBit > 32 || d->endBit < d->firstBit) return SORT_STATUS_INVALID_VALUE;
    if (d->valueCount < -1 || d->valueCount > 6) return SORT_STATUS_INVALID_VALUE;
    if (d->numElements < 0 || d->numElements > d->maxElements) return SORT_STATUS_INVALID_VALUE;
    if (numSegments > 0) {
        if (starts[0] < 0) return SORT_STATUS_INVALID_VALUE;
        for (int i = 0; i < numSegments; ++i)
            if (starts[i + 1] < starts[i]) return SORT_STATUS_INVALID_VALUE;
        if (starts[numSegments] > d->numElements) return SORT_STATUS_INVALID_VALUE;
    }
    if (!d->keys[0] || !d->keys[1]) return SORT_STATUS_INVALID_VALUE;
    // Both sides of every carried array, in index mode too: a segment's
    // passes may end on side 0, and a workgroup that runs more passes than
    // sortArray would ping-pongs through it.
    const int carried = d->valueCount == -1 ? 1 : d->valueCount;
    for (int v = 0; v < carried; ++v)
        if (!d->values[v][1] || !d->values[v][0]) return SORT_STATUS_INVALID_VALUE;
    return SORT_STATUS_SUCCESS;
}

std::vector<mgpuSegDesc> descs_from_data(const sortData_d* d, int numSegments, const int* starts) {
    std::vector<mgpuSegDesc> segs((size_t)numSegments);
    const int carried = d->valueCount == -1 ? 1 : d->valueCount;
    for (int i = 0; i < numSegments; ++i) {
        mgpuSegDesc& s = segs[i];
        std::memset(&s, 0, sizeof(s));
        const size_t off = (size_t)starts[i];
        for (int side = 0; side < 2; ++side) {
            s.keys[side] = reinterpret_cast<uint32_t*>(d->keys[side]) + off;
            for (int v = 0; v < carried; ++v)
                if (d->values[v][side])
                    s.vals[v][side] = reinterpret_cast<uint32_t*>(d->values[v][side]) + off;
        }
        s.n = (uint32_t)(starts[i + 1] - starts[i]);
    }
    return segs;
}

void finish_swap(sortData_d* d, int resultSide) {
    if (resultSide) {
        std::swap(d->keys[0], d->keys[1]);
        for (int v = 0; v < MGPU_MAX_VALUES; ++v) std::swap(d->values[v][0], d->values[v][1]);
        d->parity ^= 1;
    }
}

}  // namespace

// A reusable segmented sort.  The graph holds buffer addresses, so one graph
// is captured per orientation of the ping-pong (orientation 1 = sides
// swapped, which is where the data is after an odd-pass execute).
struct sortSegPlan_d {
    sortEngine_d* engine = nullptr;
    sortData_d layout;          // buffers as at creation (orientation 0)
    std::vector<int> starts;
    int numSegments = 0;
    SegPlanData data[2];
    bool built[2] = {false, false};
    hipStream_t captureStream = nullptr;
};

namespace {

sortStatus_t build_orientation(sortSegPlan_d* p, int o) {
    sortData_d d = p->layout;
    if (o) {
        std::swap(d.keys[0], d.keys[1]);
        for (int v = 0; v < MGPU_MAX_VALUES; ++v) std::swap(d.values[v][0], d.values[v][1]);
    }
    std::vector<mgpuSegDesc> segs = descs_from_data(&d, p->numSegments, p->starts.data());
    sortEngine_d* e = p->engine;
    SegPlanData& P = p->data[o];
    sortStatus_t st = mgpu_seg_build(e, e->lane, segs.data(), p->numSegments, d.valueCount,
                                     d.firstBit, d.endBit, &P, nullptr, /*ownScratch=*/true);
    if (st) return st;
    SortLane& lane = e->lane;
    if (lane.histDirty) {  // no memset nodes inside the graph
        if (hipMemsetAsync(lane.hist, 0, kLaneHistWords * sizeof(uint32_t), lane.stream) !=
                hipSuccess ||
            hipMemsetAsync(lane.tileCounters, 0, MGPU_LANE_COUNTERS * sizeof(uint32_t),
                           lane.stream) != hipSuccess ||
            hipStreamSynchronize(lane.stream) != hipSuccess)
            return SORT_STATUS_DEVICE_ERROR;
        lane.histDirty = false;
    }
    if (!p->captureStream &&
        hipStreamCreateWithFlags(&p->captureStream, hipStreamNonBlocking) != hipSuccess) {
        p->captureStream = nullptr;
        return SORT_STATUS_DEVICE_ERROR;
    }
    hipStream_t cs = p->captureStream;
    hipStream_t saved = lane.stream;
    const bool timing = e->timing;
    lane.stream = cs;
    e->timing = false;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        st = SORT_STATUS_DEVICE_ERROR;
    } else {
        sortStatus_t est = mgpu_seg_enqueue(P, lane);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(cs, &g);
        P.graph = g;
        if (est) st = est;
        else if (ce != hipSuccess || !g) st = SORT_STATUS_DEVICE_ERROR;
        else if (hipGraphInstantiate(&P.graphExec, g, nullptr, nullptr, 0) != hipSuccess)
            st = SORT_STATUS_DEVICE_ERROR;
    }
    lane.stream = saved;
    e->timing = timing;
    if (!st) p->built[o] = true;
    return st;
}

}  // namespace

extern "C" {

sortStatus_t SORTAPI sortSegmentedArray(sortEngine_t engine, sortData_t data, int numSegments,
                                        const int* segmentStarts) {
    if (!engine) return SORT_STATUS_INVALID_VALUE;
    sortStatus_t st = validate_segments(data, numSegments, segmentStarts);
    if (st) return st;
    if (data->endBit == data->firstBit || numSegments == 0) return SORT_STATUS_SUCCESS;
    if (mgpu_lane_take_error(engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    std::vector<mgpuSegDesc> segs = descs_from_data(data, numSegments, segmentStarts);
    SegPlanData P;
    st = mgpu_seg_build(engine, engine->lane, segs.data(), numSegments, data->valueCount,
                        data->firstBit, data->endBit, &P);
    if (!st) st = mgpu_seg_enqueue(P, engine->lane);
    if (!st && hipStreamSynchronize(engine->lane.stream) != hipSuccess)
        st = SORT_STATUS_DEVICE_ERROR;
    mgpu_seg_release(P);
    // The buffers were sorted (or not) either way: swap them as a finished
    // sort does, and report a failed self-check.
    if (!st) finish_swap(data, P.resultSide);
    if (!st) st = mgpu_lane_take_error(engine->lane);
    return st;
}

sortStatus_t SORTAPI sortSegmentedBatch(sortEngine_t engine, const sortSegment_d* segments,
                                        int numSegments, int valueCount, int firstBit, int endBit) {
    if (!engine || numSegments < 0 || (numSegments && !segments)) return SORT_STATUS_INVALID_VALUE;
    if (firstBit < 0 || endBit > 32 || endBit < firstBit || valueCount < -1 || valueCount > 6)
        return SORT_STATUS_INVALID_VALUE;
    if (endBit == firstBit || numSegments == 0) return SORT_STATUS_SUCCESS;
    std::vector<mgpuSegDesc> segs((size_t)numSegments);
    const int carried = valueCount == -1 ? 1 : valueCount;
    for (int i = 0; i < numSegments; ++i) {
        const sortSegment_d& in = segments[i];
        mgpuSegDesc& s = segs[i];
        std::memset(&s, 0, sizeof(s));
        if (in.numElements < 0 || in.numElements > SORT_MAX_ELEMENTS)
            return SORT_STATUS_INVALID_VALUE;
        if (in.numElements && (!in.keys[0] || !in.keys[1])) return SORT_STATUS_INVALID_VALUE;
        for (int side = 0; side < 2; ++side) {
            s.keys[side] = reinterpret_cast<uint32_t*>(in.keys[side]);
            for (int v = 0; v < carried; ++v) {
                if (in.numElements && !in.values[v][side]) return SORT_STATUS_INVALID_VALUE;
                s.vals[v][side] = reinterpret_cast<uint32_t*>(in.values[v][side]);
            }
        }
        s.n = (uint32_t)in.numElements;
    }
    if (mgpu_lane_take_error(engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    SegPlanData P;
    sortStatus_t st = mgpu_seg_build(engine, engine->lane, segs.data(), numSegments, valueCount,
                                     firstBit, endBit, &P);
    if (!st) st = mgpu_seg_enqueue(P, engine->lane);
    if (!st && hipStreamSynchronize(engine->lane.stream) != hipSuccess)
        st = SORT_STATUS_DEVICE_ERROR;
    mgpu_seg_release(P);
    if (!st) st = mgpu_lane_take_error(engine->lane);
    return st;
}

sortStatus_t SORTAPI sortSegmentedPlanCreate(sortEngine_t engine, const sortData_d* data,
                                             int numSegments, const int* segmentStarts,
                                             sortSegPlan_t* plan) {
    if (!engine || !plan) return SORT_STATUS_INVALID_VALUE;
    *plan = nullptr;
    sortStatus_t st = validate_segments(data, numSegments, segmentStarts);
    if (st) return st;
    sortSegPlan_d* p = new (std::nothrow) sortSegPlan_d();
    if (!p) return SORT_STATUS_HOST_ALLOC_FAILED;
    p->engine = engine;
    p->layout = *data;
    p->numSegments = numSegments;
    if (numSegments) p->starts.assign(segmentStarts, segmentStarts + numSegments + 1);
    if (data->endBit > data->firstBit && numSegments > 0) st = build_orientation(p, 0);
    if (st) {
        sortSegmentedPlanDestroy(p);
        return st;
    }
    *plan = p;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortSegmentedPlanExecute(sortSegPlan_t plan, sortData_t data) {
    if (!plan || !data) return SORT_STATUS_INVALID_VALUE;
    const sortData_d& L = plan->layout;
    if (L.endBit == L.firstBit || plan->numSegments == 0) return SORT_STATUS_SUCCESS;
    // As sortArray: an unreported failure of earlier work, before anything
    // is queued.
    if (mgpu_lane_take_error(plan->engine->lane)) return SORT_STATUS_KERNEL_ERROR;
    // The data must hold the plan's buffers, in either orientation.
    int o;
    if (data->keys[0] == L.keys[0] && data->keys[1] == L.keys[1]) o = 0;
    else if (data->keys[0] == L.keys[1] && data->keys[1] == L.keys[0]) o = 1;
    else return SORT_STATUS_INVALID_VALUE;
    if (!plan->built[o]) {
        sortStatus_t st = build_orientation(plan, o);
        if (st) return st;
    }
    SegPlanData& P = plan->data[o];
    sortEngine_d* e = plan->engine;
    if (e->timing) {
        // Per-kernel timing cannot look inside a graph: issue the launches
        // directly so that each one is event-timed.
        sortStatus_t st = mgpu_seg_enqueue(P, e->lane);
        if (st) return st;
    } else if (hipGraphLaunch(P.graphExec, e->lane.stream) != hipSuccess) {
        return SORT_STATUS_LAUNCH_ERROR;
    }
    finish_swap(data, P.resultSide);
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortSegmentedPlanDestroy(sortSegPlan_t plan) {
    if (!plan) return SORT_STATUS_SUCCESS;
    if (plan->engine) (void)hipStreamSynchronize(plan->engine->lane.stream);
    if (plan->captureStream) (void)hipStreamSynchronize(plan->captureStream);
    for (int o = 0; o < 2; ++o) mgpu_seg_release(plan->data[o]);
    if (plan->captureStream) (void)hipStreamDestroy(plan->captureStream);
    delete plan;
    return SORT_STATUS_SUCCESS;
}

}  // extern "C"
