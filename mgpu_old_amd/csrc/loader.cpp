// Code-object loading for libmgpusort: the role of LoadCountModule /
// LoadHistModule / LoadSortModule (sort/src/mgpusort/mgpusort.cpp:44-177),
// over hipModuleLoad / hipModuleGetFunction instead of cuModuleLoad.
// Names are kernelPath + "<object>.hsaco" (prefix concatenation, as the
// reference builds kernelPath + "count.cubin").
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.h"

namespace {

const char* const kRadixObject = "mgpusort_radix.hsaco";

sortStatus_t get_fn(sortEngine_d* e, hipModule_t m, const std::string& name, hipFunction_t* f) {
    if (hipModuleGetFunction(f, m, name.c_str()) != hipSuccess) return SORT_STATUS_KERNEL_ERROR;
    e->kernelNames[*f] = name;
    return SORT_STATUS_SUCCESS;
}

// Tile shapes: 1024 x 16 for large passes, 512 x 16 for small ones, unless
// MGPUSORT_TILE="<threads>x<items>" names one built shape for both (sweeps and
// tests of every shape).
sortStatus_t load_variant(sortEngine_d* e, uint32_t threads, uint32_t items, uint32_t rank,
                          PassVariant* v) {
    v->blockThreads = threads;
    v->itemsPerThread = items;
    sortStatus_t st = SORT_STATUS_SUCCESS;
    for (int m = 0; m < 4 && !st; ++m) {
        char name[96];
        std::snprintf(name, sizeof(name), "mgpu_rank_scatter_b%u_i%u_r%u_v%d", threads, items,
                      rank, m);
        st = get_fn(e, e->radixModule, name, &v->pass[m]);
        std::snprintf(name, sizeof(name), "mgpu_multi_scatter_b%u_i%u_r%u_v%d", threads, items,
                      rank, m);
        if (!st) st = get_fn(e, e->radixModule, name, &v->multi[m]);
        std::snprintf(name, sizeof(name), "mgpu_rank_scatter_b%u_i%u_r%u_x_v%d", threads, items,
                      rank, m);
        if (!st) st = get_fn(e, e->radixModule, name, &v->xseg[m]);
        std::snprintf(name, sizeof(name), "mgpu_multi_scatter_b%u_i%u_r%u_x_v%d", threads, items,
                      rank, m);
        if (!st) st = get_fn(e, e->radixModule, name, &v->multiX[m]);
    }
    if (st) return st;
    // Early-exit forms, built for the default shapes with atomic-add ranking.
    for (int m = 0; m < 4 && rank == 1; ++m) {
        char name[96];
        hipFunction_t f = nullptr;
        std::snprintf(name, sizeof(name), "mgpu_rank_scatter_b%u_i%u_r1_ee_v%d", threads, items, m);
        if (hipModuleGetFunction(&f, e->radixModule, name) == hipSuccess) {
            v->passEE[m] = f;
            e->kernelNames[f] = name;
        }
        std::snprintf(name, sizeof(name), "mgpu_rank_scatter_b%u_i%u_r1_x_ee_v%d", threads, items,
                      m);
        if (hipModuleGetFunction(&f, e->radixModule, name) == hipSuccess) {
            v->xsegEE[m] = f;
            e->kernelNames[f] = name;
        }
    }
    // Key-mapping forms of a typed sort, likewise (both, or neither).
    for (int m = 0; m < 4 && rank == 1; ++m) {
        char name[96], namex[96];
        hipFunction_t f = nullptr, fx = nullptr;
        std::snprintf(name, sizeof(name), "mgpu_rank_scatter_b%u_i%u_r1_km_v%d", threads, items, m);
        std::snprintf(namex, sizeof(namex), "mgpu_rank_scatter_b%u_i%u_r1_x_km_v%d", threads,
                      items, m);
        if (hipModuleGetFunction(&f, e->radixModule, name) != hipSuccess ||
            hipModuleGetFunction(&fx, e->radixModule, namex) != hipSuccess) {
            for (int k = 0; k < 4; ++k) v->passKM[k] = v->xsegKM[k] = nullptr;
            break;
        }
        v->passKM[m] = f;
        v->xsegKM[m] = fx;
        e->kernelNames[f] = name;
        e->kernelNames[fx] = namex;
    }
    // Packed multi-array forms, likewise.
    bool packed = rank == 1;
    for (int m = 0; m < 4 && packed; ++m) {
        char name[96];
        hipFunction_t f = nullptr, fx = nullptr;
        std::snprintf(name, sizeof(name), "mgpu_multi_packed_b%u_i%u_v%d", threads, items, m);
        char namex[96];
        std::snprintf(namex, sizeof(namex), "mgpu_multi_packed_b%u_i%u_x_v%d", threads, items, m);
        if (hipModuleGetFunction(&f, e->radixModule, name) != hipSuccess ||
            hipModuleGetFunction(&fx, e->radixModule, namex) != hipSuccess) {
            packed = false;
            break;
        }
        v->packed[m] = f;
        v->packedX[m] = fx;
        e->kernelNames[f] = name;
        e->kernelNames[fx] = namex;
    }
    if (!packed)
        for (int m = 0; m < 4; ++m) v->packed[m] = v->packedX[m] = nullptr;
    (void)hipGetLastError();
    // Segment workgroups loop over tiles: launch as many as fit at once.
    int perCU = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, v->xseg[MGPU_VALUES_SINGLE],
                                                           (int)threads, 0) != hipSuccess ||
        perCU < 1)
        perCU = 1;
    v->xsegGrid = (uint32_t)perCU * (uint32_t)e->numCUs;
    perCU = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, v->multiX[MGPU_VALUES_SINGLE],
                                                           (int)threads, 0) != hipSuccess ||
        perCU < 1)
        perCU = 1;
    v->multiXGrid = (uint32_t)perCU * (uint32_t)e->numCUs;
    if (v->packedX[MGPU_VALUES_SINGLE]) {
        perCU = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
                &perCU, v->packedX[MGPU_VALUES_SINGLE], (int)threads, 0) != hipSuccess ||
            perCU < 1)
            perCU = 1;
        v->packedXGrid = (uint32_t)perCU * (uint32_t)e->numCUs;
    }
    return SORT_STATUS_SUCCESS;
}

// Segments per XCD need: the XCD a workgroup runs on (XCC_ID), a grid's
// workgroups spread over every XCD, and lookback words that one CU writes with
// nontemporal stores seen by another CU of the same XCD polling with
// nontemporal loads.  mgpu_xcd_probe checks all three; returns the number of
// XCDs to use (1: one device-wide chain).  MGPUSORT_XCD=0 turns them off,
// MGPUSORT_XCD=1 uses them at every size (tests of small sorts).
uint32_t xcd_segments(sortEngine_d* e) {
    const char* env = std::getenv("MGPUSORT_XCD");
    if (env && env[0] == '0' && env[1] == 0) return 1;
    e->xcdForce = env && env[0] == '1' && env[1] == 0;
    hipFunction_t f = nullptr;
    if (get_fn(e, e->radixModule, "mgpu_xcd_probe", &f)) return 1;
    constexpr uint32_t kGrid = 8 * MGPU_MAX_SEGS, kRounds = 64;
    const size_t outWords = kGrid + 2 * MGPU_MAX_SEGS, areaWords = MGPU_MAX_SEGS * 128;
    uint32_t* d = nullptr;
    if (hipMalloc(&d, (outWords + areaWords) * sizeof(uint32_t)) != hipSuccess) return 1;
    std::vector<uint32_t> h(outWords);
    uint32_t* out = d;
    uint32_t* area = d + outWords;
    uint32_t rounds = kRounds;
    void* params[] = {&out, &area, &rounds};
    bool ran = hipMemsetAsync(d, 0, (outWords + areaWords) * sizeof(uint32_t), e->lane.stream) ==
               hipSuccess;
    ran = ran && hipModuleLaunchKernel(f, kGrid, 1, 1, 64, 1, 1, 0, e->lane.stream, params,
                                       nullptr) == hipSuccess;
    ran = ran && hipMemcpyAsync(h.data(), d, outWords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                e->lane.stream) == hipSuccess;
    ran = ran && hipStreamSynchronize(e->lane.stream) == hipSuccess;
    (void)hipFree(d);
    if (!ran) return 1;
    uint32_t xcds = 0;
    uint32_t per[MGPU_MAX_SEGS] = {};
    for (uint32_t b = 0; b < kGrid; ++b) {
        if (h[b] >= MGPU_MAX_SEGS) return 1;
        per[h[b]]++;
        xcds = std::max(xcds, h[b] + 1);
    }
    if (xcds < 2 || xcds > MGPU_XSEG_COUNT_SEGS) return 1;
    for (uint32_t x = 0; x < xcds; ++x) {
        if (per[x] < 2) return 1;  // an XCD without workgroups, or without a partner
        if (h[kGrid + 2 * x] != kRounds || h[kGrid + 2 * x + 1] != 0) return 1;
    }
    return xcds;
}

// Whether the device serves the lanes of one LDS atomic add that hit the same
// word in ascending lane order, which the atomic-add ranking (RANK 1 kernels)
// needs for a stable sort: mgpu_rank_probe against the counts in lane order.
bool lds_add_in_lane_order(sortEngine_d* e) {
    hipFunction_t f = nullptr;
    if (get_fn(e, e->radixModule, "mgpu_rank_probe", &f)) return false;
    constexpr uint32_t kRounds = 16, kThreads = 1024;
    uint32_t* d = nullptr;
    if (hipMalloc(&d, kRounds * kThreads * sizeof(uint32_t)) != hipSuccess) return false;
    std::vector<uint32_t> h(kRounds * kThreads);
    const bool ran = mgpu_launch(e, e->lane.stream, f, 1, kThreads, &d) == SORT_STATUS_SUCCESS &&
                     hipMemcpyAsync(h.data(), d, h.size() * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, e->lane.stream) == hipSuccess &&
                     hipStreamSynchronize(e->lane.stream) == hipSuccess;
    (void)hipFree(d);
    if (!ran) return false;
    std::vector<uint32_t> count(16 * MGPU_RADIX, 0);
    for (uint32_t r = 0; r < kRounds; ++r)
        for (uint32_t t = 0; t < kThreads; ++t) {
            const uint32_t w = t / 64, digit = mgpu_rank_probe_digit(w, r, t % 64);
            if (h[r * kThreads + t] != count[w * MGPU_RADIX + digit]++) return false;
        }
    return true;
}

}  // namespace

sortStatus_t mgpu_engine_load(sortEngine_d* e) {
    std::string path = e->kernelPath + kRadixObject;
    if (hipModuleLoad(&e->radixModule, path.c_str()) != hipSuccess) {
        e->radixModule = nullptr;
        return SORT_STATUS_KERNEL_NOT_FOUND;
    }
    sortStatus_t st = get_fn(e, e->radixModule, "mgpu_digit_count", &e->fnCount);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_digit_count_cols", &e->fnCountCols);
    // Counts of typed keys: without them (an older code object) typed sorts
    // map the keys in place around the sort.
    if (!st && (get_fn(e, e->radixModule, "mgpu_digit_count_km", &e->fnCountKM) ||
                get_fn(e, e->radixModule, "mgpu_digit_count_cols_km", &e->fnCountColsKM))) {
        e->fnCountKM = e->fnCountColsKM = nullptr;
        (void)hipGetLastError();
    }
    if (!st) st = get_fn(e, e->radixModule, "mgpu_digit_scan", &e->fnScan);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_iota", &e->fnIota);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_key_transform", &e->fnKeyTransform);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_multi_count", &e->fnMultiCount);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_multi_count_cols", &e->fnMultiCountCols);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_multi_scan", &e->fnMultiScan);
    uint32_t smallShape[2] = {512, 16}, largeShape[2] = {1024, 16};
    const char* t = std::getenv("MGPUSORT_TILE");
    unsigned b = 0, i = 0;
    if (t && std::sscanf(t, "%ux%u", &b, &i) == 2 && b && i) {
        smallShape[0] = largeShape[0] = b;
        smallShape[1] = largeShape[1] = i;
    }
    if (st) return st;
    st = mgpu_seg_load(e);
    if (st) return st;

    st = mgpu_lane_init(&e->lane);
    if (st) return st;
    // Ranking method: the single LDS atomic add per key where the device
    // orders same-word lanes (checked here), else the match masks.
    // MGPUSORT_RANK=0|1 forces one (tests of both).
    const char* r = std::getenv("MGPUSORT_RANK");
    if (r && (*r == '0' || *r == '1') && r[1] == 0) e->rankMethod = (uint32_t)(*r - '0');
    else e->rankMethod = lds_add_in_lane_order(e) ? 1u : 0u;
    e->xcdSegs = xcd_segments(e);
    mgpu_load_sort_table(e);
    st = load_variant(e, smallShape[0], smallShape[1], e->rankMethod, &e->small);
    if (!st) st = load_variant(e, largeShape[0], largeShape[1], e->rankMethod, &e->large);
    if (st) return st;
    const char* spin = std::getenv("MGPUSORT_SPIN_LIMIT");
    if (spin && *spin) e->spinLimit = (uint32_t)std::strtoul(spin, nullptr, 10);
    const char* ee = std::getenv("MGPUSORT_EE");
    if (ee && (ee[0] == '0' || ee[0] == '1') && ee[1] == 0) e->eePolicy = ee[0] - '0';
    return SORT_STATUS_SUCCESS;
}

void mgpu_engine_destroy(sortEngine_d* e) {
    if (!e) return;
    if (e->lane.stream || e->radixModule) (void)hipStreamSynchronize(e->lane.stream);
    for (auto& t : e->timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (hipEvent_t ev : e->eventPool) (void)hipEventDestroy(ev);
    mgpu_lane_free(&e->lane);
    if (e->radixModule) (void)hipModuleUnload(e->radixModule);
    mgpu_seg_unload(e);
    delete e;
}
