// Code-object loading for libmgpusort: the role of LoadCountModule /
// LoadHistModule / LoadSortModule (sort/src/mgpusort/mgpusort.cpp:44-177),
// over hipModuleLoad / hipModuleGetFunction instead of cuModuleLoad.
// Names are kernelPath + "<object>.hsaco" (prefix concatenation, as the
// reference builds kernelPath + "count.cubin").
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
    }
    return st;
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
    if (!st) st = get_fn(e, e->radixModule, "mgpu_digit_scan", &e->fnScan);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_iota", &e->fnIota);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_multi_count", &e->fnMultiCount);
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
    st = load_variant(e, smallShape[0], smallShape[1], e->rankMethod, &e->small);
    if (!st) st = load_variant(e, largeShape[0], largeShape[1], e->rankMethod, &e->large);
    if (st) return st;
    if (hipMalloc(&e->errorFlag, 16) != hipSuccess) {
        e->errorFlag = nullptr;
        return SORT_STATUS_DEVICE_ALLOC_FAILED;
    }
    if (hipMemsetAsync(e->errorFlag, 0, 16, e->lane.stream) != hipSuccess ||
        hipStreamSynchronize(e->lane.stream) != hipSuccess)
        return SORT_STATUS_DEVICE_ERROR;
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
    if (e->errorFlag) (void)hipFree(e->errorFlag);
    if (e->radixModule) (void)hipModuleUnload(e->radixModule);
    mgpu_seg_unload(e);
    delete e;
}
