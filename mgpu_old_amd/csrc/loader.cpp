// Code-object loading for libmgpusort: the role of LoadCountModule /
// LoadHistModule / LoadSortModule (sort/src/mgpusort/mgpusort.cpp:44-177),
// over hipModuleLoad / hipModuleGetFunction instead of cuModuleLoad.
// Names are kernelPath + "<object>.hsaco" (prefix concatenation, as the
// reference builds kernelPath + "count.cubin").
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "engine_internal.h"

namespace {

const char* const kRadixObject = "mgpusort_radix.hsaco";

sortStatus_t get_fn(sortEngine_d* e, hipModule_t m, const std::string& name, hipFunction_t* f) {
    if (hipModuleGetFunction(f, m, name.c_str()) != hipSuccess) return SORT_STATUS_KERNEL_ERROR;
    e->kernelNames[*f] = name;
    return SORT_STATUS_SUCCESS;
}

// Tile shape and ranking method: MGPUSORT_TILE="<threads>x<items>" and
// MGPUSORT_RANK=<0|1> pick one of the built variants (for sweeps); otherwise
// the default.
SortConfig pick_config() {
    SortConfig c = {512, 16, 0, 0};
    const char* t = std::getenv("MGPUSORT_TILE");
    if (t && *t) {
        unsigned b = 0, i = 0;
        if (std::sscanf(t, "%ux%u", &b, &i) == 2 && b && i) {
            c.blockThreads = b;
            c.itemsPerThread = i;
        }
    }
    const char* r = std::getenv("MGPUSORT_RANK");
    if (r && (*r == '0' || *r == '1')) c.rankMethod = (uint32_t)(*r - '0');
    const char* ps = std::getenv("MGPUSORT_PERSIST");
    if (ps && (*ps == '0' || *ps == '1')) c.persistent = (uint32_t)(*ps - '0');
    return c;
}

// Workgroups of f that fit on the device at once (grid of a persistent launch).
uint32_t resident_grid(const sortEngine_d* e, hipFunction_t f) {
    int blocks = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, (int)e->config.blockThreads,
                                                           0) != hipSuccess ||
        blocks < 1)
        blocks = 1;
    return (uint32_t)blocks * (uint32_t)e->numCUs;
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
    e->config = pick_config();
    if (!st) st = get_fn(e, e->radixModule, "mgpu_multi_count", &e->fnMultiCount);
    if (!st) st = get_fn(e, e->radixModule, "mgpu_multi_scan", &e->fnMultiScan);
    const char* pers = e->config.persistent ? "_p" : "";
    for (int v = 0; v < 4 && !st; ++v) {
        char name[96];
        std::snprintf(name, sizeof(name), "mgpu_multi_scatter_b%u_i%u_r%u%s_v%d",
                      e->config.blockThreads, e->config.itemsPerThread, e->config.rankMethod,
                      pers, v);
        st = get_fn(e, e->radixModule, name, &e->fnMultiPass[v]);
        if (!st) e->multiPassGrid[v] = resident_grid(e, e->fnMultiPass[v]);
    }
    for (int v = 0; v < 4 && !st; ++v) {
        char name[96];
        std::snprintf(name, sizeof(name), "mgpu_rank_scatter_b%u_i%u_r%u%s_v%d",
                      e->config.blockThreads, e->config.itemsPerThread, e->config.rankMethod,
                      pers, v);
        st = get_fn(e, e->radixModule, name, &e->fnPass[v]);
        if (!st) e->passGrid[v] = resident_grid(e, e->fnPass[v]);
    }
    if (st) return st;
    st = mgpu_seg_load(e);
    if (st) return st;

    st = mgpu_lane_init(&e->lane);
    if (st) return st;
    if (hipMalloc(&e->errorFlag, 16) != hipSuccess) {
        e->errorFlag = nullptr;
        return SORT_STATUS_DEVICE_ALLOC_FAILED;
    }
    if (hipMemset(e->errorFlag, 0, 16) != hipSuccess) return SORT_STATUS_DEVICE_ERROR;
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
