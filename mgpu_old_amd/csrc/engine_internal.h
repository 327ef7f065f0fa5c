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

// Tile shape of the rank_scatter kernel: blockThreads * itemsPerThread keys.
struct SortConfig {
    uint32_t blockThreads;
    uint32_t itemsPerThread;
};

struct sortEngine_d {
    int refcount = 1;  // sortIncEngine / sortReleaseEngine (not atomic, as util/util.h:84-98)
    int device = -1;
    int numCUs = 0;
    std::string kernelPath;
    hipStream_t stream = nullptr;

    hipModule_t radixModule = nullptr;
    hipFunction_t fnCount = nullptr;
    hipFunction_t fnScan = nullptr;
    hipFunction_t fnPass[4] = {};  // by value mode (MGPU_VALUES_*)
    SortConfig config = {256, 16};

    // Fixed scratch: hist[4][256], digitStart[4][256], tileCounters[4],
    // errorFlag -- one allocation.
    uint32_t* fixedScratch = nullptr;
    uint32_t* hist = nullptr;
    uint32_t* digitStart = nullptr;
    uint32_t* tileCounters = nullptr;
    uint32_t* errorFlag = nullptr;
    bool histDirty = false;  // hist/tileCounters may hold a partial sort's counts

    // Lookback words (grow-only).
    uint32_t* lookback = nullptr;
    uint64_t lookbackWords = 0;

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

// Loads the code objects named from e->kernelPath and allocates the fixed
// scratch.  On failure the engine is left for mgpu_engine_destroy.
sortStatus_t mgpu_engine_load(sortEngine_d* e);
void mgpu_engine_destroy(sortEngine_d* e);

// Pass split of numBits into passes of at most maxWidth bits; returns the
// pass count and fills widths (narrower passes first).
int mgpu_plan_passes(int numBits, int maxWidth, int* widths);

sortStatus_t mgpu_launch(sortEngine_d* e, hipFunction_t f, unsigned grid, unsigned block,
                         void* args, size_t argBytes);

// Runs numPasses digit passes over n keys starting at firstBit.  Pass p reads
// buffer side (p & 1) and writes side (p & 1) ^ 1 of keyBuf / valBuf, so the
// result is on side (numPasses & 1).  valueCount as in sortData_d.
sortStatus_t mgpu_sort_passes(sortEngine_d* e, uint32_t n, int firstBit, int numPasses,
                              const int* widths, int valueCount, uint32_t* const keyBuf[2],
                              uint32_t* const valBuf[MGPU_MAX_VALUES][2]);
