/*
 * mgpusort.h -- C API of the MI355X (gfx950) LSD radix sort.
 *
 * Drop-in for the reference's inc/mgpusort.h (seanbaxter/mgpu_old): the same
 * names, status codes, argument meanings and sortData_d layout, so existing
 * C/C++ host code compiles and links against libmgpusort.so unchanged.  The
 * engine loads gfx950 code objects (*.hsaco) from kernelPath through the HIP
 * module API; no CUDA header is needed (CUdeviceptr is provided below as a
 * 64-bit integer of the same size and alignment).
 *
 * Every function returns a status code; nothing throws across this boundary.
 * An engine is bound to the HIP device that is current when it is created and
 * is not reentrant: one host thread at a time per engine.
 */
#ifndef MGPUSORT_H
#define MGPUSORT_H

#define SORTAPI

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Everything declared here is exported from libmgpusort.so. */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* Device pointer as an integer, the role CUdeviceptr (64-bit) plays in the
 * reference header (inc/mgpusort.h:15,115-126). */
typedef unsigned long long sortDevicePtr_t;
#ifndef MGPUSORT_NO_CUDEVICEPTR
typedef sortDevicePtr_t CUdeviceptr;
#endif

/* Status codes: inc/mgpusort.h:21-36, same order and values. */
typedef enum {
    SORT_STATUS_SUCCESS = 0,
    SORT_STATUS_NOT_INITIALIZED,
    SORT_STATUS_DEVICE_ALLOC_FAILED,
    SORT_STATUS_HOST_ALLOC_FAILED,
    SORT_STATUS_CONFIG_NOT_SUPPORTED,
    SORT_STATUS_CONTEXT_MISMATCH,
    SORT_STATUS_INVALID_CONTEXT,
    SORT_STATUS_KERNEL_NOT_FOUND,
    SORT_STATUS_KERNEL_ERROR,
    SORT_STATUS_LAUNCH_ERROR,
    SORT_STATUS_INVALID_VALUE,
    SORT_STATUS_DEVICE_ERROR,
    SORT_STATUS_INTERNAL_ERROR,
    SORT_STATUS_UNSUPPORTED_DEVICE
} sortStatus_t;

/* valueCount that asks the first executed pass to write source indices into
 * values1 (the reference spells it -1: inc/mgpusort.h:70-73). */
#define SORT_VALUE_INDEX (-1)

/* Largest numElements one sortArray call accepts (offsets are kept in 30 bits
 * of a lookback word).  Larger arrays return SORT_STATUS_CONFIG_NOT_SUPPORTED. */
#define SORT_MAX_ELEMENTS ((1 << 30) - 1)

/* Replaces inc/mgpusort.h:39.  Returns 0 for an unknown code. */
const char* SORTAPI sortStatusString(sortStatus_t status);

struct sortEngine_d;
typedef struct sortEngine_d* sortEngine_t;

/* Replaces inc/mgpusort.h:50-51.  Binds to the current HIP device, which must
 * be gfx950 (else SORT_STATUS_UNSUPPORTED_DEVICE), and loads the code objects
 * kernelPath + "<name>.hsaco" (prefix concatenation, as the reference does
 * with cubins).  A NULL or empty kernelPath means the directory holding
 * libmgpusort.so; the environment variable MGPUSORT_KERNEL_PATH, when set,
 * overrides both. */
sortStatus_t SORTAPI sortCreateEngine(const char* kernelPath, sortEngine_t* engine);

/* Replaces inc/mgpusort.h:54-55.  Validates the arguments as the reference
 * does (128 or 256 threads, 4 or 8 values per thread); every kernel is
 * already resident after sortCreateEngine, so nothing else happens. */
sortStatus_t SORTAPI sortLoadKernel(sortEngine_t engine, int numSortThreads,
                                    int valuesPerThread, bool useTransList, int valueCount);

/* Replace inc/mgpusort.h:58-59: engine reference count. */
sortStatus_t SORTAPI sortIncEngine(sortEngine_t engine);
sortStatus_t SORTAPI sortReleaseEngine(sortEngine_t engine);

/* Alias of sortReleaseEngine (named in a comment at
 * sort/src/mgpusort/mgpusort.cpp:198 but never exported there). */
sortStatus_t SORTAPI sortDestroyEngine(sortEngine_t engine);

/* Replaces inc/mgpusort.h:63-147, same field order, LP64 layout
 * (sizeof == 144, pointers from offset 32). */
typedef struct {
    /* Capacity of every buffer, in elements. */
    int maxElements;
    /* Number of elements to sort. */
    int numElements;
    /* Number of value arrays carried (0..6), or SORT_VALUE_INDEX (-1) to write
     * source indices into values1 on the first pass and carry them after. */
    int valueCount;
    /* Sort by the bit field [firstBit, endBit) of each key. */
    int firstBit;
    int endBit;
    /* Accepted for compatibility.  This engine never reads or writes any
     * element at or beyond numElements, so the keys past the end are always
     * preserved. */
    bool preserveEndKeys;
    /* Look for already-sorted input and skip the passes that would not change
     * the order; the output is the same either way. */
    bool earlyExit;
    /* Toggled once per executed pass (buffer swap).  An array at [0] before
     * the sort is at [parity] after it; [0] always holds the sorted result. */
    int parity;
    union {
        struct {
            CUdeviceptr keys[2];
            union {
                struct {
                    CUdeviceptr values1[2];
                    CUdeviceptr values2[2];
                    CUdeviceptr values3[2];
                    CUdeviceptr values4[2];
                    CUdeviceptr values5[2];
                    CUdeviceptr values6[2];
                };
                CUdeviceptr values[6][2];
            };
        };
        struct {
            unsigned int* pKeys[2];
            union {
                struct {
                    unsigned int* pValues1[2];
                    unsigned int* pValues2[2];
                    unsigned int* pValues3[2];
                    unsigned int* pValues4[2];
                    unsigned int* pValues5[2];
                    unsigned int* pValues6[2];
                };
                unsigned int* pValues[6][2];
            };
        };
    };
} sortData_d;
typedef sortData_d* sortData_t;

/* Replace inc/mgpusort.h:155-157.  sortCreateData allocates the descriptor
 * and all its buffers (maxElements rounded up to 2048, numElements =
 * maxElements, bits [0, 32)); sortDestroyData frees both. */
sortStatus_t SORTAPI sortCreateData(sortEngine_t engine, int maxElements, int valueCount,
                                    sortData_t* data);
sortStatus_t SORTAPI sortDestroyData(sortData_t data);

/* Replace inc/mgpusort.h:164-165.  sortAllocData allocates every null buffer
 * among keys[2] and the first (valueCount == -1 ? 1 : valueCount) value
 * pairs, RoundUp(maxElements, 2048) elements each, and sets parity = 0.
 * sortFreeData frees every non-null buffer (attached ones included) and
 * zeroes the descriptor. */
sortStatus_t SORTAPI sortAllocData(sortEngine_t engine, sortData_t data);
sortStatus_t SORTAPI sortFreeData(sortData_d* data);

/* Replaces inc/mgpusort.h:167.  Stable sort of keys[0][0, numElements) by the
 * bit field [firstBit, endBit), values following the same permutation.  The
 * work is queued on the engine's stream (the null stream unless
 * sortSetStream chose another); the call returns once it is queued.
 *
 * Kernels check themselves (a lookback wait that gives up, a scatter offset
 * out of range, a tile or segment left unfinished) and record a failure in an
 * error word of the engine.  A failure of earlier, asynchronous work that has
 * not been reported yet makes sortArray, sortArrayEx and
 * sortSegmentedPlanExecute return SORT_STATUS_KERNEL_ERROR before they queue
 * anything (data unchanged; the word is cleared, so a second call runs).
 * sortHost, sortDevice, sortSegmentedArray, sortSegmentedBatch and
 * sortCheckDevice wait for the stream and return it for their own work too;
 * the sort server reports it per batch in sortCompletion_d.status. */
sortStatus_t SORTAPI sortArray(sortEngine_t engine, sortData_t data);

/* No counterpart in the reference.  sortArray for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments, kernels, result, parity
 * and error reporting as sortArray, but no hipMalloc, hipFree, memset or
 * host synchronization: the scratch is a fixed array in the library's device
 * code (about 8.2 MB per GPU, held until the process ends).  At most
 * SORT_NOALLOC_MAX_ELEMENTS elements (at the engine's default tile shapes;
 * more returns SORT_STATUS_CONFIG_NOT_SUPPORTED); earlyExit is ignored (it
 * needs a blocking readback).  Every call on one GPU shares that array, so
 * calls that overlap on the same GPU race: one at a time per device. */
#define SORT_NOALLOC_MAX_ELEMENTS (1 << 24)
sortStatus_t SORTAPI sortArrayNoAlloc(sortEngine_t engine, sortData_t data);

/* Replaces inc/mgpusort.h:169-170.  Same as sortArray with the pass split
 * chosen by the caller: ceil(bits / bitPass) passes of bitPass bits (the first
 * ones a bit narrower when it does not divide), bitPass outside 1..6 taken as
 * 6.  numSortThreads must be 128 or 256 and valuesPerThread 8, as in the
 * reference; the engine picks its own tile shape. */
sortStatus_t SORTAPI sortArrayEx(sortEngine_t engine, sortData_t data, int numSortThreads,
                                 int valuesPerThread, int bitPass, bool useTransList);

/* Replace inc/mgpusort.h:183-187.  Convenience wrappers (keys plus at most one
 * value array, bits [0, numBits)); both block until the result is in place. */
sortStatus_t SORTAPI sortHost(sortEngine_t engine, unsigned int* keys, unsigned int* values,
                              int numElements, int numBits);
sortStatus_t SORTAPI sortDevice(sortEngine_t engine, CUdeviceptr keys, CUdeviceptr values,
                                int numElements, int numBits);

/* ---- Additions with no counterpart in the reference. ---------------------- */

/* Stream the engine queues its work on (a hipStream_t; NULL = null stream).
 * Work queued after the switch waits for what was queued on the old stream
 * (an event), since both use the engine's scratch. */
sortStatus_t SORTAPI sortSetStream(sortEngine_t engine, void* stream);
sortStatus_t SORTAPI sortGetStream(sortEngine_t engine, void** stream);

/* Number of digit passes sortArray runs for a bit range of numBits bits, and
 * their widths (widths may be NULL).  Host-only; needs no device. */
int SORTAPI sortPlanPasses(int numBits, int* widths);

/* ---- Segmented sort (no counterpart in the reference). ---------------------
 * Many independent arrays sorted in one persistent launch.  Each segment ends
 * up exactly as sortArray would leave it if called on that segment alone with
 * the same bit range and value count: stable by the bit field, values
 * following, the result on side (sortPlanPasses(endBit - firstBit) & 1) of its
 * buffers.  Segments below the tuning table's device-wide threshold (4097
 * elements in the shipped kernelPath + "mgpusort_seg_tuning.txt") are sorted
 * by one wave or one workgroup each (digit width, keys per thread and
 * workgroup size per size bucket from the same table); larger ones run the
 * device-wide passes, all of them in one launch per pass.  With
 * SORT_VALUE_INDEX both sides of values[0] must be given. */

/* Sorts the segments [segmentStarts[i], segmentStarts[i+1]) of data's
 * buffers (host array of numSegments + 1 nondecreasing offsets, the last at
 * most numElements) by data's bit range and value count, then swaps buffers
 * and toggles parity as sortArray does.  Blocks until done. */
sortStatus_t SORTAPI sortSegmentedArray(sortEngine_t engine, sortData_t data, int numSegments,
                                        const int* segmentStarts);

/* One independent array of a batch with buffers of its own. */
typedef struct {
    CUdeviceptr keys[2];
    CUdeviceptr values[6][2];
    int numElements;
} sortSegment_d;

/* Sorts every segment of the batch (same bit range and value count for
 * all); each result is left on side (sortPlanPasses(endBit - firstBit) & 1)
 * of that segment's buffers.  Blocks until done. */
sortStatus_t SORTAPI sortSegmentedBatch(sortEngine_t engine, const sortSegment_d* segments,
                                        int numSegments, int valueCount, int firstBit,
                                        int endBit);

/* Reusable form of sortSegmentedArray: the work lists are built once and the
 * launches captured in a hipGraph; each execute replays it on the engine's
 * stream (asynchronously) and swaps data's buffers as sortArray does.  data
 * must hold the buffers given at creation, in either order.  The plan owns
 * the scratch of its device-wide passes (allocated at creation, freed by
 * destroy), so later sorts on the engine, which grow and replace the
 * engine's own scratch, cannot disturb a replay. */
struct sortSegPlan_d;
typedef struct sortSegPlan_d* sortSegPlan_t;
sortStatus_t SORTAPI sortSegmentedPlanCreate(sortEngine_t engine, const sortData_d* data,
                                             int numSegments, const int* segmentStarts,
                                             sortSegPlan_t* plan);
sortStatus_t SORTAPI sortSegmentedPlanExecute(sortSegPlan_t plan, sortData_t data);
sortStatus_t SORTAPI sortSegmentedPlanDestroy(sortSegPlan_t plan);

/* ---- Sort service (no counterpart in the reference). -----------------------
 * Continuous batching of sort requests: requests queue in arrival order, are
 * admitted while the ping-pong scratch they need (4 * numElements * (1 +
 * value arrays) bytes) fits in a fixed HBM budget, and every admitted request
 * is packed into the segmented kernels of the next free in-flight slot, each
 * slot with its own stream, so batches overlap on the device.  Each request is
 * sorted in place, exactly as sortArray would sort it alone. */
typedef struct {
    CUdeviceptr keys;       /* numElements keys, sorted in place */
    CUdeviceptr values[6];  /* the first (valueCount == -1 ? 1 : valueCount) are used */
    int numElements;
    int valueCount;         /* -1 (indices into values[0]) .. 6 */
    int firstBit;
    int endBit;
    unsigned long long tag; /* returned with the completion */
} sortRequest_d;

typedef struct {
    unsigned long long tag;
    sortStatus_t status;
    int batch;       /* sequence number of the batch that ran it (-1: none) */
    double submitMs; /* host clock, milliseconds since the server started */
    double launchMs;
    double doneMs;
} sortCompletion_d;

struct sortServer_d;
typedef struct sortServer_d* sortServer_t;

/* hbmBudgetBytes: scratch slab for admitted requests.  maxInFlight: batches
 * on the device at once (1..16).  maxElements: largest request accepted.
 * Besides the slab, each in-flight slot holds the scratch of its largest
 * batch (1024 requests, 2^26 keys or one request of maxElements): about
 * 65 MB of device memory and 8 MB of pinned host memory at the default
 * tile shapes, allocated here so that no batch allocates. */
sortStatus_t SORTAPI sortServerCreate(sortEngine_t engine, unsigned long long hbmBudgetBytes,
                                      int maxInFlight, int maxElements, sortServer_t* server);
/* Thread-safe.  The buffers must stay untouched until the completion. */
sortStatus_t SORTAPI sortServerSubmit(sortServer_t server, const sortRequest_d* request);
/* Up to maxCount completions, waiting at most timeoutMs for the first;
 * returns how many were written (-1 on a bad argument). */
int SORTAPI sortServerPoll(sortServer_t server, sortCompletion_d* out, int maxCount,
                           int timeoutMs);
/* Load driver: submits requests[i] (tag = i) at arrivalMs[i] after the call
 * starts, waits for all of them and writes completions[i]. */
sortStatus_t SORTAPI sortServerReplay(sortServer_t server, const sortRequest_d* requests,
                                      const double* arrivalMs, int count,
                                      sortCompletion_d* completions);
/* Finishes everything submitted, then frees the server. */
sortStatus_t SORTAPI sortServerDestroy(sortServer_t server);

/* Waits for the engine's stream and reports SORT_STATUS_KERNEL_ERROR if a
 * kernel detected an inconsistency (a scatter offset out of range, a
 * lookback wait that gave up, or a segment left without a workgroup on its
 * XCD) that no call has reported yet; SUCCESS otherwise. */
sortStatus_t SORTAPI sortCheckDevice(sortEngine_t engine);

/* What the engine chose at creation, and the last device inconsistency code
 * (0 none; 1 a lookback wait gave up; 2 a tile counter left dirty; 4 a
 * scatter offset out of range; 5 a segment with no workgroup on its XCD;
 * 6 a packed tile with more segments than it can hold; 7 a segment larger
 * than the tile of a kernel built without the global-memory routine: both
 * only from a broken plan) without clearing it.  No counterpart in the
 * reference. */
typedef struct {
    int device;
    int numCUs;
    int xcdSegments;   /* lookback chains per device-wide pass: 1, or one per XCD */
    int rankMethod;    /* 1: one LDS atomic add per key; 0: LDS match masks */
    int smallTileKeys; /* tile of passes below the switch */
    int largeTileKeys; /* tile of passes from the switch on */
    unsigned int deviceError;
    int reserved;
} sortEngineInfo_t;
sortStatus_t SORTAPI sortEngineInfo(sortEngine_t engine, sortEngineInfo_t* info);

/* Signed and floating-point keys.  sortTransformKeys maps numElements int32
 * or float32 keys in place, on the engine's stream, to unsigned keys whose
 * order is the numeric order (toRadixOrder true), or back (false): sort int
 * or float keys by mapping them, sortArray, and mapping the result back.
 * int32: the sign bit flipped.  float32: non-negative values get the sign bit
 * set, negative ones every bit flipped, so -0 sorts just below +0 and NaNs
 * by their bits at the ends.  No sortData_d field changes; the float mapping
 * is the reference's FloatToUint / UintToFloat (sort/src/kernels/common.cu:
 * 159-185), which the reference defines but never calls. */
#define SORT_KEYS_UINT32 0
#define SORT_KEYS_INT32 1
#define SORT_KEYS_FLOAT32 2
sortStatus_t SORTAPI sortTransformKeys(sortEngine_t engine, sortDevicePtr_t keys, int numElements,
                                       int keyType, bool toRadixOrder);

/* sortArray over int32 or float32 keys (keyType SORT_KEYS_INT32 or
 * SORT_KEYS_FLOAT32; SORT_KEYS_UINT32 is sortArray itself) in numeric order,
 * without the two passes over the keys that sortTransformKeys adds: the
 * count and the first pass map each key as they read it, and the last pass
 * maps it back as it writes it.  Results, buffer swaps, parity, earlyExit
 * and the bit range behave as sortTransformKeys, sortArray,
 * sortTransformKeys would (firstBit/endBit select bits of the mapped key);
 * keys come back in their own type.  No counterpart in the reference, which
 * sorts unsigned keys only. */
sortStatus_t SORTAPI sortArrayTyped(sortEngine_t engine, sortData_t data, int keyType);

/* Per-kernel timing: while enabled, every launch is bracketed by hipEvents
 * on the engine's stream.  sortReadTiming waits for the stream, fills up to
 * maxEntries aggregates (one per kernel name, in first-launch order), clears
 * the record and returns the number of kernels seen. */
typedef struct {
    char name[64];
    int launches;
    double totalMs;
    double minMs;
    double maxMs;
} sortKernelTiming_t;
sortStatus_t SORTAPI sortEnableTiming(sortEngine_t engine, bool enable);
int SORTAPI sortReadTiming(sortEngine_t engine, sortKernelTiming_t* out, int maxEntries);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* MGPUSORT_H */
