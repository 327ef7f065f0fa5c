// Sort service: continuous batching of sort requests on one GPU.
//
// No reference counterpart (the reference has only the synchronous
// sortArray, sort/src/mgpusort/mgpusort.cpp:607-616).  Each request is
// sorted exactly as sortArray would sort it alone; this file only decides
// when and next to what it runs:
//   * requests arrive on any thread (sortServerSubmit) and queue FIFO;
//   * a request is admitted when the ping-pong scratch it needs
//     (4 * n * (1 + value arrays) bytes) fits in what is left of a fixed HBM
//     budget, allocated from one slab (first fit);
//   * whenever one of `maxInFlight` slots is free, every admitted request is
//     packed into one batch: one segmented launch per (value count, bit
//     range) class on that slot's stream, so batches on different slots run
//     side by side on the device and requests that arrive meanwhile join the
//     next free slot instead of waiting for the whole device;
//   * results are sorted in place in the caller's buffers (a result that
//     ends on the scratch side is copied back on the slot's stream), and the
//     scratch returns to the budget when the slot's event completes.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "engine_internal.h"
#include "segmented_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
// Smallest batch worth a slot of its own when admitted work is split.
constexpr uint64_t kMinBatchKeys = 1ull << 22;
// A batch holds at most kMaxBatchRequests requests and, unless it is one
// request, at most kMaxBatchKeys keys.  Each slot's scratch (lookback words,
// per-segment histograms, the plan arrays of kReservedClasses classes) is
// sized for such a batch when the server starts: growing it mid-stream meant
// hipFree / hipHostMalloc calls that waited for the whole device, and batch
// launches 2-9 ms apart while every slot stalled.
constexpr uint32_t kMaxBatchRequests = 1024;
constexpr uint64_t kMaxBatchKeys = 1ull << 26;
constexpr size_t kReservedClasses = 12;

inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, size) with coalescing frees.
class SlabAllocator {
public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t bytes, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                *off = it->first;
                size_t rest = it->second - bytes;
                size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t bytes) {
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }

private:
    std::map<size_t, size_t> free_;
};

struct Pending {
    sortRequest_d req;
    double submitMs = 0;
    size_t scratchOff = 0;
    size_t scratchBytes = 0;
};

struct Slot {
    SortLane lane;
    std::vector<SegArena> arenas;  // one per class of a batch
    hipEvent_t done = nullptr;
    bool busy = false;
    int batch = -1;
    double launchMs = 0;
    std::vector<Pending> reqs;
    std::vector<SegPlanData> plans;
};

int carried(int valueCount) { return valueCount == -1 ? 1 : valueCount; }

size_t scratch_bytes(const sortRequest_d& r) {
    return (size_t)(1 + carried(r.valueCount)) * align_up(sizeof(uint32_t) * (size_t)r.numElements);
}

}  // namespace

struct sortServer_d {
    sortEngine_d* engine = nullptr;
    char* slab = nullptr;
    size_t slabBytes = 0;
    SlabAllocator alloc;
    std::vector<Slot> slots;
    int maxElements = 0;

    std::thread worker;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Pending> incoming;  // guarded by mu
    std::deque<sortCompletion_d> done;  // guarded by mu
    bool stop = false;  // guarded by mu

    // Worker-only state.
    std::deque<Pending> waiting;
    std::vector<Pending> admitted;
    int batchSeq = 0;
    Clock::time_point t0;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }
};

namespace {

void push_done(sortServer_d* s, const std::vector<sortCompletion_d>& c) {
    if (c.empty()) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->done.insert(s->done.end(), c.begin(), c.end());
    }
    s->cvDone.notify_all();
}

// Queues one batch on a free slot.
sortStatus_t launch_batch(sortServer_d* s, Slot& slot, std::vector<Pending>& reqs) {
    sortEngine_d* e = s->engine;
    slot.reqs.swap(reqs);
    slot.plans.clear();
    slot.batch = s->batchSeq++;
    slot.launchMs = s->now_ms();

    // Classes: (valueCount, firstBit, endBit).
    std::map<std::tuple<int, int, int>, std::vector<size_t>> classes;
    for (size_t i = 0; i < slot.reqs.size(); ++i) {
        const sortRequest_d& r = slot.reqs[i].req;
        classes[std::make_tuple(r.valueCount, r.firstBit, r.endBit)].push_back(i);
    }
    if (slot.arenas.size() < classes.size()) slot.arenas.resize(classes.size());
    slot.plans.resize(classes.size());
    size_t ci = 0;
    sortStatus_t st = SORT_STATUS_SUCCESS;
    for (auto& kv : classes) {
        const int vc = std::get<0>(kv.first), fb = std::get<1>(kv.first), eb = std::get<2>(kv.first);
        std::vector<mgpuSegDesc> segs(kv.second.size());
        for (size_t k = 0; k < kv.second.size(); ++k) {
            const Pending& p = slot.reqs[kv.second[k]];
            const sortRequest_d& r = p.req;
            mgpuSegDesc& d = segs[k];
            std::memset(&d, 0, sizeof(d));
            char* scratch = s->slab + p.scratchOff;
            const size_t arr = align_up(sizeof(uint32_t) * (size_t)r.numElements);
            d.keys[0] = reinterpret_cast<uint32_t*>(r.keys);
            d.keys[1] = reinterpret_cast<uint32_t*>(scratch);
            for (int v = 0; v < carried(vc); ++v) {
                d.vals[v][0] = reinterpret_cast<uint32_t*>(r.values[v]);
                d.vals[v][1] = reinterpret_cast<uint32_t*>(scratch + arr * (size_t)(1 + v));
            }
            d.n = (uint32_t)r.numElements;
        }
        SegPlanData& P = slot.plans[ci];
        st = mgpu_seg_build(e, slot.lane, segs.data(), (int)segs.size(), vc, fb, eb, &P,
                            &slot.arenas[ci]);
        if (!st) st = mgpu_seg_enqueue(P, slot.lane);
        if (st) break;
        // Results on the scratch side go back to the caller's buffers.
        if (P.resultSide == 1) {
            for (size_t k = 0; k < segs.size() && !st; ++k) {
                const mgpuSegDesc& d = segs[k];
                const size_t bytes = sizeof(uint32_t) * d.n;
                if (!bytes) continue;
                if (hipMemcpyAsync(d.keys[0], d.keys[1], bytes, hipMemcpyDeviceToDevice,
                                   slot.lane.stream) != hipSuccess)
                    st = SORT_STATUS_DEVICE_ERROR;
                for (int v = 0; v < carried(vc) && !st; ++v)
                    if (hipMemcpyAsync(d.vals[v][0], d.vals[v][1], bytes, hipMemcpyDeviceToDevice,
                                       slot.lane.stream) != hipSuccess)
                        st = SORT_STATUS_DEVICE_ERROR;
            }
        }
        if (st) break;
        ++ci;
    }
    if (!st && hipEventRecord(slot.done, slot.lane.stream) != hipSuccess)
        st = SORT_STATUS_DEVICE_ERROR;
    slot.busy = true;
    return st;
}

// Completes every request of the slot's batch with status st, or with
// SORT_STATUS_KERNEL_ERROR when a kernel of the batch failed a self-check
// (the slot's own error word, cleared for its next batch).
void finish_slot(sortServer_d* s, Slot& slot, sortStatus_t st) {
    const sortStatus_t kernel = mgpu_lane_take_error(slot.lane);
    if (!st) st = kernel;
    const double t = s->now_ms();
    std::vector<sortCompletion_d> c;
    c.reserve(slot.reqs.size());
    for (const Pending& p : slot.reqs) {
        sortCompletion_d d;
        std::memset(&d, 0, sizeof(d));
        d.tag = p.req.tag;
        d.status = st;
        d.batch = slot.batch;
        d.submitMs = p.submitMs;
        d.launchMs = slot.launchMs;
        d.doneMs = t;
        c.push_back(d);
        s->alloc.release(p.scratchOff, p.scratchBytes);
    }
    for (SegPlanData& P : slot.plans) mgpu_seg_release(P);
    slot.plans.clear();
    slot.reqs.clear();
    slot.busy = false;
    push_done(s, c);
}

void worker_main(sortServer_d* s) {
    (void)hipSetDevice(s->engine->device);
    while (true) {
        bool progress = false;
        {
            std::unique_lock<std::mutex> g(s->mu);
            while (!s->incoming.empty()) {
                s->waiting.push_back(s->incoming.front());
                s->incoming.pop_front();
            }
            bool busy = false;
            for (Slot& sl : s->slots) busy = busy || sl.busy;
            if (s->stop && s->waiting.empty() && s->admitted.empty() && !busy) break;
        }
        // Completions.
        for (Slot& slot : s->slots) {
            if (!slot.busy) continue;
            hipError_t q = hipEventQuery(slot.done);
            if (q == hipErrorNotReady) continue;
            finish_slot(s, slot, q == hipSuccess ? SORT_STATUS_SUCCESS : SORT_STATUS_DEVICE_ERROR);
            progress = true;
        }
        // Admission against the HBM budget, in arrival order.
        while (!s->waiting.empty()) {
            Pending& p = s->waiting.front();
            if (p.scratchBytes > s->slabBytes) {  // can never fit
                sortCompletion_d d;
                std::memset(&d, 0, sizeof(d));
                d.tag = p.req.tag;
                d.status = SORT_STATUS_CONFIG_NOT_SUPPORTED;
                d.batch = -1;
                d.submitMs = p.submitMs;
                d.launchMs = d.doneMs = s->now_ms();
                push_done(s, {d});
                s->waiting.pop_front();
                progress = true;
                continue;
            }
            size_t off;
            if (!s->alloc.alloc(p.scratchBytes, &off)) break;
            p.scratchOff = off;
            s->admitted.push_back(p);
            s->waiting.pop_front();
            progress = true;
        }
        // Launch what is admitted, spread over the free slots so that large
        // requests on different slots run side by side: each free slot takes
        // requests in arrival order up to an even share of the admitted keys
        // (at least kMinBatchKeys, at most the batch limits, and always at
        // least one request).
        if (!s->admitted.empty()) {
            size_t freeSlots = 0;
            for (Slot& slot : s->slots) freeSlots += !slot.busy;
            uint64_t keys = 0;
            for (const Pending& p : s->admitted) keys += (uint64_t)p.req.numElements;
            const uint64_t share =
                std::max<uint64_t>(kMinBatchKeys, freeSlots ? (keys + freeSlots - 1) / freeSlots : keys);
            size_t next = 0;
            for (Slot& slot : s->slots) {
                if (slot.busy || next >= s->admitted.size()) continue;
                std::vector<Pending> batch;
                uint64_t taken = 0;
                while (next < s->admitted.size()) {
                    const uint64_t n = (uint64_t)s->admitted[next].req.numElements;
                    if (!batch.empty() && (taken >= share || batch.size() >= kMaxBatchRequests ||
                                           taken + n > kMaxBatchKeys))
                        break;
                    taken += n;
                    batch.push_back(s->admitted[next++]);
                }
                sortStatus_t st = launch_batch(s, slot, batch);
                if (st) {
                    (void)hipStreamSynchronize(slot.lane.stream);
                    finish_slot(s, slot, st);
                }
                progress = true;
            }
            s->admitted.erase(s->admitted.begin(), s->admitted.begin() + (long)next);
        }
        if (!progress) {
            std::unique_lock<std::mutex> g(s->mu);
            s->cvWork.wait_for(g, std::chrono::microseconds(20),
                               [s] { return !s->incoming.empty() || s->stop; });
        }
    }
}

}  // namespace

extern "C" {

sortStatus_t SORTAPI sortServerCreate(sortEngine_t engine, unsigned long long hbmBudgetBytes,
                                      int maxInFlight, int maxElements, sortServer_t* server) {
    if (!engine || !server || maxInFlight < 1 || maxInFlight > 16 || maxElements < 1 ||
        maxElements > SORT_MAX_ELEMENTS || hbmBudgetBytes < kAlign)
        return SORT_STATUS_INVALID_VALUE;
    *server = nullptr;
    sortServer_d* s = new (std::nothrow) sortServer_d();
    if (!s) return SORT_STATUS_HOST_ALLOC_FAILED;
    s->engine = engine;
    s->maxElements = maxElements;
    s->slabBytes = (size_t)hbmBudgetBytes & ~(kAlign - 1);
    sortStatus_t st = SORT_STATUS_SUCCESS;
    if (hipMalloc(&s->slab, s->slabBytes) != hipSuccess) {
        s->slab = nullptr;
        st = SORT_STATUS_DEVICE_ALLOC_FAILED;
    }
    s->alloc.reset(s->slabBytes);
    s->slots.resize((size_t)maxInFlight);
    for (Slot& slot : s->slots) {
        if (st) break;
        if (hipStreamCreateWithFlags(&slot.lane.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess) {
            st = SORT_STATUS_DEVICE_ERROR;
            break;
        }
        st = mgpu_lane_init(&slot.lane);
        // Scratch for the largest batch, so that no batch grows it: lookback
        // words for every tile of its device-wide segments (each at most one
        // partial tile past its keys), a histogram per segment, plan arrays.
        const uint64_t batchKeys = std::max<uint64_t>(kMaxBatchKeys, (uint64_t)maxElements);
        const uint64_t tile = std::min(engine->small.tile(), engine->large.tile());
        const uint64_t tiles = batchKeys / tile + kMaxBatchRequests + 1;
        if (!st) st = mgpu_ensure_lookback_words(slot.lane, (uint64_t)MGPU_MAX_PASSES * tiles * MGPU_RADIX);
        if (!st) st = mgpu_ensure_multi(slot.lane, kMaxBatchRequests);
        // Each slot keeps its batch on its one stream: the side stream a
        // segmented sort forks its persistent kernels onto (a gain for one
        // sort alone) cost the server a third of its throughput, with batches
        // of up to 9 classes on 3 slots (requests all at once: 44-51 Gkeys/s
        // with it, 62-64 without, at 4 or 8 hardware queues alike).
        slot.lane.noSide = true;
        slot.arenas.resize(kReservedClasses);
        const size_t payload = mgpu_seg_payload_bound(kMaxBatchRequests, batchKeys, (uint32_t)tile);
        for (SegArena& a : slot.arenas)
            if (!st) st = mgpu_arena_reserve(&a, payload, kSegCounterWords, slot.lane.stream);
        if (!st && hipStreamSynchronize(slot.lane.stream) != hipSuccess) st = SORT_STATUS_DEVICE_ERROR;
    }
    if (st) {
        sortServerDestroy(s);
        return st;
    }
    s->t0 = Clock::now();
    s->worker = std::thread(worker_main, s);
    *server = s;
    return SORT_STATUS_SUCCESS;
}

sortStatus_t SORTAPI sortServerSubmit(sortServer_t server, const sortRequest_d* request) {
    if (!server || !request) return SORT_STATUS_INVALID_VALUE;
    const sortRequest_d& r = *request;
    if (r.numElements < 0 || r.numElements > server->maxElements || r.valueCount < -1 ||
        r.valueCount > 6 || r.firstBit < 0 || r.endBit > 32 || r.endBit < r.firstBit)
        return SORT_STATUS_INVALID_VALUE;
    if (r.numElements && !r.keys) return SORT_STATUS_INVALID_VALUE;
    for (int v = 0; v < carried(r.valueCount); ++v)
        if (r.numElements && !r.values[v]) return SORT_STATUS_INVALID_VALUE;
    Pending p;
    p.req = r;
    p.scratchBytes = std::max(kAlign, scratch_bytes(r));
    {
        std::lock_guard<std::mutex> g(server->mu);
        if (server->stop) return SORT_STATUS_NOT_INITIALIZED;
        p.submitMs = server->now_ms();
        server->incoming.push_back(p);
    }
    server->cvWork.notify_one();
    return SORT_STATUS_SUCCESS;
}

int SORTAPI sortServerPoll(sortServer_t server, sortCompletion_d* out, int maxCount,
                           int timeoutMs) {
    if (!server || !out || maxCount < 1) return -1;
    std::unique_lock<std::mutex> g(server->mu);
    if (server->done.empty() && timeoutMs > 0)
        server->cvDone.wait_for(g, std::chrono::milliseconds(timeoutMs),
                                [server] { return !server->done.empty(); });
    int n = 0;
    while (n < maxCount && !server->done.empty()) {
        out[n++] = server->done.front();
        server->done.pop_front();
    }
    return n;
}

sortStatus_t SORTAPI sortServerReplay(sortServer_t server, const sortRequest_d* requests,
                                      const double* arrivalMs, int count,
                                      sortCompletion_d* completions) {
    if (!server || count < 0 || (count && (!requests || !arrivalMs || !completions)))
        return SORT_STATUS_INVALID_VALUE;
    std::vector<char> seen((size_t)count, 0);
    sortStatus_t submitStatus = SORT_STATUS_SUCCESS;
    const double base = server->now_ms();
    std::thread submitter([&] {
        for (int i = 0; i < count; ++i) {
            const double target = base + arrivalMs[i];
            while (true) {
                const double wait = target - server->now_ms();
                if (wait <= 0) break;
                if (wait > 0.3)
                    std::this_thread::sleep_for(std::chrono::microseconds((long)((wait - 0.2) * 1e3)));
            }
            sortRequest_d r = requests[i];
            r.tag = (unsigned long long)i;
            sortStatus_t st = sortServerSubmit(server, &r);
            if (st) {
                submitStatus = st;
                sortCompletion_d d;
                std::memset(&d, 0, sizeof(d));
                d.tag = (unsigned long long)i;
                d.status = st;
                d.batch = -1;
                std::lock_guard<std::mutex> g(server->mu);
                server->done.push_back(d);
                server->cvDone.notify_all();
            }
        }
    });
    int got = 0;
    std::vector<sortCompletion_d> buf(256);
    while (got < count) {
        int n = sortServerPoll(server, buf.data(), (int)buf.size(), 100);
        for (int k = 0; k < n; ++k) {
            const unsigned long long t = buf[k].tag;
            if (t < (unsigned long long)count && !seen[t]) {
                seen[t] = 1;
                completions[t] = buf[k];
                ++got;
            }
        }
    }
    submitter.join();
    return submitStatus;
}

sortStatus_t SORTAPI sortServerDestroy(sortServer_t server) {
    if (!server) return SORT_STATUS_SUCCESS;
    {
        std::lock_guard<std::mutex> g(server->mu);
        server->stop = true;
    }
    server->cvWork.notify_all();
    if (server->worker.joinable()) server->worker.join();
    for (Slot& slot : server->slots) {
        if (slot.lane.stream) (void)hipStreamSynchronize(slot.lane.stream);
        for (SegPlanData& P : slot.plans) mgpu_seg_release(P);
        for (SegArena& a : slot.arenas) mgpu_arena_free(&a);
        mgpu_lane_free(&slot.lane);
        if (slot.done) (void)hipEventDestroy(slot.done);
        if (slot.lane.stream) (void)hipStreamDestroy(slot.lane.stream);
    }
    if (server->slab) (void)hipFree(server->slab);
    delete server;
    return SORT_STATUS_SUCCESS;
}

}  // extern "C"
