// A C++ client of the mgpusort C API written the way the reference's own
// callers use it, built with g++ against include/ and libmgpusort.so only:
//   1. simpletest2-style (sort/src/simpletest2/simpletest2.cpp:57-150):
//      caller buffers attached through MgpuSortData, bits [7, 25), two value
//      streams (uint and float), then index generation with earlyExit;
//   2. bwt-style (bwt/src/mgpubwt/mgpubwt.cpp:262-292): multi-word keys
//      sorted least significant word first by repeated sortArray calls that
//      carry the remaining words plus an index stream as values, with the
//      caller tracking buffers through `parity`;
//   3. the additions with no reference counterpart: a segmented sort of 600
//      arrays (sortSegmentedArray, then a captured plan replayed twice), 40
//      requests of mixed sizes, bit widths and value modes through the sort
//      service, and int32 and float32 keys through sortArrayTyped.
// Every result is checked against std::stable_sort on the host.  Prints
// "CLIENT OK" on success and exits non-zero on the first mismatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "mgpusort.hpp"

#define CHECK(x)                                                              \
    do {                                                                      \
        if (!(x)) {                                                           \
            std::printf("CHECK failed at line %d: %s\n", __LINE__, #x);      \
            return 1;                                                         \
        }                                                                     \
    } while (0)

static uint32_t* dalloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * 4) != hipSuccess) return nullptr;
    return static_cast<uint32_t*>(p);
}

static int simple_test(MgpuSort& engine) {
    const int n = 100000;
    std::mt19937 rng(7);
    std::vector<uint32_t> keys(n), v1(n);
    std::vector<float> v2(n);
    for (int i = 0; i < n; ++i) {
        keys[i] = (rng() & ((1u << 18) - 1)) << 7;
        v1[i] = i;
        v2[i] = std::sqrt((float)i);
    }
    uint32_t *dk = dalloc(n), *dv1 = dalloc(n), *dv2 = dalloc(n);
    CHECK(dk && dv1 && dv2);
    CHECK(hipMemcpy(dk, keys.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);
    CHECK(hipMemcpy(dv1, v1.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);
    CHECK(hipMemcpy(dv2, v2.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);

    MgpuSortData data;
    data.AttachKey(dk);
    data.AttachVal(0, dv1);
    data.AttachVal(1, dv2);
    CHECK(data.Alloc(engine, n, 2) == SORT_STATUS_SUCCESS);
    data.firstBit = 7;
    data.endBit = 7 + 18;
    CHECK(sortArray(engine, &data) == SORT_STATUS_SUCCESS);

    std::vector<uint32_t> ok(n), o1(n);
    std::vector<float> o2(n);
    CHECK(hipMemcpy(ok.data(), data.pKeys[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(o1.data(), data.pValues1[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(o2.data(), data.pValues2[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    auto bf = [](uint32_t k) { return (k >> 7) & ((1u << 18) - 1); };
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int a, int b) { return bf(keys[a]) < bf(keys[b]); });
    for (int i = 0; i < n; ++i) {
        CHECK(ok[i] == keys[perm[i]]);
        CHECK(o1[i] == v1[perm[i]]);
        CHECK(std::memcmp(&o2[i], &v2[perm[i]], 4) == 0);
    }
    data.Reset();  // frees only the library's buffers

    // Index generation with early exit over the full 32 bits.
    CHECK(hipMemcpy(dk, keys.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);
    data.AttachKey(dk);
    CHECK(data.Alloc(engine, n, SORT_VALUE_INDEX) == SORT_STATUS_SUCCESS);
    data.firstBit = 0;
    data.endBit = 32;
    data.earlyExit = true;
    CHECK(sortArray(engine, &data) == SORT_STATUS_SUCCESS);
    CHECK(hipMemcpy(ok.data(), data.pKeys[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(o1.data(), data.pValues1[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    for (int i = 0; i < n; ++i) {
        CHECK(ok[i] == keys[perm[i]]);
        CHECK(o1[i] == (uint32_t)perm[i]);
    }
    data.Reset();
    (void)hipFree(dk);
    (void)hipFree(dv1);
    (void)hipFree(dv2);
    return 0;
}

// Keys of `words` 32-bit words (word 0 most significant), sorted LSD by word.
static int bwt_style_test(MgpuSort& engine) {
    const int n = 50000, words = 3;
    std::mt19937 rng(11);
    std::vector<std::vector<uint32_t>> host(words, std::vector<uint32_t>(n));
    for (int w = 0; w < words; ++w)
        for (int i = 0; i < n; ++i) host[w][i] = rng() & (w == 0 ? 0xFFu : 0xFFFFFFFFu);
    // streams[w][p]: the two buffers of word w; streams[words] the indices.
    uint32_t* streams[words + 1][2];
    for (int w = 0; w <= words; ++w)
        for (int p = 0; p < 2; ++p) CHECK((streams[w][p] = dalloc(n)) != nullptr);
    for (int w = 0; w < words; ++w)
        CHECK(hipMemcpy(streams[w][0], host[w].data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    CHECK(hipMemcpy(streams[words][0], idx.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);

    sortData_d d;
    std::memset(&d, 0, sizeof(d));
    d.maxElements = d.numElements = n;
    d.parity = 0;
    for (int cur = words - 1; cur >= 0; --cur) {
        const int p = d.parity;
        d.pKeys[0] = streams[cur][p];
        d.pKeys[1] = streams[cur][1 - p];
        for (int i = 0; i < cur; ++i) {
            d.pValues[i][0] = streams[i][p];
            d.pValues[i][1] = streams[i][1 - p];
        }
        d.pValues[cur][0] = streams[words][p];
        d.pValues[cur][1] = streams[words][1 - p];
        d.valueCount = cur + 1;
        d.firstBit = 0;
        d.endBit = cur == 0 ? 8 : 32;  // word 0 holds 8 significant bits
        CHECK(sortArray(engine, &d) == SORT_STATUS_SUCCESS);
    }
    // After the chain, word 0 and the indices are at side parity.
    std::vector<uint32_t> outIdx(n), outW0(n);
    CHECK(hipMemcpy(outIdx.data(), streams[words][d.parity], 4 * n, hipMemcpyDeviceToHost) ==
          hipSuccess);
    CHECK(hipMemcpy(outW0.data(), streams[0][d.parity], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
        for (int w = 0; w < words; ++w)
            if (host[w][a] != host[w][b]) return host[w][a] < host[w][b];
        return false;
    });
    for (int i = 0; i < n; ++i) {
        CHECK(outIdx[i] == (uint32_t)perm[i]);
        CHECK(outW0[i] == host[0][perm[i]]);
    }
    for (int w = 0; w <= words; ++w)
        for (int p = 0; p < 2; ++p) (void)hipFree(streams[w][p]);
    return 0;
}

// Host check of one array: keys (and `vals`, gathered) stably sorted by the
// bit field [0, endBit) of the keys; indices when index is true.
static bool sorted_like_host(const std::vector<uint32_t>& keys,
                             const std::vector<std::vector<uint32_t>>& vals, int endBit,
                             bool index, const std::vector<uint32_t>& outKeys,
                             const std::vector<std::vector<uint32_t>>& outVals) {
    const size_t n = keys.size();
    const uint32_t mask = endBit >= 32 ? 0xFFFFFFFFu : (1u << endBit) - 1u;
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return (keys[a] & mask) < (keys[b] & mask); });
    for (size_t i = 0; i < n; ++i) {
        if (outKeys[i] != keys[perm[i]]) return false;
        if (index && outVals[0][i] != perm[i]) return false;
        if (!index)
            for (size_t v = 0; v < vals.size(); ++v)
                if (outVals[v][i] != vals[v][perm[i]]) return false;
    }
    return true;
}

static int segmented_test(MgpuSort& engine) {
    std::mt19937 rng(13);
    const int segs = 600;
    std::vector<int> starts(segs + 1, 0);
    std::uniform_real_distribution<double> u(0.0, 16.0);
    for (int i = 0; i < segs; ++i) starts[i + 1] = starts[i] + (int)std::pow(2.0, u(rng));
    const int n = starts[segs];
    std::vector<uint32_t> keys(n), vals(n);
    for (int i = 0; i < n; ++i) {
        keys[i] = rng();
        vals[i] = (uint32_t)i * 3u;
    }
    uint32_t *k0 = dalloc(n), *k1 = dalloc(n), *v0 = dalloc(n), *v1 = dalloc(n);
    CHECK(k0 && k1 && v0 && v1);
    sortData_d d;
    std::memset(&d, 0, sizeof(d));
    d.maxElements = d.numElements = n;
    d.valueCount = 1;
    d.firstBit = 0;
    d.endBit = 32;
    d.pKeys[0] = k0;
    d.pKeys[1] = k1;
    d.pValues[0][0] = v0;
    d.pValues[0][1] = v1;
    auto upload = [&]() {
        return hipMemcpy(d.pKeys[0], keys.data(), 4 * (size_t)n, hipMemcpyHostToDevice) ==
                   hipSuccess &&
               hipMemcpy(d.pValues[0][0], vals.data(), 4 * (size_t)n, hipMemcpyHostToDevice) ==
                   hipSuccess;
    };
    auto check = [&]() {
        std::vector<uint32_t> ok(n), ov(n);
        if (hipMemcpy(ok.data(), d.pKeys[0], 4 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ov.data(), d.pValues[0][0], 4 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
            return false;
        for (int s = 0; s < segs; ++s) {
            const int a = starts[s], b = starts[s + 1];
            std::vector<uint32_t> sk(keys.begin() + a, keys.begin() + b);
            std::vector<std::vector<uint32_t>> sv{{vals.begin() + a, vals.begin() + b}};
            std::vector<uint32_t> gk(ok.begin() + a, ok.begin() + b);
            std::vector<std::vector<uint32_t>> gv{{ov.begin() + a, ov.begin() + b}};
            if (!sorted_like_host(sk, sv, 32, false, gk, gv)) return false;
        }
        return true;
    };
    CHECK(upload());
    CHECK(sortSegmentedArray(engine, &d, segs, starts.data()) == SORT_STATUS_SUCCESS);
    CHECK(check());
    sortSegPlan_t plan = nullptr;
    CHECK(sortSegmentedPlanCreate(engine, &d, segs, starts.data(), &plan) == SORT_STATUS_SUCCESS);
    for (int rep = 0; rep < 2; ++rep) {
        CHECK(upload());
        CHECK(sortSegmentedPlanExecute(plan, &d) == SORT_STATUS_SUCCESS);
        CHECK(hipDeviceSynchronize() == hipSuccess);
        CHECK(check());
    }
    CHECK(sortSegmentedPlanDestroy(plan) == SORT_STATUS_SUCCESS);
    for (uint32_t* p : {k0, k1, v0, v1}) (void)hipFree(p);
    return 0;
}

static int server_test(MgpuSort& engine) {
    sortServer_t srv = nullptr;
    CHECK(sortServerCreate(engine, 1ull << 30, 3, 1 << 20, &srv) == SORT_STATUS_SUCCESS);
    std::mt19937 rng(17);
    std::uniform_real_distribution<double> u(0.0, 18.0);
    const int count = 40;
    static const int kWidths[3] = {16, 24, 32}, kValueCounts[4] = {0, 1, 2, -1};
    struct Req {
        std::vector<uint32_t> keys;
        std::vector<std::vector<uint32_t>> vals;
        uint32_t* dk = nullptr;
        uint32_t* dv[2] = {nullptr, nullptr};
        int vc = 0, endBit = 32;
    };
    std::vector<Req> reqs(count);
    for (int i = 0; i < count; ++i) {
        Req& r = reqs[i];
        const int n = (int)std::pow(2.0, u(rng));
        r.endBit = kWidths[i % 3];
        r.vc = kValueCounts[i % 4];
        const uint32_t mask = r.endBit >= 32 ? 0xFFFFFFFFu : (1u << r.endBit) - 1u;
        r.keys.resize(n);
        for (int j = 0; j < n; ++j) r.keys[j] = rng() & mask;
        CHECK((r.dk = dalloc(n)) != nullptr);
        CHECK(hipMemcpy(r.dk, r.keys.data(), 4 * (size_t)n, hipMemcpyHostToDevice) == hipSuccess);
        const int arrays = r.vc == -1 ? 1 : r.vc;
        for (int v = 0; v < arrays; ++v) {
            CHECK((r.dv[v] = dalloc(n)) != nullptr);
            if (r.vc > 0) {
                r.vals.emplace_back(n);
                for (int j = 0; j < n; ++j) r.vals.back()[j] = rng();
                CHECK(hipMemcpy(r.dv[v], r.vals.back().data(), 4 * (size_t)n,
                                hipMemcpyHostToDevice) == hipSuccess);
            }
        }
        sortRequest_d q;
        std::memset(&q, 0, sizeof(q));
        q.keys = (CUdeviceptr)(uintptr_t)r.dk;
        for (int v = 0; v < arrays; ++v) q.values[v] = (CUdeviceptr)(uintptr_t)r.dv[v];
        q.numElements = n;
        q.valueCount = r.vc;
        q.firstBit = 0;
        q.endBit = r.endBit;
        q.tag = (unsigned long long)i;
        CHECK(sortServerSubmit(srv, &q) == SORT_STATUS_SUCCESS);
    }
    int done = 0;
    sortCompletion_d comps[64];
    for (int polls = 0; done < count && polls < 120; ++polls) {
        const int got = sortServerPoll(srv, comps, 64, 1000);
        CHECK(got >= 0);
        for (int c = 0; c < got; ++c) CHECK(comps[c].status == SORT_STATUS_SUCCESS);
        done += got;
    }
    CHECK(done == count);
    CHECK(sortServerDestroy(srv) == SORT_STATUS_SUCCESS);
    for (Req& r : reqs) {
        const size_t n = r.keys.size();
        const int arrays = r.vc == -1 ? 1 : r.vc;
        std::vector<uint32_t> ok(n);
        std::vector<std::vector<uint32_t>> ov(arrays, std::vector<uint32_t>(n));
        CHECK(hipMemcpy(ok.data(), r.dk, 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
        for (int v = 0; v < arrays; ++v)
            CHECK(hipMemcpy(ov[v].data(), r.dv[v], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
        CHECK(sorted_like_host(r.keys, r.vals, r.endBit, r.vc == -1, ok, ov));
        (void)hipFree(r.dk);
        for (int v = 0; v < arrays; ++v) (void)hipFree(r.dv[v]);
    }
    return 0;
}

// sortArrayTyped with index generation: int32 keys, then float32 keys (no
// NaNs or zeros, so numeric order is the sort's order), each against
// std::stable_sort by value.
static int typed_test(MgpuSort& engine) {
    const int n = 70001;
    std::mt19937 rng(21);
    std::normal_distribution<float> normal(0.f, 1000.f);
    for (int type : {SORT_KEYS_INT32, SORT_KEYS_FLOAT32}) {
        std::vector<uint32_t> keys(n);
        for (int i = 0; i < n; ++i) {
            if (type == SORT_KEYS_INT32) {
                keys[i] = rng();
            } else {
                float f = normal(rng);
                if (f == 0.f) f = 1.f;
                std::memcpy(&keys[i], &f, 4);
            }
        }
        uint32_t* dk = dalloc(n);
        CHECK(dk);
        CHECK(hipMemcpy(dk, keys.data(), 4 * n, hipMemcpyHostToDevice) == hipSuccess);
        MgpuSortData data;
        data.AttachKey(dk);
        CHECK(data.Alloc(engine, n, SORT_VALUE_INDEX) == SORT_STATUS_SUCCESS);
        data.firstBit = 0;
        data.endBit = 32;
        CHECK(sortArrayTyped(engine, &data, type) == SORT_STATUS_SUCCESS);
        std::vector<uint32_t> ok(n), oi(n);
        CHECK(hipMemcpy(ok.data(), data.pKeys[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
        CHECK(hipMemcpy(oi.data(), data.pValues1[0], 4 * n, hipMemcpyDeviceToHost) == hipSuccess);
        std::vector<uint32_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0u);
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
            if (type == SORT_KEYS_INT32) return (int32_t)keys[a] < (int32_t)keys[b];
            float fa, fb;
            std::memcpy(&fa, &keys[a], 4);
            std::memcpy(&fb, &keys[b], 4);
            return fa < fb;
        });
        for (int i = 0; i < n; ++i) {
            CHECK(ok[i] == keys[perm[i]]);
            CHECK(oi[i] == perm[i]);
        }
        data.Reset();
        (void)hipFree(dk);
    }
    return 0;
}

int main(int argc, char** argv) {
    const char* path = argc > 1 ? argv[1] : nullptr;
    try {
        MgpuSort engine(path);
        if (simple_test(engine)) return 1;
        if (bwt_style_test(engine)) return 1;
        if (segmented_test(engine)) return 1;
        if (server_test(engine)) return 1;
        if (typed_test(engine)) return 1;
    } catch (const MgpuException& e) {
        std::printf("MgpuException: %s\n", e.what());
        return 2;
    }
    std::printf("CLIENT OK\n");
    return 0;
}
