// CPU oracle for the mgpusort LSD radix sort path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file: it is the checker that tests/, the smoke check and the
// CPU-baseline leg of bench.py compare the HIP path against.
//
// Reference path (seanbaxter/mgpu_old), what each function here restates:
//   * oracle_stable_sort      -- the closed form of sortArray
//       (sort/src/mgpusort/mgpusort.cpp:560-616): an LSD radix sort whose every
//       pass ranks stably is a stable sort of the keys by the bit field
//       bfe(key, firstBit, endBit - firstBit) (sort/src/kernels/common.cu:37-41),
//       carrying 0..6 value arrays through the same permutation
//       (mgpusort.cpp:491-507), or emitting source indices into values1 when
//       valueCount == -1 (inc/mgpusort.h:70-73, mgpusort.cpp:553).  Written as
//       std::stable_sort of (bitfield, source index) pairs, then a gather.
//   * oracle_sort1            -- cpusort.cpp:18-72 (Sort1Pass / Sort1), the
//       sequential counting-sort pass, generalised to any [firstBit, endBit)
//       and to value arrays.
//   * oracle_sort2            -- cpusort.cpp:85-192 (Sort2CountPass,
//       Sort2SortPass, Sort2Pass, Sort2): per-block count, sequential reduce,
//       exclusive scan, per-block scatter.  The two "PARALLEL PART" loops
//       (cpusort.cpp:129-138, 171-182) run on `threads` std::threads.  The
//       original does not compile on LP64 (std::min(size_t, uint) at :136),
//       so it is restated, not built.
//   * oracle_block_counts     -- the per-block digit histograms of
//       cpusort.cpp:133-138 / count.cu (known answers for a device digit
//       count), and oracle_digit_offsets the scatter offsets of
//       cpusort.cpp:144-169 (hist1/hist2/hist3).
//   * oracle_splitmix64_*     -- the seeded input generator shared with the
//       workload module (SplitMix64, Steele/Lea/Flood 2014; counter form).
//   * oracle_key_transform    -- signed and floating keys to the unsigned
//       order the sort uses and back: int32 by flipping the sign bit, float32
//       in the steps of the reference's FloatToUint / UintToFloat
//       (sort/src/kernels/common.cu:159-185; defined there, never called).
//
// Pinning: tests/test_oracle.py checks oracle_key_transform against numpy's
// int32 and float32 orderings and for a round trip; it checks oracle_stable_sort against numpy's
// stable argsort, against brute force over every permutation of small
// inputs, against the invariants that fix a stable sort uniquely, and
// oracle_sort1/oracle_sort2 against oracle_stable_sort; splitmix64 against
// its published first output for seed 0.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint32_t bitfield(uint32_t k, int firstBit, int numBits) {
    uint32_t x = k >> firstBit;
    return numBits >= 32 ? x : (x & ((1u << numBits) - 1u));
}

inline uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

const uint64_t kGamma = 0x9e3779b97f4a7c15ULL;

int check_range(int firstBit, int endBit) {
    if (firstBit < 0 || endBit > 32 || endBit < firstBit) return -1;
    return 0;
}

// One counting-sort pass of cpusort.cpp Sort2Pass over [0, count) with blocks
// of blockSize keys.  perm is permuted alongside the keys so that any number
// of value arrays can be gathered afterwards.
void sort2_pass(std::vector<uint32_t>& keys, std::vector<uint32_t>& perm,
                int bit, int numBits, size_t blockSize, int threads) {
    const size_t count = keys.size();
    const uint32_t numBuckets = 1u << numBits;
    const size_t numBlocks = (count + blockSize - 1) / blockSize;
    std::vector<uint32_t> countsBlock(numBuckets * numBlocks, 0);

    auto parallel_blocks = [&](auto&& body) {
        if (threads <= 1 || numBlocks < 2) {
            for (size_t b = 0; b < numBlocks; ++b) body(b);
            return;
        }
        std::vector<std::thread> pool;
        int nt = std::min<int>(threads, (int)numBlocks);
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                size_t lo = numBlocks * t / nt, hi = numBlocks * (t + 1) / nt;
                for (size_t b = lo; b < hi; ++b) body(b);
            });
        for (auto& th : pool) th.join();
    };

    // PARALLEL PART (cpusort.cpp:129-138): per-block digit histogram.
    parallel_blocks([&](size_t block) {
        size_t index = block * blockSize;
        size_t n = std::min(count - index, blockSize);
        uint32_t* c = &countsBlock[block * numBuckets];
        for (size_t i = 0; i < n; ++i) ++c[bitfield(keys[index + i], bit, numBits)];
    });

    // REDUCTION PART (cpusort.cpp:144-169): global totals, exclusive scan over
    // digits, then running per-block offsets.
    std::vector<uint64_t> countsGlobal(numBuckets, 0), excScanGlobal(numBuckets, 0);
    for (size_t block = 0; block < numBlocks; ++block)
        for (uint32_t d = 0; d < numBuckets; ++d)
            countsGlobal[d] += countsBlock[block * numBuckets + d];
    for (uint32_t d = 1; d < numBuckets; ++d)
        excScanGlobal[d] = excScanGlobal[d - 1] + countsGlobal[d - 1];
    std::vector<uint64_t> excScanBlock(numBuckets * numBlocks);
    std::vector<uint64_t> digitOffsets(numBuckets, 0);
    for (size_t block = 0; block < numBlocks; ++block)
        for (uint32_t d = 0; d < numBuckets; ++d) {
            excScanBlock[block * numBuckets + d] = digitOffsets[d] + excScanGlobal[d];
            digitOffsets[d] += countsBlock[block * numBuckets + d];
        }

    // PARALLEL PART (cpusort.cpp:171-182): per-block scatter, keys visited in
    // order so that equal digits keep their order (the stable rank).
    std::vector<uint32_t> sortedKeys(count), sortedPerm(count);
    parallel_blocks([&](size_t block) {
        size_t index = block * blockSize;
        size_t n = std::min(count - index, blockSize);
        std::vector<uint64_t> off(excScanBlock.begin() + block * numBuckets,
                                  excScanBlock.begin() + (block + 1) * numBuckets);
        for (size_t i = 0; i < n; ++i) {
            uint32_t k = keys[index + i];
            uint64_t dst = off[bitfield(k, bit, numBits)]++;
            sortedKeys[dst] = k;
            sortedPerm[dst] = perm[index + i];
        }
    });
    keys.swap(sortedKeys);
    perm.swap(sortedPerm);
}

void gather_outputs(const uint32_t* const* values, int valueCount, size_t n,
                    const std::vector<uint32_t>& perm, uint32_t* const* valuesOut) {
    if (valueCount == -1) {
        if (n) std::memcpy(valuesOut[0], perm.data(), n * sizeof(uint32_t));
        return;
    }
    for (int v = 0; v < valueCount; ++v)
        for (size_t i = 0; i < n; ++i) valuesOut[v][i] = values[v][perm[i]];
}

}  // namespace

extern "C" {

// i-th output of SplitMix64 seeded with `seed` (counter form: the state after
// i+1 steps is seed + (i+1)*gamma).
uint64_t oracle_splitmix64_at(uint64_t seed, uint64_t i) {
    return splitmix_mix(seed + (i + 1) * kGamma);
}

// Workload keys: word(i) = high 32 bits of the i-th SplitMix64 output.
//   kind 0: uniform u32 keys, key[i] = word(i)
//   kind 1: low-entropy keys, key[i] = word(3i) & word(3i+1) & word(3i+2)
//           (each bit set with probability 1/8)
// Bits at and above maskBits are cleared (maskBits 32 keeps all).
// type 1: int32, 2: float32; toRadix 1: to the unsigned order, 0: back.
// float32 follows common.cu:159-185 step by step: to the order, all but the
// sign bit flipped when negative (x ^ (0x7fffffff & (x >> 31))), then 2^31
// added; back, 2^31 subtracted and the same conditional flip.
int oracle_key_transform(const uint32_t* in, uint64_t n, int type, int toRadix, uint32_t* out) {
    if (type != 1 && type != 2) return -1;
    for (uint64_t i = 0; i < n; ++i) {
        if (type == 1) {
            out[i] = in[i] ^ 0x80000000u;
        } else if (toRadix) {
            const int32_t x = (int32_t)in[i];
            const int32_t bits = x >> 31;
            const int32_t flipped = x ^ (0x7fffffff & bits);
            out[i] = (uint32_t)flipped + 0x80000000u;
        } else {
            const uint32_t flippedU = in[i] - 0x80000000u;
            const int32_t flipped = (int32_t)flippedU;
            const int32_t bits = flipped >> 31;
            out[i] = (uint32_t)(flipped ^ (0x7fffffff & bits));
        }
    }
    return 0;
}

void oracle_fill_keys(uint32_t* out, uint64_t n, uint64_t seed, int kind, int maskBits) {
    uint32_t mask = maskBits >= 32 ? 0xffffffffu : ((1u << maskBits) - 1u);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t k;
        if (kind == 1)
            k = (uint32_t)(oracle_splitmix64_at(seed, 3 * i) >> 32) &
                (uint32_t)(oracle_splitmix64_at(seed, 3 * i + 1) >> 32) &
                (uint32_t)(oracle_splitmix64_at(seed, 3 * i + 2) >> 32);
        else
            k = (uint32_t)(oracle_splitmix64_at(seed, i) >> 32);
        out[i] = k & mask;
    }
}

// Closed form of sortArray: stable sort by bitfield(key, firstBit, endBit).
// values: valueCount input arrays (ignored when valueCount <= 0).
// valuesOut: valueCount output arrays, or one array of source indices when
// valueCount == -1.  Returns 0, or -1 on a bad argument.
int oracle_stable_sort(const uint32_t* keys, const uint32_t* const* values, int valueCount,
                       uint64_t n, int firstBit, int endBit, uint32_t* keysOut,
                       uint32_t* const* valuesOut) {
    if (check_range(firstBit, endBit) || valueCount < -1 || valueCount > 6) return -1;
    const int numBits = endBit - firstBit;
    std::vector<std::pair<uint32_t, uint32_t>> pairs(n);
    for (uint64_t i = 0; i < n; ++i)
        pairs[i] = {numBits ? bitfield(keys[i], firstBit, numBits) : 0u, (uint32_t)i};
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const std::pair<uint32_t, uint32_t>& a,
                        const std::pair<uint32_t, uint32_t>& b) { return a.first < b.first; });
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = pairs[i].second;
    for (uint64_t i = 0; i < n; ++i) keysOut[i] = keys[perm[i]];
    gather_outputs(values, valueCount, n, perm, valuesOut);
    return 0;
}

// cpusort.cpp Sort1 (:18-72): sequential counting-sort passes of digitBits
// bits from firstBit up to endBit (the last pass narrower when the range is
// not a multiple of digitBits).
int oracle_sort1(const uint32_t* keys, const uint32_t* const* values, int valueCount,
                 uint64_t n, int firstBit, int endBit, int digitBits, uint32_t* keysOut,
                 uint32_t* const* valuesOut) {
    if (check_range(firstBit, endBit) || digitBits < 1 || digitBits > 16 ||
        valueCount < -1 || valueCount > 6)
        return -1;
    std::vector<uint32_t> k(keys, keys + n), perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    for (int bit = firstBit; bit < endBit; bit += digitBits) {
        int nb = std::min(digitBits, endBit - bit);
        // One block covering the whole array is exactly Sort1Pass.
        sort2_pass(k, perm, bit, nb, n ? n : 1, 1);
    }
    if (n) std::memcpy(keysOut, k.data(), n * sizeof(uint32_t));
    gather_outputs(values, valueCount, n, perm, valuesOut);
    return 0;
}

// cpusort.cpp Sort2 (:85-192) with a blockSize-key block decomposition and
// `threads` worker threads for the two parallel parts.
int oracle_sort2(const uint32_t* keys, const uint32_t* const* values, int valueCount,
                 uint64_t n, int firstBit, int endBit, int digitBits, uint64_t blockSize,
                 int threads, uint32_t* keysOut, uint32_t* const* valuesOut) {
    if (check_range(firstBit, endBit) || digitBits < 1 || digitBits > 16 ||
        blockSize == 0 || valueCount < -1 || valueCount > 6)
        return -1;
    std::vector<uint32_t> k(keys, keys + n), perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    for (int bit = firstBit; bit < endBit; bit += digitBits) {
        int nb = std::min(digitBits, endBit - bit);
        sort2_pass(k, perm, bit, nb, blockSize, threads);
    }
    if (n) std::memcpy(keysOut, k.data(), n * sizeof(uint32_t));
    gather_outputs(values, valueCount, n, perm, valuesOut);
    return 0;
}

// Per-block digit histograms (cpusort.cpp:133-138): counts[block][digit] for
// blocks of blockSize keys, digit = bitfield(key, bit, numBits).
int oracle_block_counts(const uint32_t* keys, uint64_t n, int bit, int numBits,
                        uint64_t blockSize, uint32_t* counts) {
    if (numBits < 1 || numBits > 16 || bit < 0 || bit + numBits > 32 || !blockSize) return -1;
    uint64_t numBuckets = 1ull << numBits;
    uint64_t numBlocks = (n + blockSize - 1) / blockSize;
    if (numBlocks) std::memset(counts, 0, numBlocks * numBuckets * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; ++i)
        ++counts[(i / blockSize) * numBuckets + bitfield(keys[i], bit, numBits)];
    return 0;
}

// Scatter offsets of cpusort.cpp:144-169: offsets[block][digit] = start of
// block's digit run in the output of one pass.
int oracle_digit_offsets(const uint32_t* counts, uint64_t numBlocks, int numBits,
                         uint64_t* offsets) {
    if (numBits < 1 || numBits > 16) return -1;
    uint64_t numBuckets = 1ull << numBits;
    std::vector<uint64_t> total(numBuckets, 0), start(numBuckets, 0);
    for (uint64_t b = 0; b < numBlocks; ++b)
        for (uint64_t d = 0; d < numBuckets; ++d) total[d] += counts[b * numBuckets + d];
    for (uint64_t d = 1; d < numBuckets; ++d) start[d] = start[d - 1] + total[d - 1];
    for (uint64_t b = 0; b < numBlocks; ++b)
        for (uint64_t d = 0; d < numBuckets; ++d) {
            offsets[b * numBuckets + d] = start[d];
            start[d] += counts[b * numBuckets + d];
        }
    return 0;
}

}  // extern "C"
