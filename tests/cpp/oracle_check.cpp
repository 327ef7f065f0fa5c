// Runs the CPU oracle (oracle/oracle.cpp) over seeded random cases so that
// tests/test_sanitizers.py can build both with AddressSanitizer and
// UndefinedBehaviorSanitizer and catch a memory or arithmetic error in the
// checker itself.  Each case also cross-checks the oracle's three sorts
// (closed form, Sort1 and Sort2 restatements) and the key transforms' round
// trip, so a sanitizer-clean run is a correct run.  Exit status 0 on success.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
uint64_t oracle_splitmix64_at(uint64_t seed, uint64_t i);
void oracle_fill_keys(uint32_t* out, uint64_t n, uint64_t seed, int kind, int maskBits);
int oracle_key_transform(const uint32_t* in, uint64_t n, int type, int toRadix, uint32_t* out);
int oracle_stable_sort(const uint32_t* keys, const uint32_t* const* values, int valueCount,
                       uint64_t n, int firstBit, int endBit, uint32_t* keysOut,
                       uint32_t* const* valuesOut);
int oracle_sort1(const uint32_t* keys, const uint32_t* const* values, int valueCount, uint64_t n,
                 int firstBit, int endBit, int digitBits, uint32_t* keysOut,
                 uint32_t* const* valuesOut);
int oracle_sort2(const uint32_t* keys, const uint32_t* const* values, int valueCount, uint64_t n,
                 int firstBit, int endBit, int digitBits, uint64_t blockSize, int threads,
                 uint32_t* keysOut, uint32_t* const* valuesOut);
int oracle_block_counts(const uint32_t* keys, uint64_t n, int bit, int numBits,
                        uint64_t blockSize, uint32_t* counts);
int oracle_digit_offsets(const uint32_t* counts, uint64_t numBlocks, int numBits,
                         uint64_t* offsets);
}

namespace {

int failures = 0;

void check(bool ok, const char* what, int c) {
    if (!ok) {
        std::fprintf(stderr, "case %d: %s\n", c, what);
        ++failures;
    }
}

uint32_t word(uint64_t seed, uint64_t i) { return (uint32_t)(oracle_splitmix64_at(seed, i) >> 32); }

}  // namespace

int main() {
    for (int c = 0; c < 300; ++c) {
        const uint64_t n = word(1, c) % (c < 20 ? 8 : 5000);  // tiny and empty cases first
        const int vc = (int)(word(2, c) % 8) - 1;             // -1..6
        const int fb = (int)(word(3, c) % 32);
        const int eb = fb + 1 + (int)(word(4, c) % (32 - fb));
        const int kind = (int)(word(5, c) % 2);
        std::vector<uint32_t> keys(n);
        oracle_fill_keys(keys.data(), n, 100 + c, kind, 32);
        const int nv = vc < 0 ? 0 : vc, nout = vc == -1 ? 1 : vc;
        std::vector<std::vector<uint32_t>> vals(nv, std::vector<uint32_t>(n));
        std::vector<const uint32_t*> vin(7, nullptr);
        for (int v = 0; v < nv; ++v) {
            oracle_fill_keys(vals[v].data(), n, 200 + 7 * c + v, 0, 32);
            vin[v] = vals[v].data();
        }
        std::vector<uint32_t> k0(n), k1(n), k2(n);
        std::vector<std::vector<uint32_t>> o0(nout, std::vector<uint32_t>(n)), o1 = o0, o2 = o0;
        std::vector<uint32_t*> p0(7, nullptr), p1(7, nullptr), p2(7, nullptr);
        for (int v = 0; v < nout; ++v) {
            p0[v] = o0[v].data();
            p1[v] = o1[v].data();
            p2[v] = o2[v].data();
        }
        const int digitBits = 1 + (int)(word(6, c) % 8);
        const uint64_t block = 1 + word(7, c) % 300;
        check(oracle_stable_sort(keys.data(), vin.data(), vc, n, fb, eb, k0.data(), p0.data()) == 0,
              "stable_sort rejected a valid case", c);
        check(oracle_sort1(keys.data(), vin.data(), vc, n, fb, eb, digitBits, k1.data(),
                           p1.data()) == 0,
              "sort1 rejected a valid case", c);
        check(oracle_sort2(keys.data(), vin.data(), vc, n, fb, eb, digitBits, block,
                           1 + c % 3, k2.data(), p2.data()) == 0,
              "sort2 rejected a valid case", c);
        check(k0 == k1 && k0 == k2, "keys differ between the oracle's sorts", c);
        check(o0 == o1 && o0 == o2, "values differ between the oracle's sorts", c);

        // Per-block counts sum to the keys, offsets start each digit's run.
        const int nb = 1 + c % 8;
        const uint64_t numBlocks = (n + block - 1) / block;
        std::vector<uint32_t> counts(numBlocks << nb);
        std::vector<uint64_t> offs(numBlocks << nb);
        check(oracle_block_counts(keys.data(), n, fb + nb <= 32 ? fb : 32 - nb, nb, block,
                                  counts.data()) == 0,
              "block_counts rejected a valid case", c);
        check(oracle_digit_offsets(counts.data(), numBlocks, nb, offs.data()) == 0,
              "digit_offsets rejected a valid case", c);
        uint64_t total = 0;
        for (uint32_t x : counts) total += x;
        check(total == n, "block counts do not sum to n", c);

        // Signed and float keys: the transform and its inverse round trip.
        std::vector<uint32_t> t(n), back(n);
        for (int type = 1; type <= 2; ++type) {
            check(oracle_key_transform(keys.data(), n, type, 1, t.data()) == 0, "transform", c);
            check(oracle_key_transform(t.data(), n, type, 0, back.data()) == 0, "inverse", c);
            check(back == keys, "transform round trip", c);
        }
    }
    // Arguments the oracle must refuse.
    uint32_t k = 1, out = 0;
    check(oracle_stable_sort(&k, nullptr, 0, 1, 5, 4, &out, nullptr) != 0, "bad bit range", -1);
    check(oracle_stable_sort(&k, nullptr, 7, 1, 0, 32, &out, nullptr) != 0, "bad value count", -1);
    check(oracle_key_transform(&k, 1, 3, 1, &out) != 0, "bad key type", -1);
    if (failures) return 1;
    std::printf("oracle_check: 300 cases clean\n");
    return 0;
}
