// mgpu_key_map (the two-word form digit_count uses for signed and float keys)
// against mgpu_key_to_radix / mgpu_key_from_radix (the form key_transform and
// the passes use), over every 32-bit pattern whose low 12 bits are in a fixed
// set plus the special values: built and run by tests/test_key_map.py.
#include <cstdint>
#include <cstdio>

#include "kernel_args.h"

int main() {
    // (type, lo, hi) of the forward and backward forms.
    const uint32_t types[2] = {MGPU_KEYS_INT32, MGPU_KEYS_FLOAT32};
    const uint32_t fwd[2][2] = {{0x80000000u, 0x80000000u}, {0x80000000u, 0xFFFFFFFFu}};
    const uint32_t bwd[2][2] = {{0x80000000u, 0x80000000u}, {0xFFFFFFFFu, 0x80000000u}};
    const uint32_t lows[] = {0u, 1u, 0x7FFu, 0x800u, 0xFFFu, 0x123u};
    const uint32_t specials[] = {0u, 0x80000000u, 0x7F800000u, 0xFF800000u, 0x7FC00000u,
                                 0xFFC00000u, 0x00000001u, 0x80000001u, 0x7FFFFFFFu, 0xFFFFFFFFu};
    long checked = 0;
    for (int t = 0; t < 2; ++t) {
        auto check = [&](uint32_t k) {
            const uint32_t u = mgpu_key_to_radix(k, types[t]);
            if (mgpu_key_map(k, fwd[t][0], fwd[t][1]) != u) return false;
            if (mgpu_key_map(u, bwd[t][0], bwd[t][1]) != k) return false;
            if (mgpu_key_from_radix(u, types[t]) != k) return false;
            if (mgpu_key_map(k, 0u, 0u) != k) return false;
            ++checked;
            return true;
        };
        for (uint32_t hi = 0; hi < (1u << 20); ++hi)
            for (uint32_t lo : lows)
                if (!check(hi << 12 | lo)) {
                    std::printf("mismatch: type %u key %08x\n", types[t], hi << 12 | lo);
                    return 1;
                }
        for (uint32_t k : specials)
            if (!check(k)) {
                std::printf("mismatch: type %u key %08x\n", types[t], k);
                return 1;
            }
    }
    std::printf("key_map_check: %ld keys agree\n", checked);
    return 0;
}
