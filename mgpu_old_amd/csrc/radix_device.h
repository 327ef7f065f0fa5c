// Device building blocks shared by the radix-sort code objects (gfx950,
// wave64): lane masks, scans, cross-XCD-coherent loads/stores and the stable
// in-wave digit ranking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgpu {

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Agent-scope loads/stores: coherent across the XCDs' L2s, which is what the
// cross-workgroup lookback needs.
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A pointer read from memory (an array's address in a segment descriptor) is
// a generic pointer to the compiler, which accesses it with flat
// instructions.  A flat access counts on the LDS counter as well as the
// memory one, so every later wait for an LDS result also waits for it (a
// load's data, a store's write), and a wait for its data also drains the LDS
// atomics in flight.  to_global gives the same address as a global-memory
// pointer, accessed with global instructions.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* to_global(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Records a failed self-check in the launch's error word, which lives in
// host memory mapped into the device (system scope: the host reads it).
__device__ __forceinline__ void set_error(uint32_t* flag, uint32_t code) {
    __hip_atomic_store(flag, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Inclusive scan over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x, uint32_t lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    return x;
}

// Exclusive scan of x over all threads of the block (threads that do not
// take part pass 0).  s_tmp holds one word per wave.  Contains barriers:
// every thread of the block must call it.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_tmp,
                                                         uint32_t lane, uint32_t wave) {
    constexpr int W = BLOCK / 64;
    uint32_t inc = wave_inclusive_scan(x, lane);
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)
        if ((uint32_t)w < wave) before += s_tmp[w];
    __syncthreads();
    return before + inc - x;
}

// Stable rank of one key per lane among the 64 lanes of a wave by digit d:
// returns the number of earlier keys of the wave with the same digit (earlier
// slots, counted in `count`, plus lower lanes of this slot) and advances
// count[d] by the number of lanes holding d.  `match` is a per-wave table of
// 64-bit lane masks, all zero on entry and left zero.  All 64 lanes must be
// active.
__device__ __forceinline__ uint32_t wave_rank(uint32_t d, uint32_t lane, uint32_t* count,
                                              uint64_t* match) {
    __hip_atomic_fetch_or(&match[d], 1ull << lane, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint64_t m =
        __hip_atomic_load(&match[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t below = lanes_below(m);
    const uint32_t old = count[d];
    if (below == 0) {  // the lowest lane of the group advances the count
        count[d] = old + (uint32_t)__builtin_popcountll(m);
        __hip_atomic_store(&match[d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return old + below;
}

}  // namespace mgpu
