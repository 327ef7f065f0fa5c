// Segmented LSD radix sort for gfx950: thousands of independent arrays of
// uneven sizes sorted in one persistent launch.  Each segment's result is
// exactly what sortArray would produce on it alone (stable by the bit field,
// values following), left on side `result_side` of its ping-pong buffers.
//
// No reference counterpart: the reference sorts one array per sortArray call
// (sort/src/mgpusort/mgpusort.cpp:607-616); its per-block ranking and scatter
// steps (sort/src/kernels/sort.cu:54-479) appear here inside one workgroup.
//
// Work is split by size on the host:
//   small  (n <= 64*IPT)     one wave sorts the segment in LDS, all passes;
//   block  (n <= BLOCK*IPT)  one workgroup sorts it in LDS, all passes;
//   global (larger)          one workgroup runs every pass through global
//                            memory, tile by tile, with running digit offsets
//                            (no cross-workgroup lookback needed).
// Workgroups first drain the block/global list (largest first, from a
// shared cursor), then their waves take the small list round-robin, one
// segment per wave at a time.  The last workgroup to finish puts the cursor
// back to zero, so a captured graph of this launch can be replayed as is.
//
// In-LDS routines carry (key, payload) through the passes, one 64-bit LDS
// word per slot: the payload is the value with one value array, else the
// source index (the value in index mode; several arrays are gathered through
// it once at the end).  The global routine carries the values through each
// pass.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_args.h"
#include "radix_device.h"

namespace {

using mgpu::block_exclusive_scan;
using mgpu::st_agent;
using mgpu::wave_inclusive_scan;
using mgpu::wave_rank;


// Orders this wave's LDS accesses before and after the call.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stable rank of this lane's digit d among the active lanes of the wave
// with the same digit, counted from count[d], which it advances.  rank1: one
// LDS atomic add per lane (served in ascending lane order, as the engine
// checked at creation), or a single add when every active lane holds the
// same digit (sorted or constant keys would serialize 64 adds on one word);
// else by match masks (wave_rank).
__device__ __forceinline__ uint32_t seg_rank(uint32_t d, uint32_t lane, uint32_t* count,
                                             uint64_t* match, bool rank1) {
    if (!rank1) return wave_rank(d, lane, count, match);
    const uint32_t f = __builtin_amdgcn_readfirstlane(d);
    const uint64_t exec = __builtin_amdgcn_read_exec();
    if (__builtin_amdgcn_ballot_w64(d == f) == exec) {
        const uint32_t below = mgpu::lanes_below(exec);
        uint32_t b = 0;
        if (below == 0)
            b = __hip_atomic_fetch_add(&count[f], (uint32_t)__builtin_popcountll(exec),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return __builtin_amdgcn_readfirstlane(b) + below;
    }
    return __hip_atomic_fetch_add(&count[d], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Bit field of pass p: shift and mask (the split of mgpuSegArgs).
__device__ __forceinline__ void pass_field(const mgpuSegArgs& a, uint32_t p, uint32_t& shift,
                                           uint32_t& mask) {
    const uint32_t np = a.num_passes, base = a.num_bits / np, extra = a.num_bits % np;
    const uint32_t wide_from = np - extra;  // passes from here take base + 1 bits
    shift = a.first_bit + p * base + (p > wide_from ? p - wide_from : 0u);
    mask = (1u << (base + (p >= wide_from ? 1u : 0u))) - 1u;
}

// Shared memory of one workgroup of BLOCK threads (W waves).  The in-LDS
// routines stage a slot's key and payload together, one 64-bit word; the
// global routine uses the stage's first TILE 32-bit words.  Wave w's slice
// (64 * IPT words from w * 64 * IPT) also holds, in its first RADIX words,
// the wave's match-mask table of the match-mask ranking (match_table), which
// the staging overwrites: that ranking clears it first (clear_match).
template <int RB, int IPT, int BLOCK>
struct SegShared {
    static constexpr int W = BLOCK / 64;
    static constexpr int RADIX = 1 << RB;
    static constexpr int TILE = BLOCK * IPT;
    static constexpr int MAXP = (32 + RB - 1) / RB;
    static_assert(64 * IPT >= RADIX, "a wave's stage slice holds its match-mask table");
    uint64_t stage[TILE];
    uint32_t count[W][RADIX];
    uint32_t base[MAXP][RADIX];  // global routine: per-pass digit starts
    uint32_t run[RADIX];         // global routine: running digit offsets
    uint32_t dst[RADIX];         // global routine: output base per digit for a tile
    uint32_t tmp[W];
    uint32_t item;
};

template <int RB, int IPT, int BLOCK>
__device__ __forceinline__ uint64_t* match_table(SegShared<RB, IPT, BLOCK>& sh, uint32_t wave) {
    return sh.stage + wave * 64 * IPT;
}

// Zeroes the wave's match-mask table before a ranking by match masks (the
// atomic-add ranking does not use it).  Orders the wave's LDS accesses.
template <int RB, int IPT, int BLOCK>
__device__ __forceinline__ void clear_match(const mgpuSegArgs& a, SegShared<RB, IPT, BLOCK>& sh,
                                            uint32_t lane, uint32_t wave) {
    if (a.rank1) return;
    uint64_t* m = match_table(sh, wave);
    for (uint32_t i = lane; i < (uint32_t)SegShared<RB, IPT, BLOCK>::RADIX; i += 64) m[i] = 0;
    wave_sync();
}

// The payload a slot carries through the in-LDS passes: its value when the
// sort has exactly one value array (read once, in order, at the start, and
// written at the end with no gather), else its source index (the value
// itself in index mode; the gather index of every array otherwise).
__device__ __forceinline__ bool carries_value(const mgpuSegArgs& a) {
    return !a.index_mode && a.num_values == 1;
}

// Writes the sorted keys and the values: the carried payload when it is the
// value (one array, or indices), else every value array gathered through the
// source indices, out[pos] = in[src[pos]].  `sync` orders the gathers of the
// whole group before any store (the result side may be the input side).
template <int IPT, typename Sync>
__device__ __forceinline__ void emit(const mgpuSegArgs& a, const mgpuSegDesc& s, uint32_t n,
                                     const uint32_t* key, const uint32_t* pay, const uint32_t* idx,
                                     Sync sync) {
    const uint32_t side = a.result_side;
    const auto ko = mgpu::to_global(s.keys[side]);
#pragma unroll
    for (int j = 0; j < IPT; ++j)
        if (idx[j] < n) ko[idx[j]] = key[j];
    if (a.index_mode || carries_value(a)) {
        const auto vo = mgpu::to_global(s.vals[0][side]);
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) vo[idx[j]] = pay[j];
        return;
    }
    for (uint32_t v = 0; v < a.num_values; ++v) {
        const auto vi = mgpu::to_global(s.vals[v][0]);
        const auto vo = mgpu::to_global(s.vals[v][side]);
        uint32_t val[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) val[j] = vi[pay[j]];
        sync();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) vo[idx[j]] = val[j];
        sync();
    }
}

// Initial registers of an in-LDS sort: slot j's key and payload.
template <int IPT>
__device__ __forceinline__ void load_slots(const mgpuSegArgs& a, const mgpuSegDesc& s, uint32_t n,
                                           const uint32_t* idx, uint32_t* key, uint32_t* pay) {
    const auto ki = mgpu::to_global(s.keys[0]);
    const bool value = carries_value(a);
    const auto vi = mgpu::to_global(s.vals[0][0]);
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        key[j] = idx[j] < n ? ki[idx[j]] : 0u;
        pay[j] = idx[j] < n && value ? vi[idx[j]] : idx[j];
    }
}

// ---- one wave, whole segment in registers + this wave's LDS slice ---------
template <int RB, int IPT, int BLOCK>
__device__ void wave_sort(const mgpuSegArgs& a, const mgpuSegDesc& s, SegShared<RB, IPT, BLOCK>& sh,
                          uint32_t lane, uint32_t wave) {
    constexpr int RADIX = SegShared<RB, IPT, BLOCK>::RADIX;
    constexpr int DPL = RADIX >= 64 ? RADIX / 64 : 1;  // digits per lane in the scan
    const uint32_t n = s.n;
    uint64_t* stage = sh.stage + wave * 64 * IPT;
    uint32_t* count = sh.count[wave];
    uint64_t* match = match_table(sh, wave);

    uint32_t key[IPT], pay[IPT], idx[IPT], pos[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) idx[j] = j * 64 + lane;
    load_slots<IPT>(a, s, n, idx, key, pay);
    for (uint32_t p = 0; p < a.num_passes; ++p) {
        uint32_t shift, mask;
        pass_field(a, p, shift, mask);
#pragma unroll
        for (int i = 0; i < DPL; ++i)
            if (lane * DPL + i < (uint32_t)RADIX) count[lane * DPL + i] = 0;
        wave_sync();
        clear_match(a, sh, lane, wave);
        uint32_t rank[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n)
                rank[j] = seg_rank((key[j] >> shift) & mask, lane, count, match, a.rank1 != 0);
        wave_sync();
        // Exclusive scan of the digit counts across the wave.
        uint32_t c[DPL], sum = 0;
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            c[i] = lane * DPL + i < (uint32_t)RADIX ? count[lane * DPL + i] : 0u;
            sum += c[i];
        }
        uint32_t run = wave_inclusive_scan(sum, lane) - sum;
#pragma unroll
        for (int i = 0; i < DPL; ++i)
            if (lane * DPL + i < (uint32_t)RADIX) {
                count[lane * DPL + i] = run;
                run += c[i];
            }
        wave_sync();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) {
                pos[j] = count[(key[j] >> shift) & mask] + rank[j];
                stage[pos[j]] = (uint64_t)pay[j] << 32 | key[j];
            }
        wave_sync();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) {
                const uint64_t t = stage[idx[j]];
                key[j] = (uint32_t)t;
                pay[j] = (uint32_t)(t >> 32);
            }
        wave_sync();
    }
    emit<IPT>(a, s, n, key, pay, idx, [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        wave_sync();
    });
}

// ---- one wave, at most 64 keys: a bitonic network in registers -------------
// Lane i holds key i.  The network orders 64-bit compare keys (the sort's bit
// field above the lane index, all ones past n), so equal fields keep their
// input order and the result is the stable sort by the field, as the passes
// would leave it; each lane then takes its key and payload from the lane its
// compare key came from.  No LDS and no atomics: two shuffles per step, 21
// steps.
__device__ void wave_bitonic64(const mgpuSegArgs& a, const mgpuSegDesc& s, uint32_t lane) {
    const uint32_t n = s.n;
    uint32_t idx[1] = {lane}, key[1], pay[1];
    load_slots<1>(a, s, n, idx, key, pay);
    const uint32_t field = a.num_bits >= 32 ? key[0] >> a.first_bit
                                            : (key[0] >> a.first_bit) & ((1u << a.num_bits) - 1u);
    uint64_t ck = lane < n ? ((uint64_t)field << 6) | lane : ~0ull;
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t olo = (uint32_t)__shfl_xor((int)(uint32_t)ck, (int)j, 64);
            const uint32_t ohi = (uint32_t)__shfl_xor((int)(uint32_t)(ck >> 32), (int)j, 64);
            const uint64_t other = ((uint64_t)ohi << 32) | olo;
            // Ascending blocks of k (all of them at k = 64); the lower lane
            // of a pair keeps the smaller key in an ascending block.
            const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
            if (keep_min ? other < ck : other > ck) ck = other;
        }
    }
    const int src = (int)((uint32_t)ck & 63u);
    uint32_t k2[1] = {(uint32_t)__shfl((int)key[0], src, 64)};
    uint32_t p2[1] = {(uint32_t)__shfl((int)pay[0], src, 64)};
    emit<1>(a, s, n, k2, p2, idx, [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        wave_sync();
    });
}

// ---- one workgroup, whole segment in registers + LDS ----------------------
template <int RB, int IPT, int BLOCK>
__device__ void block_sort(const mgpuSegArgs& a, const mgpuSegDesc& s, SegShared<RB, IPT, BLOCK>& sh,
                           uint32_t tid, uint32_t lane, uint32_t wave) {
    constexpr int RADIX = SegShared<RB, IPT, BLOCK>::RADIX;
    constexpr int W = BLOCK / 64;
    const uint32_t n = s.n;
    uint32_t key[IPT], pay[IPT], idx[IPT], pos[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) idx[j] = wave * 64 * IPT + j * 64 + lane;
    load_slots<IPT>(a, s, n, idx, key, pay);
    for (uint32_t p = 0; p < a.num_passes; ++p) {
        uint32_t shift, mask;
        pass_field(a, p, shift, mask);
        // (The previous pass's reads of the stage end before this barrier,
        // and its writes of the next pass come after two more.)
        for (uint32_t i = tid; i < (uint32_t)(W * RADIX); i += BLOCK) (&sh.count[0][0])[i] = 0;
        __syncthreads();
        clear_match(a, sh, lane, wave);
        uint32_t rank[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n)
                rank[j] = seg_rank((key[j] >> shift) & mask, lane, sh.count[wave],
                                   match_table(sh, wave), a.rank1 != 0);
        __syncthreads();
        uint32_t total = 0;
        if (tid < (uint32_t)RADIX) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                uint32_t c = sh.count[w][tid];
                sh.count[w][tid] = total;
                total += c;
            }
        }
        const uint32_t start = block_exclusive_scan<BLOCK>(total, sh.tmp, lane, wave);
        if (tid < (uint32_t)RADIX) {
#pragma unroll
            for (int w = 0; w < W; ++w) sh.count[w][tid] += start;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) {
                pos[j] = sh.count[wave][(key[j] >> shift) & mask] + rank[j];
                sh.stage[pos[j]] = (uint64_t)pay[j] << 32 | key[j];
            }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (idx[j] < n) {
                const uint64_t t = sh.stage[idx[j]];
                key[j] = (uint32_t)t;
                pay[j] = (uint32_t)(t >> 32);
            }
    }
    emit<IPT>(a, s, n, key, pay, idx, [] { __syncthreads(); });
}

// ---- one workgroup, passes through global memory ---------------------------
template <int RB, int IPT, int BLOCK>
__device__ void global_sort(const mgpuSegArgs& a, const mgpuSegDesc& s, SegShared<RB, IPT, BLOCK>& sh,
                            uint32_t tid, uint32_t lane, uint32_t wave) {
    constexpr int RADIX = SegShared<RB, IPT, BLOCK>::RADIX;
    uint32_t* const st = reinterpret_cast<uint32_t*>(sh.stage);  // TILE words of it
    constexpr int W = BLOCK / 64;
    constexpr uint32_t TILE = SegShared<RB, IPT, BLOCK>::TILE;
    const uint32_t n = s.n;
    const uint32_t np = a.num_passes;

    // Digit histograms of every pass in one read, then their exclusive scans.
    for (uint32_t i = tid; i < np * RADIX; i += BLOCK) (&sh.base[0][0])[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += BLOCK) {
        const uint32_t k = mgpu::to_global(s.keys[0])[i];
        for (uint32_t p = 0; p < np; ++p) {
            uint32_t shift, mask;
            pass_field(a, p, shift, mask);
            __hip_atomic_fetch_add(&sh.base[p][(k >> shift) & mask], 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
    for (uint32_t p = 0; p < np; ++p) {
        uint32_t c = tid < (uint32_t)RADIX ? sh.base[p][tid] : 0u;
        uint32_t e = block_exclusive_scan<BLOCK>(c, sh.tmp, lane, wave);
        if (tid < (uint32_t)RADIX) sh.base[p][tid] = e;
    }

    const uint32_t nv = a.num_values;
    for (uint32_t p = 0; p < np; ++p) {
        const uint32_t in = p & 1, out = in ^ 1;
        uint32_t shift, mask;
        pass_field(a, p, shift, mask);
        const auto ki = mgpu::to_global(s.keys[in]);
        const auto ko = mgpu::to_global(s.keys[out]);
        if (tid < (uint32_t)RADIX) sh.run[tid] = sh.base[p][tid];
        for (uint32_t tb = 0; tb < n; tb += TILE) {
            for (uint32_t i = tid; i < (uint32_t)(W * RADIX); i += BLOCK) (&sh.count[0][0])[i] = 0;
            __syncthreads();
            clear_match(a, sh, lane, wave);
            uint32_t key[IPT], rank[IPT], pos[IPT], dest[IPT];
            const uint32_t wb = tb + wave * 64 * IPT;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t i = wb + j * 64 + lane;
                if (i < n) {
                    key[j] = ki[i];
                    rank[j] = seg_rank((key[j] >> shift) & mask, lane, sh.count[wave],
                                       match_table(sh, wave), a.rank1 != 0);
                }
            }
            __syncthreads();
            uint32_t total = 0;
            if (tid < (uint32_t)RADIX) {
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint32_t c = sh.count[w][tid];
                    sh.count[w][tid] = total;
                    total += c;
                }
            }
            const uint32_t start = block_exclusive_scan<BLOCK>(total, sh.tmp, lane, wave);
            if (tid < (uint32_t)RADIX) {
#pragma unroll
                for (int w = 0; w < W; ++w) sh.count[w][tid] += start;
                sh.dst[tid] = sh.run[tid] - start;
                sh.run[tid] += total;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t i = wb + j * 64 + lane;
                if (i < n) {
                    pos[j] = sh.count[wave][(key[j] >> shift) & mask] + rank[j];
                    st[pos[j]] = key[j];
                }
            }
            __syncthreads();
            const uint32_t valid = min(TILE, n - tb);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t q = j * BLOCK + tid;
                dest[j] = n;
                if (q < valid) {
                    const uint32_t k = st[q];
                    dest[j] = sh.dst[(k >> shift) & mask] + q;
                    if (dest[j] < n) ko[dest[j]] = k;
                    else mgpu::set_error(a.error_flag, 4u);
                }
            }
            for (uint32_t v = 0; v < nv; ++v) {
                const bool gen = a.index_mode && p == 0;
                const auto vi = mgpu::to_global(s.vals[v][in]);
                uint32_t val[IPT];
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const uint32_t i = wb + j * 64 + lane;
                    if (i < n) val[j] = gen ? i : vi[i];
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const uint32_t i = wb + j * 64 + lane;
                    if (i < n) st[pos[j]] = val[j];
                }
                __syncthreads();
                const auto vo = mgpu::to_global(s.vals[v][out]);
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const uint32_t q = j * BLOCK + tid;
                    if (q < valid && dest[j] < n) vo[dest[j]] = st[q];
                }
            }
            __syncthreads();
        }
        // This pass's stores must be seen by the next pass's loads of the
        // same workgroup (the L1 may hold lines read two passes ago).
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }

    // The result is on side (np & 1); move it when the launch wants the other.
    if ((np & 1) != a.result_side) {
        const uint32_t from = np & 1, to = a.result_side;
        const auto kt = mgpu::to_global(s.keys[to]);
        const auto kf = mgpu::to_global(s.keys[from]);
        for (uint32_t i = tid; i < n; i += BLOCK) kt[i] = kf[i];
        for (uint32_t v = 0; v < nv; ++v)
            for (uint32_t i = tid; i < n; i += BLOCK)
                mgpu::to_global(s.vals[v][to])[i] = mgpu::to_global(s.vals[v][from])[i];
    }
}

// GLOBAL: the kernel carries the one-workgroup routine for segments larger
// than a workgroup's LDS tile.  That routine sets the kernel's register
// count (r8, 16 keys per thread, 512 threads: 221 VGPRs and 96 bytes of
// scratch with it, 141 and none without; 8 keys per thread: 132 -> 89), so
// launches whose block-list segments all fit the tile take the kernel
// without it, which fits more waves per SIMD.
template <int RB, int IPT, int BLOCK, bool GLOBAL>
__device__ void segsort_body(const mgpuSegArgs& a) {
    static_assert(BLOCK >= (1 << RB), "a thread per digit in the block scans");
    constexpr int W = BLOCK / 64;
    __shared__ SegShared<RB, IPT, BLOCK> sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // Workgroup phase: block and global segments, largest first, from a
    // shared cursor.  (Dealt round-robin like the small ones below, they
    // measured slower where they are many: 20000 segments log-uniform in
    // [2^6, 2^12] 0.44 -> 0.68 ms, since a workgroup that starts late still
    // holds a full share; an atomic per segment of this size is cheap.)
    while (true) {
        if (tid == 0)
            sh.item = __hip_atomic_fetch_add(&a.counters[0], 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t item = sh.item;
        __syncthreads();
        if (item >= a.num_big) break;
        const mgpuSegDesc& s = a.segs[a.big_list[item]];
        if (s.n <= (uint32_t)SegShared<RB, IPT, BLOCK>::TILE) {
            block_sort<RB, IPT, BLOCK>(a, s, sh, tid, lane, wave);
        } else if constexpr (GLOBAL) {
            global_sort<RB, IPT, BLOCK>(a, s, sh, tid, lane, wave);
        } else if (tid == 0) {
            mgpu::set_error(a.error_flag, 7u);  // the host never sends one here
        }
        __syncthreads();
    }

    // Wave phase: one small segment per wave at a time, dealt round-robin
    // over the grid's waves.  (A shared cursor costs a device-wide atomic on
    // one word per segment: ~25 ns each, serialized, which bounded a batch
    // of 130K segments of up to 64 keys at 3.3 ms whatever the kernel.)
    for (uint32_t item = blockIdx.x * W + wave; item < a.num_small; item += gridDim.x * W) {
        const mgpuSegDesc& s = a.segs[a.small_list[item]];
        if (s.n <= 64) wave_bitonic64(a, s, lane);
        else wave_sort<RB, IPT, BLOCK>(a, s, sh, lane, wave);
    }

    __syncthreads();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t done = __hip_atomic_fetch_add(&a.counters[2], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        if (done == gridDim.x - 1) {  // every workgroup is past both queues
            st_agent(&a.counters[0], 0u);
            st_agent(&a.counters[2], 0u);
        }
    }
}

}  // namespace

// Kernel variants: digit width RB (4..8 bits) x keys per thread IPT (8, 16)
// x workgroup size BLOCK (128 for digits up to 7 bits, 256, 512), each with
// the one-workgroup global-memory routine and without it (_ng).
#define MGPU_SEGSORT(RB, IPT, B)                                                                \
    extern "C" __global__ void __launch_bounds__(B)                                            \
        mgpu_segsort_r##RB##_i##IPT##_b##B(mgpuSegArgs a) {                                    \
        segsort_body<RB, IPT, B, true>(a);                                                     \
    }                                                                                          \
    extern "C" __global__ void __launch_bounds__(B)                                            \
        mgpu_segsort_r##RB##_i##IPT##_b##B##_ng(mgpuSegArgs a) {                               \
        segsort_body<RB, IPT, B, false>(a);                                                    \
    }
#define MGPU_SEGSORT_BLOCKS(RB, IPT)                                                           \
    MGPU_SEGSORT(RB, IPT, 256)                                                                 \
    MGPU_SEGSORT(RB, IPT, 512)

MGPU_SEGSORT(4, 8, 128)
MGPU_SEGSORT(4, 16, 128)
MGPU_SEGSORT(5, 8, 128)
MGPU_SEGSORT(5, 16, 128)
MGPU_SEGSORT(6, 8, 128)
MGPU_SEGSORT(6, 16, 128)
MGPU_SEGSORT(7, 8, 128)
MGPU_SEGSORT(7, 16, 128)
MGPU_SEGSORT_BLOCKS(4, 8)
MGPU_SEGSORT_BLOCKS(4, 16)
MGPU_SEGSORT_BLOCKS(5, 8)
MGPU_SEGSORT_BLOCKS(5, 16)
MGPU_SEGSORT_BLOCKS(6, 8)
MGPU_SEGSORT_BLOCKS(6, 16)
MGPU_SEGSORT_BLOCKS(7, 8)
MGPU_SEGSORT_BLOCKS(7, 16)
MGPU_SEGSORT_BLOCKS(8, 8)
MGPU_SEGSORT_BLOCKS(8, 16)
