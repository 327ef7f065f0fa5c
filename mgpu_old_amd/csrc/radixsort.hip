// LSD radix sort pass kernels for gfx950 (CDNA4, wave64).
//
// One sort = digit_count (one read of the keys: the digit histogram of every
// pass) -> digit_scan (exclusive scan of those histograms) -> one
// rank_scatter launch per digit pass.
//
// Reference kernels this replaces (seanbaxter/mgpu_old, sort/src/kernels):
//   CountBuckets_n            count.cu:48-333      -> mgpu_digit_count,
//                                                  mgpu_digit_count_cols
//   HistogramReduce_{1,2,3}_n hist1.cu, hist2.cu,  -> mgpu_digit_scan plus the
//                             hist3.cu             decoupled lookback inside
//                                                  rank_scatter
//   RadixSort_n (4 modes)     sort.cu:54-479       -> mgpu_rank_scatter_*
// The reference counts per sort block and scans (block, digit) counts in three
// launches every pass; here every pass's global digit totals come from one
// read up front, and the per-tile offsets come from a lookback across tiles
// inside the scatter kernel, so a pass reads its keys and values once and
// writes them once.
//
// Stable ranking inside a tile (wave64): keys are loaded striped, lane l of
// wave w holding key  tile_base + w*64*IPT + j*64 + l  in slot j, so the
// order (wave, slot, lane) is the input order.  For each slot the wave ranks
// its 64 keys by digit: a key's rank is the number of lower lanes of the slot
// with the same digit plus the wave's running count of that digit.  RANK 0
// builds a match mask of those lanes with LDS atomic ORs into a per-wave
// table of 256 lane masks; RANK 1 takes the rank straight from one LDS atomic
// add per key on the wave's count, where the device serves same-word lanes in
// ascending order (checked by the engine at creation).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_args.h"
#include "radix_device.h"

namespace {

constexpr uint32_t RADIX = MGPU_RADIX;

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Records a failed self-check in the launch's error word, which lives in
// host memory mapped into the device (system scope: the host reads it).
__device__ __forceinline__ void set_error(uint32_t* flag, uint32_t code) {
    __hip_atomic_store(flag, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Lane i gets v of lane i - 1, lane 0 gets lane0 (DPP wave_shr:1).
__device__ __forceinline__ uint32_t prev_lane(uint32_t v, uint32_t lane0) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)lane0, (int)v, 0x138, 0xF, 0xF, false);
}

// Lane i gets v of lane i + 1, lane 63 gets lane63 (DPP wave_shl:1).
__device__ __forceinline__ uint32_t next_lane(uint32_t v, uint32_t lane63) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)lane63, (int)v, 0x130, 0xF, 0xF, false);
}

// Orders this wave's LDS accesses before and after the call.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// Maximum over the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off, 64);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off, 64);
        const uint64_t o = ((uint64_t)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

// Exclusive scan of x over the first 256 threads of the block (threads past
// 256 must pass 0).  s_tmp holds one word per wave.  Contains barriers: every
// thread of the block must call it.
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

// Exclusive scan of x within each group of RADIX consecutive threads (RADIX /
// 64 waves).  s_tmp holds one word per wave.  Contains barriers: every thread
// of the block must call it.
template <int BLOCK>
__device__ __forceinline__ uint32_t group_exclusive_scan(uint32_t x, uint32_t* s_tmp,
                                                         uint32_t lane, uint32_t wave) {
    constexpr uint32_t GW = RADIX / 64;  // waves per group
    const uint32_t inc = wave_inclusive_scan(x, lane);
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    const uint32_t g0 = wave & ~(GW - 1u);
#pragma unroll
    for (uint32_t w = 0; w < GW; ++w)
        if (g0 + w < wave) before += s_tmp[g0 + w];
    __syncthreads();
    return before + inc - x;
}

// Adds 4 keys per lane (16-byte loads) to a per-wave LDS histogram of one
// pass's digit.  Sorted and constant input put every lane on one counter, and
// an LDS atomic serializes the lanes that hit the same word (2^28 sorted keys:
// 2.6 ms instead of 0.3): when every active lane's 4 keys share one digit the
// wave adds them with a single atomic.  (Merging equal digits within a lane
// as well helps skewed input but costs uniform input more than it saves.)
__device__ __forceinline__ void count4(uint32_t* h, uint32_t k0, uint32_t k1, uint32_t k2,
                                       uint32_t k3, uint32_t shift, uint32_t mask) {
    const uint32_t d0 = (k0 >> shift) & mask, d1 = (k1 >> shift) & mask;
    const uint32_t d2 = (k2 >> shift) & mask, d3 = (k3 >> shift) & mask;
    const uint32_t f = __builtin_amdgcn_readfirstlane(d0);
    const uint64_t exec = __builtin_amdgcn_read_exec();
    const bool one = d0 == f && d1 == f && d2 == f && d3 == f;
    if (__builtin_amdgcn_ballot_w64(one) == exec) {
        if (lanes_below(exec) == 0)
            __hip_atomic_fetch_add(&h[f], 4u * (uint32_t)__builtin_popcountll(exec),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
    }
    __hip_atomic_fetch_add(&h[d0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&h[d1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&h[d2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&h[d3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Adds 1 per active lane to LDS counter h[i].  The lanes of one LDS atomic
// that hit the same word are serialized, so when some lanes hold `pref` (the
// caller's guess at the most frequent value) the lanes with the first such
// lane's i are merged first: that lane adds the group's size, the others add
// 0 to a word of their own in spare[64], every other lane adds 1.  Otherwise
// (the common case on spread-out input) each lane adds 1.
__device__ __forceinline__ void count1(uint32_t* h, uint32_t* spare, uint32_t i, bool pref,
                                       uint32_t lane) {
    const uint64_t pm = __builtin_amdgcn_ballot_w64(pref);
    if (pm == 0) {
        __hip_atomic_fetch_add(&h[i], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
    }
    const uint32_t leader = (uint32_t)__builtin_ctzll(pm);
    const uint32_t f = __builtin_amdgcn_readlane(i, leader);
    const uint64_t m = __builtin_amdgcn_ballot_w64(i == f);
    const bool peer = i == f && lane != leader;
    uint32_t* const w = peer ? &spare[lane] : &h[i];
    const uint32_t v = peer ? 0u : (lane == leader ? (uint32_t)__builtin_popcountll(m) : 1u);
    __hip_atomic_fetch_add(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------------------
// digit_count: LDS histograms of all passes' digits, merged into the global
// histogram with one atomic per (pass, digit) per workgroup.  Also zeroes the
// lookback words the passes of this sort will use.
//
// COLS = 0: one histogram per wave ([wave][pass][digit]); a wave adds 4 keys
// per lane with one atomic when all of them share a digit (count4).
// COLS = 32: one histogram per workgroup with a column per lane slot
// ([pass][digit][lane % 32]): the 32 lanes of a half-wave always hit different
// LDS banks, and lanes with equal digits different words, so an LDS atomic
// never serializes more than the two lanes of a wave that share a column (an
// atomic add per key costs ~11 CU cycles per wave on random banks, ~5.6 on
// distinct ones, ~54 when a third of the lanes share one word).  128 KB of
// LDS: 1024 threads, one workgroup per CU.
template <int BLOCK, int COLS, bool KMAP>
__device__ void digit_count_body(const mgpuCountArgs& a) {
    constexpr int W = BLOCK / 64;
    constexpr uint32_t WORDS = COLS ? MGPU_MAX_PASSES * RADIX * COLS : W * MGPU_MAX_PASSES * RADIX;
    static_assert(COLS == 0 || (COLS == 32 && BLOCK == MGPU_MAX_PASSES * RADIX),
                  "the column layout merges one (pass, digit) per thread");
    __shared__ uint32_t s_hist[WORDS];
    __shared__ uint32_t s_desc[2];  // early-exit descents of this workgroup
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    for (uint32_t i = tid; i < WORDS; i += BLOCK) s_hist[i] = 0;
    if (tid < 2) s_desc[tid] = 0;

    // Zero this sort's lookback words (grid-stride, 16-byte stores).
    {
        uint4* c = reinterpret_cast<uint4*>(a.clear);
        const uint64_t nvec = a.clear_words >> 2;
        for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + tid; i < nvec;
             i += (uint64_t)gridDim.x * BLOCK)
            c[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();

    // Word of (pass, digit) for this thread: pass p's histogram starts at
    // p * PSTRIDE, digits are DSTRIDE apart, and the thread's own histogram
    // (COLS = 0) or column (COLS = 32) is at `mine`.
    constexpr uint32_t PSTRIDE = COLS ? RADIX * COLS : RADIX;
    constexpr uint32_t DSTRIDE = COLS ? COLS : 1;
    uint32_t* const mine = COLS ? s_hist + (tid & (COLS - 1)) : s_hist + wave * MGPU_MAX_PASSES * RADIX;
    const uint32_t np = a.num_passes;
    auto add = [&](uint32_t p, uint32_t d) {
        __hip_atomic_fetch_add(&mine[p * PSTRIDE + d * DSTRIDE], 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto count_key = [&](uint32_t k) {
#pragma unroll
        for (int p = 0; p < MGPU_MAX_PASSES; ++p)
            if ((uint32_t)p < np) add(p, (k >> a.shift[p]) & a.mask[p]);
    };

    // Chunk of this workgroup inside its segment; the segment's histograms.
    const uint32_t seg = blockIdx.x / a.chunks_per_seg;
    const uint32_t seg_begin =
        seg == 0 ? 0u : mgpu_seg_first_tile(seg, a.num_tiles, a.num_segs) * a.tile;
    const uint32_t seg_end =
        seg + 1 >= a.num_segs ? a.n : mgpu_seg_first_tile(seg + 1, a.num_tiles, a.num_segs) * a.tile;
    const uint32_t begin = min(seg_end, seg_begin + (blockIdx.x % a.chunks_per_seg) * a.chunk);
    const uint32_t end = min(seg_end, begin + a.chunk);
    uint32_t* const hist = a.hist + seg * (MGPU_MAX_PASSES * RADIX);
    // Early-exit check: descents between neighbours (each key against the one
    // after it, across chunk ends too) of the sort's bit field and of the
    // first pass's digit (what the reference's per-pass count checks: the
    // whole key and the current digit).  In the vector body a wave reads 64
    // consecutive vectors per step, so the key after a lane's vector is the
    // next lane's first (DPP), or for lane 63 the first of lane 0's vector
    // of the next step (readlane), or after the wave's last step the next
    // wave's first key, which lane 63 loads with the round's vectors.  No
    // lane waits on memory inside the loop.  The pair that joins the vector
    // body to the keys after it is taken once, after the loop.
    // Fields are compared in place: x & M > y & M orders two keys as their
    // extracted fields would (both sit at the same bit position).
    const bool check = a.unsorted != nullptr;
    // Signed or float keys (KMAP) are counted and checked in their unsigned
    // order.
    auto key_at = [&](uint32_t i) {
        return KMAP ? mgpu_key_map(a.keys[i], a.map_lo, a.map_hi) : a.keys[i];
    };
    const uint32_t mfull = a.check_mask << a.check_shift, mfirst = a.mask[0] << a.shift[0];
    uint32_t descents = 0, first_desc = 0;  // per lane (scalar tails)
    uint32_t wdesc = 0, wfirst = 0;         // per wave, by ballot (vector body)
    auto pair = [&](uint32_t x, uint32_t y) {  // x just before y in the input
        descents += (x & mfull) > (y & mfull);
        first_desc += (x & mfirst) > (y & mfirst);
    };
    auto ballots = [](bool p) { return (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(p)); };
    if (begin < end) {
        if (a.vec_ok) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            const u32x4* v = reinterpret_cast<const u32x4*>(a.keys + begin);
            const uint32_t nvec = (end - begin) >> 2;
            // UNROLL independent 16-byte loads in flight per thread before any
            // of them is used: the loop is bound by memory latency otherwise
            // (the one-workgroup-per-CU column kernel: 8, 0.23 -> 0.21 ms at
            // 2^28; 16 spills).
            constexpr int UNROLL = COLS ? 8 : 4;
            const uint32_t lane = tid & 63;
            // Vectors of a round: wave w's step u covers the 64 vectors from
            // round + (w * UNROLL + u) * 64.  Every lane runs the same rounds
            // (lanes past the chunk's vectors are masked off), so that the
            // descent counts can be taken by wave-wide ballots.
            const uint32_t wbase = wave * 64 * UNROLL + lane;
            for (uint32_t round = 0; round < nvec; round += UNROLL * BLOCK) {
                const uint32_t i0 = round + wbase;
                u32x4 q[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    const uint32_t i = i0 + u * 64;
                    q[u] = i < nvec ? __builtin_nontemporal_load(v + i) : u32x4{0u, 0u, 0u, 0u};
                }
                const uint32_t ilast = i0 + (UNROLL - 1) * 64;
                uint32_t edge = 0;
                if (check && lane == 63 && ilast + 1 < nvec) edge = key_at(begin + 4 * ilast + 4);
                if constexpr (KMAP) {
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) {
                        q[u].x = mgpu_key_map(q[u].x, a.map_lo, a.map_hi);
                        q[u].y = mgpu_key_map(q[u].y, a.map_lo, a.map_hi);
                        q[u].z = mgpu_key_map(q[u].z, a.map_lo, a.map_hi);
                        q[u].w = mgpu_key_map(q[u].w, a.map_lo, a.map_hi);
                    }
                }
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    const uint32_t i = i0 + u * 64;
                    const bool valid = i < nvec;
                    if (valid) {
#pragma unroll
                        for (int p = 0; p < MGPU_MAX_PASSES; ++p) {
                            if ((uint32_t)p >= np) continue;
                            if constexpr (COLS == 0) {
                                count4(mine + p * PSTRIDE, q[u].x, q[u].y, q[u].z, q[u].w,
                                       a.shift[p], a.mask[p]);
                            } else {
                                // The digit by one bit-field extract (mask =
                                // 2^width - 1).
                                const uint32_t w = (uint32_t)__builtin_popcount(a.mask[p]);
                                add(p, __builtin_amdgcn_ubfe(q[u].x, a.shift[p], w));
                                add(p, __builtin_amdgcn_ubfe(q[u].y, a.shift[p], w));
                                add(p, __builtin_amdgcn_ubfe(q[u].z, a.shift[p], w));
                                add(p, __builtin_amdgcn_ubfe(q[u].w, a.shift[p], w));
                            }
                        }
                    }
                    if (check) {
                        // The key after this vector (used only when vector
                        // i + 1 is one of the chunk's: then it is in a lane).
                        const uint32_t after63 =
                            u + 1 < UNROLL ? __builtin_amdgcn_readlane(q[u + 1 < UNROLL ? u + 1 : u].x, 0)
                                           : edge;
                        const uint32_t nx = next_lane(q[u].x, after63);
                        const bool has = i + 1 < nvec;
                        const u32x4 f = q[u] & mfull, g = q[u] & mfirst;
                        wdesc += ballots(valid && f.x > f.y) + ballots(valid && f.y > f.z) +
                                 ballots(valid && f.z > f.w) + ballots(has && f.w > (nx & mfull));
                        wfirst += ballots(valid && g.x > g.y) + ballots(valid && g.y > g.z) +
                                  ballots(valid && g.z > g.w) + ballots(has && g.w > (nx & mfirst));
                    }
                }
            }
            if (check && tid == 0 && nvec > 0 && begin + 4 * nvec < a.n)
                pair(key_at(begin + 4 * nvec - 1), key_at(begin + 4 * nvec));
            for (uint32_t i = begin + (nvec << 2) + tid; i < end; i += BLOCK) {
                const uint32_t k = key_at(i);
                count_key(k);
                if (check && i + 1 < a.n) pair(k, key_at(i + 1));
            }
        } else {
            for (uint32_t i = begin + tid; i < end; i += BLOCK) {
                const uint32_t k = key_at(i);
                count_key(k);
                if (check && i + 1 < a.n) pair(k, key_at(i + 1));
            }
        }
    }
    if (check) {
        // Sums over the wave, then over the workgroup in LDS: one global
        // atomic per counter and workgroup.
        auto wave_sum = [](uint32_t v) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            return v;
        };
        const uint32_t dsum = wave_sum(descents) + wdesc, fsum = wave_sum(first_desc) + wfirst;
        if ((tid & 63) == 0) {
            if (dsum) __hip_atomic_fetch_add(&s_desc[0], dsum, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
            if (fsum) __hip_atomic_fetch_add(&s_desc[1], fsum, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
    if (check && tid < 2 && s_desc[tid])
        __hip_atomic_fetch_add(a.unsorted + tid, s_desc[tid], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);

    if constexpr (COLS == 0) {
        for (uint32_t i = tid; i < (uint32_t)(MGPU_MAX_PASSES * RADIX); i += BLOCK) {
            if (i / RADIX >= np) break;
            uint32_t sum = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) sum += s_hist[w * MGPU_MAX_PASSES * RADIX + i];
            if (sum) __hip_atomic_fetch_add(&hist[i], sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else {
        // Thread i sums the columns of (pass, digit) i, starting at column i
        // so that the lanes of a wave read different banks.
        const uint32_t i = tid;
        if (i / RADIX < np) {
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t c = 0; c < (uint32_t)COLS; ++c)
                sum += s_hist[i * COLS + ((c + i) & (COLS - 1))];
            if (sum) __hip_atomic_fetch_add(&hist[i], sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// digit_scan: one workgroup of 256 threads per pass (workgroup p: pass p;
// thread d: digit d).  A pass's output holds the digits in order and, within
// a digit, the segments in order: run (x, d) starts at the keys of smaller
// digits plus the keys of digit d in segments before x.
__device__ void digit_scan_body(const mgpuScanArgs& a) {
    __shared__ uint32_t s_tmp[4];
    __shared__ uint64_t s_max[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t p = blockIdx.x;
    uint32_t c[MGPU_MAX_SEGS];
    uint32_t total = 0;
#pragma unroll
    for (uint32_t x = 0; x < MGPU_MAX_SEGS; ++x) {
        c[x] = x < a.num_segs ? a.hist[(x * MGPU_MAX_PASSES + p) * RADIX + tid] : 0u;
        total += c[x];
    }
    // The pass's most frequent digit: its value for segments per XCD, its
    // count for the early-exit report (a pass whose digit holds every key is
    // the identity).
    if (a.seg_count_words || a.report) {
        const uint64_t best = wave_max_u64(((uint64_t)total << 8) | tid);
        if (lane == 0) s_max[wave] = best;
        __syncthreads();
        if (tid == 0) {
            uint64_t m = s_max[0];
            for (int w = 1; w < 4; ++w) m = s_max[w] > m ? s_max[w] : m;
            a.tile_counters[MGPU_MODE_SLOT + p] = (uint32_t)(m & 0xFFu);
            if (a.report) {
                __hip_atomic_store(&a.report[2 + p], (uint32_t)(m >> 8), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                if (p == 0) {
                    __hip_atomic_store(&a.report[0], a.tile_counters[MGPU_UNSORTED_SLOT],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    __hip_atomic_store(&a.report[1], a.tile_counters[MGPU_UNSORTED_SLOT + 1],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
        }
        __syncthreads();  // the counters' reads above come before block 0 clears them
    }
    uint32_t run = block_exclusive_scan<256>(total, s_tmp, lane, wave);
#pragma unroll
    for (uint32_t x = 0; x < MGPU_MAX_SEGS; ++x) {
        if (x < a.num_segs) {
            const uint32_t i = (x * MGPU_MAX_PASSES + p) * RADIX + tid;
            a.digit_start[i] = run;
            run += c[x];
            a.hist[i] = 0;
        }
    }
    if (p == 0)
        for (uint32_t i = tid; i < MGPU_MODE_SLOT; i += 256) a.tile_counters[i] = 0;
    for (uint32_t i = p * 256 + tid; i < a.seg_count_words; i += gridDim.x * 256) a.seg_count[i] = 0;
}

// ---------------------------------------------------------------------------
// Device-wide passes over several arrays: digit_count per chunk of an array
// (one chunk per workgroup), digit_scan per array.
// Keys are read through to_global pointers (a descriptor's pointer would be
// read with flat loads).

template <int BLOCK>
__device__ void multi_count_body(const mgpuMultiCountArgs& a) {
    constexpr int W = BLOCK / 64;
    __shared__ uint32_t s_hist[W][MGPU_MAX_PASSES][RADIX];
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    uint32_t* flat = &s_hist[0][0][0];
    for (uint32_t i = tid; i < (uint32_t)(W * MGPU_MAX_PASSES * RADIX); i += BLOCK) flat[i] = 0;
    {
        uint4* c = reinterpret_cast<uint4*>(a.clear);
        const uint64_t nvec = a.clear_words >> 2;
        for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + tid; i < nvec;
             i += (uint64_t)gridDim.x * BLOCK)
            c[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (blockIdx.x >= a.num_chunks) return;  // clearing only (uniform per workgroup)
    __syncthreads();
    const uint32_t* ch = a.chunks + 4 * blockIdx.x;
    const uint32_t sg = ch[0], begin = ch[1], end = ch[2];
    const uint32_t* kp = a.segs[sg].keys[0];
    const auto keys = mgpu::to_global(kp);
    const uint32_t np = a.num_passes;
    uint32_t (*h)[RADIX] = s_hist[wave];
    auto count_key = [&](uint32_t k) {
#pragma unroll
        for (int p = 0; p < MGPU_MAX_PASSES; ++p)
            if ((uint32_t)p < np)
                __hip_atomic_fetch_add(&h[p][(k >> a.shift[p]) & a.mask[p]], 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    // Scalar head up to 16-byte alignment, vector body, scalar tail.
    const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(kp + begin) >> 2) & 3);
    const uint32_t vbeg = min(end, begin + ((4 - mis) & 3));
    const uint32_t nvec = (end - vbeg) >> 2;
    for (uint32_t i = begin + tid; i < vbeg; i += BLOCK) count_key(keys[i]);
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const auto v = mgpu::to_global(reinterpret_cast<const u32x4*>(kp + vbeg));
    // A 16K-key chunk is 16 vectors per thread: all of them in flight at once
    // (4096 segments in [1, 2^16]: 78 -> 73 us per launch).
    constexpr int UNROLL = 16;
    for (uint32_t i0 = tid; i0 < nvec; i0 += UNROLL * BLOCK) {
        u32x4 q[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (i0 + u * BLOCK < nvec) q[u] = __builtin_nontemporal_load(v + i0 + u * BLOCK);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            if (i0 + u * BLOCK >= nvec) break;
#pragma unroll
            for (int p = 0; p < MGPU_MAX_PASSES; ++p)
                if ((uint32_t)p < a.num_passes)
                    count4(h[p], q[u].x, q[u].y, q[u].z, q[u].w, a.shift[p], a.mask[p]);
        }
    }
    for (uint32_t i = vbeg + 4 * nvec + tid; i < end; i += BLOCK) count_key(keys[i]);
    __syncthreads();
    uint32_t* hist = a.hist + (size_t)sg * MGPU_MAX_PASSES * RADIX;
    for (uint32_t i = tid; i < (uint32_t)(MGPU_MAX_PASSES * RADIX); i += BLOCK) {
        if (i / RADIX >= np) break;
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) s += s_hist[w][i / RADIX][i % RADIX];
        if (s) __hip_atomic_fetch_add(&hist[i], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The same over whole arrays (or 64K-key pieces of larger ones), one
// 1024-thread workgroup each, with lane-column histograms: per (pass, digit
// pair) a row of 32 words, word c counting the keys of the lanes with
// lane % 32 == c in two 16-bit halves (even digit low, odd digit high).  The
// 32 lanes of a half-wave then always add to distinct banks, and the lanes
// that share a word (lane and lane + 32 of a wave) are at most two, where the
// per-wave histograms had same-digit lanes serialize on one word.  A column
// counts at most 32 threads x 64 keys of a 64K-key piece: 16 bits suffice.
__device__ void multi_count_cols_body(const mgpuMultiCountArgs& a) {
    constexpr int BLOCK = 1024;
    constexpr uint32_t COLS = 32, ROWS = MGPU_MAX_PASSES * RADIX / 2;
    __shared__ __attribute__((aligned(16))) uint32_t s_hist[ROWS * COLS];  // 64 KB
    const uint32_t tid = threadIdx.x, col = tid & (COLS - 1);
    for (uint32_t i = tid; i < ROWS * COLS / 4; i += BLOCK)
        reinterpret_cast<uint4*>(s_hist)[i] = make_uint4(0u, 0u, 0u, 0u);
    {
        uint4* c = reinterpret_cast<uint4*>(a.clear);
        const uint64_t nvec = a.clear_words >> 2;
        for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + tid; i < nvec;
             i += (uint64_t)gridDim.x * BLOCK)
            c[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (blockIdx.x >= a.num_chunks) return;  // clearing only (uniform per workgroup)
    __syncthreads();
    const uint32_t* ch = a.chunks + 4 * blockIdx.x;
    const uint32_t sg = ch[0], begin = ch[1], end = ch[2];
    const uint32_t* kp = a.segs[sg].keys[0];
    const auto keys = mgpu::to_global(kp);
    const uint32_t np = a.num_passes;
    // Pass p's word for key k: row d / 2 of the pass's RADIX / 2 (digit d =
    // the field of `width` bits at `shift`), the parity of d picking the
    // 16-bit half.  Both come from bit-field extracts: the row is the field
    // without its low bit, the increment 1 or 1 << 16 by that bit.
    uint32_t row_at[MGPU_MAX_PASSES], row_bits[MGPU_MAX_PASSES];
#pragma unroll
    for (int p = 0; p < MGPU_MAX_PASSES; ++p) {
        const uint32_t width = (uint32_t)__builtin_popcount(a.mask[p]);
        row_at[p] = a.shift[p] + 1;
        row_bits[p] = width > 0 ? width - 1 : 0;
    }
    auto add = [&](int p, uint32_t k) {
        const uint32_t row = __builtin_amdgcn_ubfe(k, row_at[p], row_bits[p]);
        const uint32_t inc = __builtin_amdgcn_ubfe(k, a.shift[p], 1) * 0xFFFFu + 1u;
        __hip_atomic_fetch_add(&s_hist[((uint32_t)p * (RADIX / 2) + row) * COLS + col], inc,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto count_key = [&](uint32_t k) {
#pragma unroll
        for (int p = 0; p < MGPU_MAX_PASSES; ++p)
            if ((uint32_t)p < np) add(p, k);
    };
    // Scalar head up to 16-byte alignment, vector body, scalar tail.
    const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(kp + begin) >> 2) & 3);
    const uint32_t vbeg = min(end, begin + ((4 - mis) & 3));
    const uint32_t nvec = (end - vbeg) >> 2;
    for (uint32_t i = begin + tid; i < vbeg; i += BLOCK) count_key(keys[i]);
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const auto v = mgpu::to_global(reinterpret_cast<const u32x4*>(kp + vbeg));
    constexpr int UNROLL = 16;  // a 64K-key piece: every vector of a thread in flight
    for (uint32_t i0 = tid; i0 < nvec; i0 += UNROLL * BLOCK) {
        u32x4 q[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (i0 + u * BLOCK < nvec) q[u] = __builtin_nontemporal_load(v + i0 + u * BLOCK);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            if (i0 + u * BLOCK >= nvec) break;
            // The pass test once per pass and vector, not per key.
#pragma unroll
            for (int p = 0; p < MGPU_MAX_PASSES; ++p) {
                if ((uint32_t)p >= np) break;
                add(p, q[u].x);
                add(p, q[u].y);
                add(p, q[u].z);
                add(p, q[u].w);
            }
        }
    }
    for (uint32_t i = vbeg + 4 * nvec + tid; i < end; i += BLOCK) count_key(keys[i]);
    __syncthreads();
    // Thread (p, d) = (tid / 256, tid % 256) and its partner tid ^ 1 hold the
    // two digits of row tid / 2: each sums half the row's columns, both
    // halves of every word, and hands the partner the sum of the partner's
    // half.  Reads start at column tid / 2 of the thread's 16 so that a
    // half-wave reads 32 distinct banks.
    const uint32_t p = tid / RADIX, d = tid % RADIX;
    const uint32_t h = tid & 1u, row = tid >> 1;
    uint32_t lo = 0, hi = 0;
    if (p < np) {
#pragma unroll
        for (uint32_t j = 0; j < COLS / 2; ++j) {
            const uint32_t w = s_hist[row * COLS + h * (COLS / 2) + ((j + row) & (COLS / 2 - 1))];
            lo += w & 0xFFFFu;
            hi += w >> 16;
        }
    }
    const uint32_t sum = (h ? hi : lo) + (uint32_t)__shfl_xor((int)(h ? lo : hi), 1, 64);
    if (a.direct) {
        // The piece is the whole array: its digit starts, pass by pass (a
        // scan per group of 256 threads), with no global histogram and no
        // separate scan launch.  Workgroup 0 zeroes the passes' tile
        // counters, as that launch would.
        __shared__ uint32_t s_tmp[BLOCK / 64];
        const uint32_t ex = group_exclusive_scan<BLOCK>(sum, s_tmp, tid & 63, tid >> 6);
        if (p < np) a.digit_start[((size_t)sg * MGPU_MAX_PASSES + p) * RADIX + d] = ex;
        if (blockIdx.x == 0 && tid < MGPU_LANE_COUNTERS) a.tile_counters[tid] = 0;
    } else if (sum) {
        __hip_atomic_fetch_add(&a.hist[((size_t)sg * MGPU_MAX_PASSES + p) * RADIX + d], sum,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ void multi_scan_body(const mgpuMultiScanArgs& a) {
    __shared__ uint32_t s_tmp[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * MGPU_MAX_PASSES * RADIX;
    for (uint32_t p = 0; p < a.num_passes; ++p) {
        const uint32_t c = a.hist[base + p * RADIX + tid];
        const uint32_t e = block_exclusive_scan<256>(c, s_tmp, lane, wave);
        a.digit_start[base + p * RADIX + tid] = e;
        a.hist[base + p * RADIX + tid] = 0;
    }
    if (blockIdx.x == 0 && tid < MGPU_LANE_COUNTERS) a.tile_counters[tid] = 0;
}

// ---------------------------------------------------------------------------
// rank_scatter: one digit pass, one tile of BLOCK*IPT keys per workgroup.
// MULTI: tiles of several arrays (mgpuMultiPassArgs), else of the single array
// of mgpuPassArgs.

// Decoupled lookback for digit d of local tile `tile` by a group of G
// consecutive lanes (k = 0..G-1), each holding K words: one round reads the
// lookback words of the G*K nearest predecessors not yet counted (lane k, word
// q: tile t0 - k*K - q, nearest first), then adds up the counts up to and
// including the nearest inclusive prefix, or all of them and goes on.  A round
// whose needed words are not all published yet is read again.  Tile 0 always
// publishes a prefix, so the walk ends there at the latest; positions before
// tile 0 read as an empty prefix and are never counted.  Loads are relaxed:
// a word's count is all that is used of the tile that published it.
#ifndef MGPU_LOOKBACK_WINDOW
#define MGPU_LOOKBACK_WINDOW 16
#endif

// LOCAL chains (one XCD's segment, read and written only by that XCD's
// workgroups) use nontemporal loads and stores, which reach the XCD's L2 but
// do not stay in a CU's L1, so a poll sees another CU's store at L2 latency.
// The engine checks that behaviour (mgpu_xcd_probe) before it picks LOCAL
// kernels.  Other chains use agent-scope atomics, coherent across XCDs.
template <bool LOCAL>
__device__ __forceinline__ uint32_t lookback_word(const uint32_t* lb, int32_t t, uint32_t d) {
    if (t < 0) return MGPU_LB_PREFIX;
    const uint32_t* p = &lb[(uint32_t)t * RADIX + d];
    if constexpr (LOCAL) return __builtin_nontemporal_load(p);
    else return ld_agent(p);
}

template <bool LOCAL>
__device__ __forceinline__ void lookback_publish(uint32_t* p, uint32_t v) {
    if constexpr (LOCAL) __builtin_nontemporal_store(v, p);
    else st_agent(p, v);
}

template <uint32_t K, bool LOCAL>
__device__ __forceinline__ void lookback_read(const uint32_t* lb, int32_t t0, uint32_t d,
                                              uint32_t k, uint32_t (&w)[K]) {
    // Plain (nontemporal) loads: keep the compiler from reusing an earlier
    // round's values.
    if constexpr (LOCAL) asm volatile("" ::: "memory");
#pragma unroll
    for (uint32_t q = 0; q < K; ++q) w[q] = lookback_word<LOCAL>(lb, t0 - (int32_t)(k * K + q), d);
}

// Returns the exclusive prefix of the tile for digit d.  `w` holds the first
// round's words, read by the caller with lookback_read(lb, tile - 1, ...).
// A word still unpublished after spin_limit polls (about a second at the
// default; only a broken tile counter could make a wait that long) ends the
// wait with what has been summed so far and sets error code 1.  spin_limit 0
// ends every wait at once, before anything is summed (tests of the error
// path: every tile after a chain's first one gives up).
template <uint32_t G, uint32_t K, bool LOCAL>
__device__ __forceinline__ uint32_t lookback_window(const uint32_t* lb, uint32_t tile, uint32_t d,
                                                    uint32_t k, uint32_t lane,
                                                    uint32_t* error_flag, uint32_t spin_limit,
                                                    uint32_t (&w)[K]) {
    if (tile == 0) return 0;
    if (spin_limit == 0) {
        set_error(error_flag, 1u);
        return 0;
    }
    constexpr uint32_t GM = (1u << G) - 1u;
    const uint32_t gfirst = lane & ~(G - 1u);  // the group's first lane in the wave
    uint32_t prefix = 0, spins = 0;
    int32_t t0 = (int32_t)tile - 1;  // nearest tile of this round
    while (true) {
        // This lane's words up to its nearest prefix: all published?  Sum.
        bool ready = true, has_pref = false;
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t q = 0; q < K; ++q) {
            if (!has_pref) {
                ready = ready && (w[q] & MGPU_LB_STATE) != 0;
                sum += w[q] & MGPU_LB_VALUE;
                has_pref = (w[q] & MGPU_LB_PREFIX) != 0;
            }
        }
        const uint32_t rdy = (uint32_t)(__builtin_amdgcn_ballot_w64(ready) >> gfirst) & GM;
        const uint32_t pref = (uint32_t)(__builtin_amdgcn_ballot_w64(has_pref) >> gfirst) & GM;
        // Lanes whose words count: up to the first with a prefix, or all.
        const uint32_t upto = pref ? ((pref & (0u - pref)) << 1) - 1u : GM;
        if ((rdy & upto) != upto) {
            if (++spins > spin_limit) {
                set_error(error_flag, 1u);
                return prefix;
            }
            __builtin_amdgcn_s_sleep(1);
            lookback_read<K, LOCAL>(lb, t0, d, k, w);
            continue;
        }
        uint32_t v = ((upto >> k) & 1u) ? sum : 0u;
#pragma unroll
        for (uint32_t off = 1; off < G; off <<= 1) v += __shfl_xor(v, off, G);
        prefix += v;
        if (pref) return prefix;
        t0 -= (int32_t)(G * K);
        lookback_read<K, LOCAL>(lb, t0, d, k, w);
    }
}

// Where one tile reads and writes.
struct TileJob {
    const uint32_t* keys_in;
    uint32_t* keys_out;
    const mgpuSegDesc* sd;  // MULTI: the array's descriptor
    const uint32_t* digit_start;
    uint32_t* lookback;     // lookback words from the chain's first tile on
    uint32_t n;             // keys of the array
    uint32_t tile;          // tile index within the chain
    uint32_t first;         // tiles of the array before the chain's first tile
};

// Global tile t of a pass: which array (MULTI) and where its data lives.
// One load of the host-built tile map: a binary search over tile_begin was a
// chain of ~10 dependent global loads ahead of every tile.
template <bool MULTI, typename Args>
__device__ __forceinline__ uint32_t find_array(const Args& a, uint32_t t) {
    if constexpr (MULTI) {
        return a.tile_seg[t];
    } else {
        (void)a;
        (void)t;
        return 0;
    }
}

template <bool MULTI, typename Args>
__device__ __forceinline__ TileJob make_job(const Args& a, uint32_t gtile, uint32_t seg) {
    TileJob j;
    if constexpr (MULTI) {
        j.sd = &a.segs[seg];
        const uint32_t tb = a.tile_begin[seg];
        j.keys_in = j.sd->keys[a.in_side];
        j.keys_out = j.sd->keys[a.in_side ^ 1];
        j.n = j.sd->n;
        j.tile = gtile - tb;
        j.first = 0;
        j.digit_start = a.digit_start + (seg * MGPU_MAX_PASSES + a.pass) * RADIX;
        j.lookback = a.lookback + (size_t)tb * RADIX;
    } else {
        (void)seg;
        j.sd = nullptr;
        j.keys_in = a.keys_in;
        j.keys_out = a.keys_out;
        j.n = a.n;
        j.tile = gtile;
        j.first = 0;
        j.digit_start = a.digit_start;
        j.lookback = a.lookback;
    }
    return j;
}

template <bool MULTI, typename Args>
__device__ __forceinline__ const uint32_t* job_vals_in(const Args& a, const TileJob& j,
                                                       uint32_t v) {
    if constexpr (MULTI) return j.sd->vals[v][a.in_side];
    else return a.vals_in[v];
}

template <bool MULTI, typename Args>
__device__ __forceinline__ uint32_t* job_vals_out(const Args& a, const TileJob& j, uint32_t v) {
    if constexpr (MULTI) return j.sd->vals[v][a.in_side ^ 1];
    else return a.vals_out[v];
}

// One tile per workgroup over several arrays whose tiles come in
// a.num_xcds groups (whole arrays, group g holding global tiles
// [xcd_tile_begin[g], xcd_tile_begin[g + 1])), each group with its own
// counter: workgroup b asks group b % groups first, then the groups after
// it.  One counter for every tile of a launch serializes the workgroups'
// atomics on one word (~11 ns each: 3000 tiles, 38 us); spread over 8 words
// they overlap.  A tile only waits on earlier tiles of its own array, which
// its group's counter handed out before it, so no tile waits on one that is
// not running.  Returns num_tiles when every group is taken.
__device__ __forceinline__ uint32_t take_grouped_tile(const mgpuMultiPassArgs& a) {
    const uint32_t groups = a.num_xcds;
    uint32_t g = blockIdx.x % groups;
    for (uint32_t k = 0; k < groups; ++k) {
        const uint32_t b = a.xcd_tile_begin[g], e = a.xcd_tile_begin[g + 1];
        if (b < e) {
            const uint32_t x = __hip_atomic_fetch_add(a.tile_counter + g, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
            if (x < e - b) return b + x;
        }
        g = g + 1 == groups ? 0u : g + 1;
    }
    return a.num_tiles;
}

// Striped load of one array of a tile: slot j of lane l of wave w holds
// element tile_base + w*64*IPT + j*64 + l.  Raw buffer loads with the array
// length as the range: the hardware returns 0 for elements at or past n, so a
// partial tile takes the same branch-free path as a full one (callers never
// use those slots), and the compiler can wait for each slot separately.  One
// per-lane byte offset plus constant offsets (j * 256 bytes) address the
// slots.  Offsets past 2^32 bytes (only past n, in the last tile of the
// largest arrays) wrap to earlier elements of the same array: still in range,
// read and ignored.
constexpr int kBufferDword3 = 0x00020000;  // 32-bit data format, no swizzle
// Cache policy of the tile loads: the default.  Marking them nontemporal
// made a 2^28 key+value sort 12% slower (4.28 against 3.76 ms); marking the
// scattered stores nontemporal made it 57% slower (their partial lines are
// merged in L2).
constexpr int kTileLoadAux = 0;

// Buffer resource over an array of n words (n < 2^30, so byte offsets fit in
// 32 bits): loads at or past n return 0, stores there are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t array_rsrc(const uint32_t* p, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(p), 0, (int)(n * 4u),
                                             kBufferDword3);
}

template <int IPT>
__device__ __forceinline__ void load_striped(const uint32_t* src, uint32_t n, uint32_t wave_base,
                                             uint32_t lane, uint32_t* out) {
    const __amdgpu_buffer_rsrc_t r = array_rsrc(src, n);
    const uint32_t off = (wave_base + lane) * 4u;
#pragma unroll
    for (int j = 0; j < IPT; ++j)
        out[j] = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + j * 256u), 0, kTileLoadAux);
}

// RANK 0: in-wave ranking by LDS match masks (documented behaviour only).
// RANK 1: one LDS atomic add per key on the wave's count of its digit: the
// returned old count is the key's stable rank when the lanes of one
// instruction that hit the same count are served in ascending lane order,
// which the engine checks at creation (mgpu_rank_probe) before it picks
// these kernels.
// XSEG: the workgroup takes tiles of its XCD's share until it is done.  A
// single array is cut into one segment of tiles per XCD with a lookback chain
// of its own (each pass counting the next one's per-segment digit starts);
// with MULTI, whole arrays are assigned to XCDs, so each array's chain stays
// on one XCD and nothing else changes.
// EE (early exit, RANK 1 only): a wave whose keys are already in digit order
// ranks them by position instead of by LDS atomics (the role of the
// reference's sorted-block path, sort.cu:135-158, with the check made here
// per wave instead of per block in a count kernel).  Its keys then stage to
// consecutive LDS words, free of bank conflicts; a tile whose waves all are in
// order, and in order across waves, stages every key to its own place.
template <int BLOCK, int IPT, int RANK, int VMODE, bool MULTI, bool XSEG, bool EE, bool KMAP,
          typename Args>
__device__ void rank_scatter_body(const Args& a) {
    static_assert(!EE || RANK == 1, "early exit builds on the atomic-add ranking");
    static_assert(!(EE && MULTI), "early exit is for single arrays");
    constexpr bool XSINGLE = XSEG && !MULTI;  // one array cut into segments per XCD
    static_assert(BLOCK % 64 == 0 && BLOCK >= (int)RADIX, "block must cover the radix");
    constexpr int W = BLOCK / 64;
    constexpr uint32_t TILE = BLOCK * IPT;
    constexpr uint32_t WAVE_TILE = 64 * IPT;

    // One value array or generated indices are staged beside the keys.  The
    // in-wave match masks are only live during ranking, before anything is
    // staged, so they share the value stage's memory.
    constexpr bool JOINT = VMODE == MGPU_VALUES_SINGLE || VMODE == MGPU_VALUES_INDEX;
    constexpr uint32_t MATCH_WORDS = RANK == 0 ? (uint32_t)W * RADIX * 2 : 0;
    constexpr uint32_t AUX_WORDS0 = JOINT && TILE > MATCH_WORDS ? TILE : MATCH_WORDS;
    constexpr uint32_t AUX_WORDS = AUX_WORDS0 > 0 ? AUX_WORDS0 : 2;
    __shared__ uint32_t s_stage[TILE];
    __shared__ __attribute__((aligned(16))) uint32_t s_aux[AUX_WORDS];
    __shared__ uint32_t s_count[W][RADIX];
    uint64_t(*const s_match)[RADIX] = reinterpret_cast<uint64_t(*)[RADIX]>(s_aux);
    uint32_t* const s_vstage = s_aux;
    __shared__ uint32_t s_global[RADIX];
    __shared__ uint32_t s_tmp[W];
    __shared__ uint32_t s_tile[2];
    __shared__ uint32_t s_lstart[RADIX];
    __shared__ uint32_t s_tcount[RADIX];
    // XSEG: this segment's digit starts, and the next pass's digit counts of
    // the keys this workgroup writes, by destination segment.
    constexpr uint32_t NSEG = MGPU_XSEG_COUNT_SEGS;
    __shared__ uint32_t s_dstart[XSINGLE ? RADIX : 1];
    __shared__ uint32_t s_next[XSINGLE ? NSEG * RADIX : 1];
    __shared__ uint32_t s_spare[XSINGLE ? 64 : 1];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63;
    const uint32_t wave = tid >> 6;
    const uint32_t shift = a.shift, mask = a.mask;
    const uint32_t width = (uint32_t)__builtin_popcount(mask);  // mask = 2^width - 1

    // The segment of this workgroup's XCD (XSEG), else the whole array.
    uint32_t seg = 0, seg_first = 0, seg_tiles = a.num_tiles;
    if constexpr (XSEG) {
        uint32_t xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        seg = xcc & 0xFu;
        if constexpr (MULTI) {
            if (seg < a.num_xcds) {
                seg_first = a.xcd_tile_begin[seg];
                seg_tiles = a.xcd_tile_begin[seg + 1] - seg_first;
            } else {
                seg_tiles = 0;
            }
        } else if (seg < a.num_segs) {
            seg_first = mgpu_seg_first_tile(seg, a.num_tiles, a.num_segs);
            seg_tiles = mgpu_seg_first_tile(seg + 1, a.num_tiles, a.num_segs) - seg_first;
        } else {
            seg_tiles = 0;  // no segment for this XCD: nothing to take
        }
    }
    for (uint32_t i = tid; i < (uint32_t)(W * RADIX); i += BLOCK) {
        (&s_count[0][0])[i] = 0;
        if constexpr (RANK == 0) (&s_match[0][0])[i] = 0;
    }
    // XSEG: whether this pass counts for the next one, and this segment's
    // digit starts.
    bool count_next = false;
    // Segment of output position q: ((q / TILE + 1) * S - 1) / T for S
    // segments of T tiles (the largest x with floor(x * T / S) * TILE <= q),
    // as fma(num, 1/T, 0.5/T) truncated: num < 2^21 is exact in a float, the
    // quotient is at least 0.5/T >= 2^-17 from an integer, and the rounding of
    // 1/T and of the fma errs by less than 2^-19.
    const float seg_rcp = 1.0f / (float)a.num_tiles;
    const float seg_half = 0.5f * seg_rcp;
    uint32_t next_mode = RADIX;  // most frequent digit of the next pass
    uint32_t next_width = 0;
    if constexpr (XSINGLE) {
        count_next = a.seg_count_out != nullptr;
        if (count_next) {
            next_mode = *a.next_mode;
            next_width = (uint32_t)__builtin_popcount(a.next_mask);
            for (uint32_t i = tid; i < NSEG * RADIX; i += BLOCK) s_next[i] = 0;
        }
        if (tid < RADIX && seg < a.num_segs) {
            uint32_t st;
            if (a.seg_count_in) {
                st = a.digit_start[tid];
                for (uint32_t x = 0; x < seg; ++x) st += a.seg_count_in[x * RADIX + tid];
            } else {
                st = a.digit_start[seg * (MGPU_MAX_PASSES * RADIX) + tid];
            }
            s_dstart[tid] = st;
        }
    }

    while (true) {
        // The tile comes from the counter in start order (and is sorted right
        // away), so a tile only ever waits on tiles that are already running.
        if (tid == 0) {
            uint32_t t = seg_tiles;  // a workgroup of an XCD without a segment takes none
            bool grouped = false;
            if constexpr (MULTI && !XSEG) grouped = a.num_xcds > 1;
            if (grouped) {
                if constexpr (MULTI && !XSEG) t = take_grouped_tile(a);
            } else if (seg_tiles > 0) {
                t = __hip_atomic_fetch_add(a.tile_counter + seg, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
            s_tile[0] = t;
            s_tile[1] = MULTI && t < seg_tiles ? find_array<MULTI>(a, seg_first + t) : 0u;
        }
        __syncthreads();
        // Tile and array indices are the same for every lane: keep them (and all
        // that is derived from them) in scalar registers.
        const uint32_t gtile = __builtin_amdgcn_readfirstlane(s_tile[0]);
        if (gtile >= seg_tiles) {
            // One tile per workgroup: only a counter left dirty gets here.  A
            // segment's workgroup: the segment is done.
            if (!XSEG && tid == 0) set_error(a.error_flag, 2u);
            break;
        }
        // MULTI: tiles are numbered across the arrays; XSINGLE: within the
        // segment, whose chain starts at its first tile.
        TileJob job = make_job<MULTI>(a, MULTI ? seg_first + gtile : gtile,
                                      __builtin_amdgcn_readfirstlane(s_tile[1]));
        if constexpr (XSINGLE) {
            job.first = seg_first;
            job.lookback += (size_t)seg_first * RADIX;
        }
        const uint32_t n = job.n, tile = job.tile;
        const uint32_t base = (job.first + tile) * TILE;
        const uint32_t wave_base = base + wave * WAVE_TILE;
        const uint32_t tile_valid = min(TILE, n - base);
        const bool full = tile_valid == TILE;

        // Keys and the single value array, striped, requested together: ranking
        // starts as soon as each key slot is in while the values still load.
        uint32_t key[IPT], val[IPT];
        load_striped<IPT>(job.keys_in, n, wave_base, lane, key);
        if constexpr (VMODE == MGPU_VALUES_SINGLE)
            load_striped<IPT>(job_vals_in<MULTI>(a, job, 0), n, wave_base, lane, val);
        // Signed or float keys (KMAP), in the sort's first pass: to their
        // unsigned order.
        if constexpr (KMAP) {
            const uint32_t tin = a.key_types & 0xFFu;
            if (tin != 0) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) key[j] = mgpu_key_to_radix(key[j], tin);
            }
        }

        // Keys past n (last tile only) get every bit set: they take the pass's
        // top digit (mask) and, last in input order, rank after every real key
        // of the tile.  Digit of slot j: one bit-field extract.
        const int32_t rem = (int32_t)(n - (wave_base + lane));  // slots j*64 < rem are real
        if (!full) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) key[j] = j * 64 < rem ? key[j] : 0xFFFFFFFFu;
        }
        auto digit_of = [&](int j) -> uint32_t { return __builtin_amdgcn_ubfe(key[j], shift, width); };

        // Rank within the wave: pos[j] = earlier same-digit keys of this wave.
        uint32_t pos[IPT];
        bool wave_sorted = false;
        if constexpr (EE) {
            // Key q = j * 64 + lane of the wave against key q - 1.
            uint64_t desc = 0;
            uint32_t tail = 0;  // digit of the previous slot's lane 63 (none before slot 0)
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t d = digit_of(j);
                desc |= __builtin_amdgcn_ballot_w64(prev_lane(d, tail) > d);
                tail = __builtin_amdgcn_readlane(d, 63);
            }
            wave_sorted = desc == 0;
        }
        if (EE && wave_sorted) {
            // Runs of equal digits in key order: a key's rank is its distance
            // from its run's first key, and a run's length is the digit's
            // count.  The run starts live in the staging area, unused until
            // after the next barrier.
            uint32_t* const first = s_stage + wave * RADIX;
            uint32_t tail = RADIX;  // no digit: key 0 starts a run
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t d = digit_of(j);
                if (prev_lane(d, tail) != d) first[d] = (uint32_t)j * 64 + lane;
                tail = __builtin_amdgcn_readlane(d, 63);
            }
            wave_sync();
            tail = RADIX;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t d = digit_of(j), q = (uint32_t)j * 64 + lane;
                pos[j] = q - first[d];
                const uint32_t pd = prev_lane(d, tail);
                if (q != 0 && pd != d) s_count[wave][pd] = q - first[pd];  // run of pd ends at q
                tail = __builtin_amdgcn_readlane(d, 63);
            }
            if (lane == 63) {
                const uint32_t d = digit_of(IPT - 1);
                s_count[wave][d] = WAVE_TILE - first[d];
            }
        } else {
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t d = digit_of(j);
            if constexpr (RANK == 1) {
                // A slot whose 64 keys share one digit (sorted or low-entropy
                // input) takes its ranks with one add instead of 64 adds to one
                // word, which the LDS would serialize.
                const uint32_t f = __builtin_amdgcn_readfirstlane(d);
                if (__builtin_amdgcn_ballot_w64(d == f) == ~0ull) {
                    uint32_t base = 0;
                    if (lane == 0)
                        base = __hip_atomic_fetch_add(&s_count[wave][f], 64u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                    pos[j] = __builtin_amdgcn_readfirstlane(base) + lane;
                } else {
                    pos[j] = __hip_atomic_fetch_add(&s_count[wave][d], 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            } else {
                // Each lane ORs its bit into the wave's mask for its digit and
                // reads the mask back; LDS operations of one wave complete in
                // order, so the mask holds exactly the lanes of this slot with
                // that digit.
                __hip_atomic_fetch_or(&s_match[wave][d], 1ull << lane, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint64_t m = __hip_atomic_load(&s_match[wave][d], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t below = lanes_below(m);
                const uint32_t old = s_count[wave][d];
                pos[j] = old + below;
                if (below == 0) {  // lowest lane of the group advances the count
                    s_count[wave][d] = old + (uint32_t)__builtin_popcountll(m);
                    __hip_atomic_store(&s_match[wave][d], 0ull, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
        }  // ranking
        __syncthreads();

        // Per digit (thread d): wave-exclusive offsets and the tile total,
        // published at once for the tiles after this one; then the tile-local
        // digit starts (block scan) are added to the wave offsets.
        uint32_t total = 0, tile_count = 0;
        if (tid < RADIX) {
            uint32_t sum = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint32_t c = s_count[w][tid];
                s_count[w][tid] = sum;
                sum += c;
            }
            total = sum;
            tile_count = (tid == mask) ? sum - (TILE - tile_valid) : sum;
            lookback_publish<XSEG>(&job.lookback[tile * RADIX + tid],
                                   (tile == 0 ? MGPU_LB_PREFIX : MGPU_LB_AGGREGATE) | tile_count);
        }
        const uint32_t local_start = block_exclusive_scan<BLOCK>(total, s_tmp, lane, wave);
        if (tid < RADIX) {
            s_lstart[tid] = local_start;
            s_tcount[tid] = tile_count;
#pragma unroll
            for (int w = 0; w < W; ++w) s_count[w][tid] += local_start;
        }
        __syncthreads();

        // BLOCK / RADIX threads per digit walk back over the predecessors.  The
        // first round's words are requested before the staging below so that
        // their latency overlaps it (starting earlier only waits longer for the
        // predecessors' own totals).
        constexpr uint32_t G = BLOCK / RADIX;
        constexpr uint32_t K = MGPU_LOOKBACK_WINDOW / G > 0 ? MGPU_LOOKBACK_WINDOW / G : 1;
        const uint32_t lb_digit = tid / G, lb_k = tid % G;
        uint32_t lb_words[K];
        lookback_read<K, XSEG>(job.lookback, (int32_t)tile - 1, lb_digit, lb_k, lb_words);

        // Tile-local position of every key; stage the keys in digit order (and,
        // with one value array or generated indices, the values beside them, so
        // that both leave after a single barrier).
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            pos[j] += s_count[wave][digit_of(j)];
            s_stage[pos[j]] = key[j];
            if constexpr (JOINT) {
                if constexpr (VMODE == MGPU_VALUES_INDEX) s_vstage[pos[j]] = wave_base + lane + j * 64;
                else s_vstage[pos[j]] = val[j];
            }
        }

        {
            const uint32_t prefix = lookback_window<G, K, XSEG>(
                job.lookback, tile, lb_digit, lb_k, lane, a.error_flag, a.spin_limit, lb_words);
            if (lb_k == 0) {
                if (tile > 0)
                    lookback_publish<XSEG>(&job.lookback[tile * RADIX + lb_digit],
                                           MGPU_LB_PREFIX | (prefix + s_tcount[lb_digit]));
                const uint32_t dstart = XSINGLE ? s_dstart[lb_digit] : job.digit_start[lb_digit];
                s_global[lb_digit] = dstart + prefix - s_lstart[lb_digit];
            }
        }
        __syncthreads();

        // Write out in tile order: consecutive threads hold consecutive
        // positions, so each digit's run of the tile leaves as one stretch.
        // The LDS reads of every slot are issued before the first store, so
        // that their latencies overlap instead of adding up slot by slot
        // (slots past the tile's keys read LDS words that are not used).
        const int32_t vrem = (int32_t)(tile_valid - tid);  // positions i*BLOCK < vrem are real
        uint32_t* const vout0 = JOINT ? job_vals_out<MULTI>(a, job, 0) : nullptr;
        uint32_t kout[IPT], dest[IPT], vout[JOINT ? IPT : 1];
#pragma unroll
        for (int i = 0; i < IPT; ++i) kout[i] = s_stage[i * BLOCK + tid];
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            dest[i] = s_global[__builtin_amdgcn_ubfe(kout[i], shift, width)] + i * BLOCK + tid;
            if constexpr (JOINT) vout[i] = s_vstage[i * BLOCK + tid];
        }
        // Signed or float keys (KMAP), in the sort's last pass: back to their
        // type (no pass after it reads their digits).
        if constexpr (KMAP) {
            const uint32_t tout = a.key_types >> 8;
            if (tout != 0) {
#pragma unroll
                for (int i = 0; i < IPT; ++i) kout[i] = mgpu_key_from_radix(kout[i], tout);
            }
        }
        // Offsets outside [0, n) can only come from a broken count or
        // lookback: they are flagged instead of written.  Slots with nothing
        // to write get dest = n, where the buffer stores' range check drops
        // them (one 32-bit offset per key instead of 64-bit addresses and a
        // branch).
        const __amdgpu_buffer_rsrc_t rk = array_rsrc(job.keys_out, n);
        const __amdgpu_buffer_rsrc_t rv = JOINT ? array_rsrc(vout0, n) : rk;
        bool bad = false;
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            const bool real = full || i * BLOCK < vrem;
            const bool ok = real && dest[i] < n;
            bad = bad || (real && !ok);
            dest[i] = ok ? dest[i] : n;
            __builtin_amdgcn_raw_buffer_store_b32(kout[i], rk, (int)(dest[i] * 4u), 0, 0);
            if constexpr (JOINT)
                __builtin_amdgcn_raw_buffer_store_b32(vout[i], rv, (int)(dest[i] * 4u), 0, 0);
        }
        if (bad) set_error(a.error_flag, 4u);
        // Segments per XCD: the next pass's digit of every key written, by
        // the segment of the next pass that its position falls in.
        // (Computing the segment costs less than a per-digit LDS lookup.)
        if constexpr (XSINGLE) {
            if (count_next) {
#pragma unroll
                for (int i = 0; i < IPT; ++i) {
                    if (dest[i] < n) {
                        // dest / TILE < 2^16, num_segs <= 8: a 24-bit multiply.
                        const uint32_t num = __umul24(dest[i] / TILE, a.num_segs) + a.num_segs - 1;
                        const uint32_t xo =
                            (uint32_t)__builtin_fmaf((float)num, seg_rcp, seg_half);
                        const uint32_t nd = __builtin_amdgcn_ubfe(kout[i], a.next_shift, next_width);
                        count1(s_next, s_spare, xo * RADIX + nd, nd == next_mode, lane);
                    }
                }
            }
        }

        // Two or more value arrays: one after the other through the key stage.
        if constexpr (VMODE == MGPU_VALUES_MULTI) {
            for (uint32_t v = 0; v < a.num_values; ++v) {
                load_striped<IPT>(job_vals_in<MULTI>(a, job, v), n, wave_base, lane, val);
                uint32_t* dstv = job_vals_out<MULTI>(a, job, v);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < IPT; ++j) s_stage[pos[j]] = val[j];
                __syncthreads();
#pragma unroll
                for (int i = 0; i < IPT; ++i)
                    if (dest[i] < n) dstv[dest[i]] = s_stage[i * BLOCK + tid];
            }
        }

        if constexpr (!XSEG) break;
        __syncthreads();  // every read of this tile's LDS state is done
        for (uint32_t i = tid; i < (uint32_t)(W * RADIX); i += BLOCK) {
            (&s_count[0][0])[i] = 0;
            // The value stage wrote over the match masks.
            if constexpr (RANK == 0 && JOINT) (&s_match[0][0])[i] = 0;
        }
    }  // while

    // The next pass's counts of this workgroup's keys, and the last
    // workgroup to finish checks that every segment was taken to its end
    // (each needs a workgroup on its XCD).
    if constexpr (XSINGLE) {
        if (count_next) {
            // Two digits' counts per 64-bit atomic (half the device-scope
            // atomics, which every workgroup issues at the same moment): a
            // word's total is at most n < 2^31, so the low half never
            // carries into the high one.  seg_count_out is 8-byte aligned.
            __syncthreads();
            uint64_t* const out2 = reinterpret_cast<uint64_t*>(a.seg_count_out);
            for (uint32_t i = tid; i < NSEG * RADIX / 2; i += BLOCK) {
                const uint64_t c = (uint64_t)s_next[2 * i + 1] << 32 | s_next[2 * i];
                if (c) __hip_atomic_fetch_add(&out2[i], c, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    if constexpr (XSEG) {
        if (tid == 0) {
            const uint32_t done =
                __hip_atomic_fetch_add(a.finished, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (done + 1 == gridDim.x) {
                uint32_t segs;
                if constexpr (MULTI) segs = a.num_xcds;
                else segs = a.num_segs;
                for (uint32_t x = 0; x < segs; ++x) {
                    uint32_t tiles;
                    if constexpr (MULTI) tiles = a.xcd_tile_begin[x + 1] - a.xcd_tile_begin[x];
                    else tiles = mgpu_seg_first_tile(x + 1, a.num_tiles, a.num_segs) -
                                 mgpu_seg_first_tile(x, a.num_tiles, a.num_segs);
                    if (ld_agent(a.tile_counter + x) < tiles) set_error(a.error_flag, 5u);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Packed device-wide passes of a segmented sort (mgpu_multi_packed_*).  The
// arrays are cut into wave chunks of 64 * IPT keys and laid out back to back,
// BLOCK / 64 chunks to a tile, so an array's last, partial chunk shares its
// tile with the arrays after it instead of leaving the rest of a tile empty
// (the tile-per-array layout of rank_scatter_body<MULTI> ran 3460 tiles for
// the keys of 2783 on 4096 segments log-uniform in [1, 2^16]).  A run of
// chunks of one array in a tile is a part.  Each wave ranks its chunk as in
// rank_scatter_body (RANK 1); the tile is staged part by part, each part in
// digit order, and written out in rounds of BLOCK consecutive positions,
// each inside one chunk-sized stretch of the stage and so inside one part:
// the array a round writes to is the same for every lane.  Only a tile whose
// first chunk continues an array begun earlier walks back along the chain.
// Every tile publishes the digit counts of its last part: as a prefix when
// that array starts in the tile (nothing before it counts), else as an
// aggregate, and as a prefix once its own walk has ended.
template <int BLOCK, int IPT, int VMODE, bool XSEG>
__device__ void packed_scatter_body(const mgpuMultiPassArgs& a) {
    constexpr int W = BLOCK / 64;
    constexpr uint32_t WT = 64 * IPT;  // keys of a wave chunk
    constexpr uint32_t TILE = BLOCK * IPT;
    constexpr uint32_t PMAX = MGPU_PACK_MAX_PARTS(W);
    constexpr uint32_t GROUPS = BLOCK / RADIX;  // parts scanned at once
    static_assert(BLOCK % RADIX == 0 && BLOCK <= (int)WT && WT % BLOCK == 0,
                  "a write-out round must lie inside one chunk");
    static_assert(PMAX >= 1, "at least one part per tile");
    constexpr bool JOINT = VMODE == MGPU_VALUES_SINGLE || VMODE == MGPU_VALUES_INDEX;
    __shared__ uint32_t s_stage[TILE];
    __shared__ uint32_t s_vstage[JOINT ? TILE : 1];
    __shared__ uint32_t s_count[W][RADIX];
    // Per (part, digit): the part's digit total, then its first staging
    // position, then (output position - staging position).
    __shared__ uint32_t s_part[PMAX][RADIX];
    __shared__ uint32_t s_prefix[RADIX];  // first part: its digits in earlier tiles
    __shared__ uint32_t s_last[RADIX];    // last part: its digit counts in this tile
    __shared__ uint32_t s_valid[W];       // keys of each chunk
    __shared__ uint32_t s_pfirst[PMAX];   // first chunk of each part
    __shared__ uint32_t s_pend[PMAX];     // end of each part's keys in the stage
    __shared__ uint32_t s_tmp[W];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t shift = a.shift, mask = a.mask;
    const uint32_t width = (uint32_t)__builtin_popcount(mask);  // mask = 2^width - 1
    const uint32_t in = a.in_side, out = a.in_side ^ 1u;

    // XSEG: this XCD's range of tiles (whole arrays), else every tile.
    uint32_t xslot = 0, xfirst = 0, xtiles = a.num_tiles;
    if constexpr (XSEG) {
        uint32_t xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        xslot = xcc & 0xFu;
        if (xslot < a.num_xcds) {
            xfirst = a.xcd_tile_begin[xslot];
            xtiles = a.xcd_tile_begin[xslot + 1] - xfirst;
        } else {
            xtiles = 0;
        }
    }
    for (uint32_t i = tid; i < (uint32_t)(W * RADIX); i += BLOCK) (&s_count[0][0])[i] = 0;

    while (true) {
        if (tid == 0) {
            uint32_t t = xtiles;
            if (!XSEG && a.num_xcds > 1)
                t = take_grouped_tile(a);
            else if (xtiles > 0)
                t = __hip_atomic_fetch_add(a.tile_counter + xslot, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            s_tile = t;
        }
        __syncthreads();
        const uint32_t lt = __builtin_amdgcn_readfirstlane(s_tile);
        if (lt >= xtiles) {
            if (!XSEG && tid == 0) set_error(a.error_flag, 2u);
            break;
        }
        const uint32_t tile = xfirst + lt;  // global tile: its lookback words
        const uint32_t first_arr = a.tile_info[2 * tile];
        const uint32_t info = a.tile_info[2 * tile + 1];
        const uint32_t starts = info & 0xFFFEu, live = (info >> 16) & 0xFFu;
        const bool cont = ((info >> 24) & 1u) != 0;
        const uint32_t nparts = (uint32_t)__builtin_popcount(starts) + 1u;
        if (nparts > PMAX) {  // the host bounds the parts of every tile
            if (tid == 0) set_error(a.error_flag, 6u);
            break;
        }
        // Part of chunk w: the arrays that start at chunks 1..w.
        auto part_of = [&](uint32_t w) -> uint32_t {
            return (uint32_t)__builtin_popcount(starts & ((2u << w) - 1u));
        };

        // This wave's chunk: keys and the single value array, striped.
        // Padding chunks (past `live`) read nothing: every slot is a filler.
        const mgpuSegDesc& sd = a.segs[first_arr + part_of(wave)];
        const uint32_t n = sd.n;
        const bool has = wave < live;
        const uint32_t wbase =
            has ? (tile * W + wave - a.chunk_begin[first_arr + part_of(wave)]) * WT : n;
        uint32_t key[IPT], val[IPT];
        load_striped<IPT>(sd.keys[in], n, wbase, lane, key);
        if constexpr (VMODE == MGPU_VALUES_SINGLE) load_striped<IPT>(sd.vals[0][in], n, wbase, lane, val);
        // Fillers past the array's end get every bit set: the pass's top digit,
        // ranked after every real key of their part.
        const int32_t rem = (int32_t)(n - (wbase + lane));
#pragma unroll
        for (int j = 0; j < IPT; ++j) key[j] = j * 64 < rem ? key[j] : 0xFFFFFFFFu;
        auto digit_of = [&](int j) -> uint32_t { return __builtin_amdgcn_ubfe(key[j], shift, width); };

        // Rank within the wave (atomic adds, served in ascending lane order).
        uint32_t pos[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t d = digit_of(j);
            const uint32_t f = __builtin_amdgcn_readfirstlane(d);
            if (__builtin_amdgcn_ballot_w64(d == f) == ~0ull) {
                uint32_t b = 0;
                if (lane == 0)
                    b = __hip_atomic_fetch_add(&s_count[wave][f], 64u, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                pos[j] = __builtin_amdgcn_readfirstlane(b) + lane;
            } else {
                pos[j] = __hip_atomic_fetch_add(&s_count[wave][d], 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (lane == 0) s_valid[wave] = has ? min(WT, n - wbase) : 0u;
        __syncthreads();

        // Per digit (thread d): offsets of each wave among the tile's keys of
        // d, and each part's total of d (a difference of the offsets at the
        // part's first chunk).  The last part's count (fillers taken off the
        // top digit) is published at once.  Thread RADIX (thread 0 when the
        // block has no more threads than digits): the parts' first chunks
        // and stage ends.
        uint32_t total = 0;
        if (tid < RADIX) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint32_t c = s_count[w][tid];
                s_count[w][tid] = total;
                total += c;
            }
            uint32_t m = starts, before = 0, q = 0;
            while (m) {
                const uint32_t w = (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                const uint32_t b = s_count[w][tid];
                s_part[q][tid] = b - before;
                before = b;
                ++q;
            }
            if (nparts > 1) s_part[q][tid] = total - before;
            uint32_t last = total - before;
            if (tid == mask) {
                const uint32_t lw = starts ? 31u - (uint32_t)__builtin_clz(starts) : 0u;
                for (uint32_t w = lw; w < (uint32_t)W; ++w) last -= WT - s_valid[w];
            }
            s_last[tid] = last;
            lookback_publish<XSEG>(&a.lookback[(size_t)tile * RADIX + tid],
                                   (nparts == 1 && cont ? MGPU_LB_AGGREGATE : MGPU_LB_PREFIX) | last);
        }
        if (tid == (BLOCK > (int)RADIX ? RADIX : 0u)) {
            uint32_t q = 0, w0 = 0, keys = 0;
            for (uint32_t w = 0; w < (uint32_t)W; ++w) {
                if (w > 0 && ((starts >> w) & 1u)) {
                    s_pfirst[q] = w0;
                    s_pend[q] = w0 * WT + keys;
                    ++q;
                    w0 = w;
                    keys = 0;
                }
                keys += s_valid[w];
            }
            s_pfirst[q] = w0;
            s_pend[q] = w0 * WT + keys;
        }

        if (nparts == 1) {
            // One part, as in rank_scatter_body: the tile's digit starts (a
            // block scan of the totals) added to the wave offsets.
            const uint32_t start = block_exclusive_scan<BLOCK>(total, s_tmp, lane, wave);
            if (tid < RADIX) {
                s_part[0][tid] = start;
#pragma unroll
                for (int w = 0; w < W; ++w) s_count[w][tid] += start;
            }
        } else {
            __syncthreads();
            // Each part's digit starts in the stage: its first chunk's
            // position plus the exclusive scan of its digit totals (GROUPS
            // parts a round).
            for (uint32_t pb = 0; pb < nparts; pb += GROUPS) {
                const uint32_t q = pb + tid / RADIX, d = tid % RADIX;
                const uint32_t v = q < nparts ? s_part[q][d] : 0u;
                const uint32_t ex = group_exclusive_scan<BLOCK>(v, s_tmp, lane, wave);
                if (q < nparts) s_part[q][d] = s_pfirst[q] * WT + ex;
            }
            __syncthreads();
            // Wave offsets become stage positions: within its part a wave's
            // keys of d follow the part's earlier waves' (offset at w minus
            // offset at the part's first chunk, then the part's start of d).
            // Waves in descending order, so a part's first-chunk offset is
            // read before it is replaced.
            if (tid < RADIX) {
#pragma unroll
                for (int w = W - 1; w >= 0; --w) {
                    const uint32_t q = part_of((uint32_t)w);
                    const uint32_t fw =
                        31u - (uint32_t)__builtin_clz((starts | 1u) & ((2u << w) - 1u));
                    s_count[w][tid] = s_count[w][tid] - s_count[fw][tid] + s_part[q][tid];
                }
            }
        }
        __syncthreads();

        // Only a tile continuing an array walks back; the first round's words
        // are requested before the staging, which overlaps their latency.
        constexpr uint32_t G = BLOCK / RADIX;
        constexpr uint32_t K = MGPU_LOOKBACK_WINDOW / G > 0 ? MGPU_LOOKBACK_WINDOW / G : 1;
        const uint32_t lb_digit = tid / G, lb_k = tid % G;
        uint32_t lb_words[K];
        if (cont) lookback_read<K, XSEG>(a.lookback, (int32_t)tile - 1, lb_digit, lb_k, lb_words);

#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            pos[j] += s_count[wave][digit_of(j)];
            s_stage[pos[j]] = key[j];
            if constexpr (VMODE == MGPU_VALUES_INDEX) s_vstage[pos[j]] = wbase + lane + j * 64;
            else if constexpr (VMODE == MGPU_VALUES_SINGLE) s_vstage[pos[j]] = val[j];
        }

        // Per (part, digit): output position minus stage position.  With one
        // part the threads that end the walk (or the digits' threads) set it
        // at once.
        const uint32_t* const ds0 =
            a.digit_start + ((size_t)first_arr * MGPU_MAX_PASSES + a.pass) * RADIX;
        if (cont) {
            const uint32_t prefix = lookback_window<G, K, XSEG>(
                a.lookback, tile, lb_digit, lb_k, lane, a.error_flag, a.spin_limit, lb_words);
            if (lb_k == 0) {
                if (nparts == 1) {
                    lookback_publish<XSEG>(&a.lookback[(size_t)tile * RADIX + lb_digit],
                                           MGPU_LB_PREFIX | (prefix + s_last[lb_digit]));
                    s_part[0][lb_digit] = ds0[lb_digit] + prefix - s_part[0][lb_digit];
                } else {
                    s_prefix[lb_digit] = prefix;
                }
            }
        } else if (tid < RADIX) {
            if (nparts == 1) s_part[0][tid] = ds0[tid] - s_part[0][tid];
            else s_prefix[tid] = 0;
        }
        __syncthreads();
        if (nparts > 1) {
            for (uint32_t pb = 0; pb < nparts; pb += GROUPS) {
                const uint32_t q = pb + tid / RADIX, d = tid % RADIX;
                if (q < nparts)
                    s_part[q][d] = ds0[(size_t)q * MGPU_MAX_PASSES * RADIX + d] +
                                   (q == 0 ? s_prefix[d] : 0u) - s_part[q][d];
            }
            __syncthreads();
        }

        // Write out: round i covers stage positions [i*BLOCK, (i+1)*BLOCK),
        // inside chunk i*BLOCK/WT of the stage, hence inside one part; the
        // rounds go in order and switch arrays where a part starts.
        // Positions past the part's keys (fillers) get dest = the array's
        // length, which the buffer stores' range check drops; offsets outside
        // the array from a broken count or lookback are flagged instead.
        uint32_t kout[IPT], dest[IPT];
#pragma unroll
        for (int i = 0; i < IPT; ++i) kout[i] = s_stage[i * BLOCK + tid];
        bool bad = false;
        {
            uint32_t q = 0;
            const mgpuSegDesc* od = &a.segs[first_arr];
            uint32_t on = od->n, pend = s_pend[0];
            __amdgpu_buffer_rsrc_t rk = array_rsrc(od->keys[out], on);
            __amdgpu_buffer_rsrc_t rv = JOINT ? array_rsrc(od->vals[0][out], on) : rk;
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                constexpr uint32_t R = WT / BLOCK;  // rounds per chunk
                const uint32_t c = (uint32_t)i / R;
                if ((uint32_t)i % R == 0 && c > 0 && ((starts >> c) & 1u)) {
                    ++q;
                    od = &a.segs[first_arr + q];
                    on = od->n;
                    pend = s_pend[q];
                    rk = array_rsrc(od->keys[out], on);
                    if constexpr (JOINT) rv = array_rsrc(od->vals[0][out], on);
                }
                const uint32_t sp = (uint32_t)(i * BLOCK) + tid;
                const bool real = c < live && sp < pend;
                const uint32_t dd = s_part[q][__builtin_amdgcn_ubfe(kout[i], shift, width)] + sp;
                const bool ok = real && dd < on;
                bad = bad || (real && !ok);
                dest[i] = ok ? dd : on;
                __builtin_amdgcn_raw_buffer_store_b32(kout[i], rk, (int)(dest[i] * 4u), 0, 0);
                if constexpr (JOINT)
                    __builtin_amdgcn_raw_buffer_store_b32(s_vstage[sp], rv, (int)(dest[i] * 4u), 0, 0);
            }
        }
        if (bad) set_error(a.error_flag, 4u);

        // Two or more value arrays: one after the other through the key stage.
        if constexpr (VMODE == MGPU_VALUES_MULTI) {
            for (uint32_t v = 0; v < a.num_values; ++v) {
                load_striped<IPT>(sd.vals[v][in], n, wbase, lane, val);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < IPT; ++j) s_stage[pos[j]] = val[j];
                __syncthreads();
                uint32_t q = 0;
                const mgpuSegDesc* od = &a.segs[first_arr];
                __amdgpu_buffer_rsrc_t rv = array_rsrc(od->vals[v][out], od->n);
#pragma unroll
                for (int i = 0; i < IPT; ++i) {
                    constexpr uint32_t R = WT / BLOCK;
                    const uint32_t c = (uint32_t)i / R;
                    if ((uint32_t)i % R == 0 && c > 0 && ((starts >> c) & 1u)) {
                        ++q;
                        od = &a.segs[first_arr + q];
                        rv = array_rsrc(od->vals[v][out], od->n);
                    }
                    __builtin_amdgcn_raw_buffer_store_b32(s_stage[i * BLOCK + tid], rv,
                                                          (int)(dest[i] * 4u), 0, 0);
                }
            }
        }

        if constexpr (!XSEG) break;
        __syncthreads();  // every read of this tile's LDS state is done
        for (uint32_t i = tid; i < (uint32_t)(W * RADIX); i += BLOCK) (&s_count[0][0])[i] = 0;
    }  // while

    // The last workgroup to finish checks that every XCD's range was taken
    // to its end.
    if constexpr (XSEG) {
        if (tid == 0) {
            const uint32_t done =
                __hip_atomic_fetch_add(a.finished, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (done + 1 == gridDim.x) {
                for (uint32_t x = 0; x < a.num_xcds; ++x)
                    if (ld_agent(a.tile_counter + x) < a.xcd_tile_begin[x + 1] - a.xcd_tile_begin[x])
                        set_error(a.error_flag, 5u);
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Exported entry points (names looked up by the engine).

extern "C" __global__ void __launch_bounds__(256) mgpu_digit_count(mgpuCountArgs a) {
    digit_count_body<256, 0, false>(a);
}

extern "C" __global__ void __launch_bounds__(1024) mgpu_digit_count_cols(mgpuCountArgs a) {
    digit_count_body<1024, 32, false>(a);
}

// The same over signed or float keys, counted in their unsigned order.
extern "C" __global__ void __launch_bounds__(256) mgpu_digit_count_km(mgpuCountArgs a) {
    digit_count_body<256, 0, true>(a);
}

extern "C" __global__ void __launch_bounds__(1024) mgpu_digit_count_cols_km(mgpuCountArgs a) {
    digit_count_body<1024, 32, true>(a);
}

extern "C" __global__ void __launch_bounds__(256) mgpu_digit_scan(mgpuScanArgs a) {
    digit_scan_body(a);
}

extern "C" __global__ void __launch_bounds__(256) mgpu_iota(mgpuIotaArgs a) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += gridDim.x * 256) a.out[i] = i;
}

// Grid-stride, 16-byte accesses where the pointer allows.
extern "C" __global__ void __launch_bounds__(256) mgpu_key_transform(mgpuKeyTransformArgs a) {
    const uint32_t stride = gridDim.x * 256;
    auto f = [&](uint32_t k) {
        return a.to_radix ? mgpu_key_to_radix(k, a.type) : mgpu_key_from_radix(k, a.type);
    };
    if ((reinterpret_cast<uintptr_t>(a.keys) & 15) == 0) {
        uint4* v = reinterpret_cast<uint4*>(a.keys);
        const uint32_t nv = a.n >> 2;
        for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nv; i += stride) {
            uint4 q = v[i];
            q.x = f(q.x);
            q.y = f(q.y);
            q.z = f(q.z);
            q.w = f(q.w);
            v[i] = q;
        }
        for (uint32_t i = (nv << 2) + blockIdx.x * 256 + threadIdx.x; i < a.n; i += stride)
            a.keys[i] = f(a.keys[i]);
    } else {
        for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += stride) a.keys[i] = f(a.keys[i]);
    }
}

extern "C" __global__ void __launch_bounds__(256) mgpu_multi_count(mgpuMultiCountArgs a) {
    multi_count_body<256>(a);
}

extern "C" __global__ void __launch_bounds__(1024) mgpu_multi_count_cols(mgpuMultiCountArgs a) {
    multi_count_cols_body(a);
}

extern "C" __global__ void __launch_bounds__(256) mgpu_multi_scan(mgpuMultiScanArgs a) {
    multi_scan_body(a);
}

#define MGPU_RANK_SCATTER_MODE(B, I, R, V, MODE)                                             \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_rank_scatter_b##B##_i##I##_r##R##_v##V(mgpuPassArgs a) {                         \
        rank_scatter_body<B, I, R, MODE, false, false, false, false>(a);                      \
    }                                                                                         \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_rank_scatter_b##B##_i##I##_r##R##_x_v##V(mgpuPassArgs a) {                       \
        rank_scatter_body<B, I, R, MODE, false, true, false, false>(a);                       \
    }                                                                                         \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_multi_scatter_b##B##_i##I##_r##R##_v##V(mgpuMultiPassArgs a) {                   \
        rank_scatter_body<B, I, R, MODE, true, false, false, false>(a);                       \
    }                                                                                         \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_multi_scatter_b##B##_i##I##_r##R##_x_v##V(mgpuMultiPassArgs a) {                 \
        rank_scatter_body<B, I, R, MODE, true, true, false, false>(a);                        \
    }
// Early-exit forms (atomic-add ranking only): one chain, and a chain per XCD.
#define MGPU_RANK_SCATTER_EE_MODE(B, I, V, MODE)                                              \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_rank_scatter_b##B##_i##I##_r1_ee_v##V(mgpuPassArgs a) {                          \
        rank_scatter_body<B, I, 1, MODE, false, false, true, false>(a);                       \
    }                                                                                         \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_rank_scatter_b##B##_i##I##_r1_x_ee_v##V(mgpuPassArgs a) {                        \
        rank_scatter_body<B, I, 1, MODE, false, true, true, false>(a);                        \
    }
// Forms that map signed or float keys to their unsigned order as they read
// them, or back as they write them (atomic-add ranking only): the first and
// last passes of a typed sort.
#define MGPU_RANK_SCATTER_KM_MODE(B, I, V, MODE)                                              \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_rank_scatter_b##B##_i##I##_r1_km_v##V(mgpuPassArgs a) {                          \
        rank_scatter_body<B, I, 1, MODE, false, false, false, true>(a);                       \
    }                                                                                         \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_rank_scatter_b##B##_i##I##_r1_x_km_v##V(mgpuPassArgs a) {                        \
        rank_scatter_body<B, I, 1, MODE, false, true, false, true>(a);                        \
    }
#define MGPU_RANK_SCATTER_KM(B, I)                                                            \
    MGPU_RANK_SCATTER_KM_MODE(B, I, 0, MGPU_VALUES_NONE)                                      \
    MGPU_RANK_SCATTER_KM_MODE(B, I, 1, MGPU_VALUES_INDEX)                                     \
    MGPU_RANK_SCATTER_KM_MODE(B, I, 2, MGPU_VALUES_SINGLE)                                    \
    MGPU_RANK_SCATTER_KM_MODE(B, I, 3, MGPU_VALUES_MULTI)
#define MGPU_RANK_SCATTER_EE(B, I)                                                            \
    MGPU_RANK_SCATTER_EE_MODE(B, I, 0, MGPU_VALUES_NONE)                                      \
    MGPU_RANK_SCATTER_EE_MODE(B, I, 1, MGPU_VALUES_INDEX)                                     \
    MGPU_RANK_SCATTER_EE_MODE(B, I, 2, MGPU_VALUES_SINGLE)                                    \
    MGPU_RANK_SCATTER_EE_MODE(B, I, 3, MGPU_VALUES_MULTI)
#define MGPU_RANK_SCATTER(B, I, R)                                                            \
    MGPU_RANK_SCATTER_MODE(B, I, R, 0, MGPU_VALUES_NONE)                                      \
    MGPU_RANK_SCATTER_MODE(B, I, R, 1, MGPU_VALUES_INDEX)                                     \
    MGPU_RANK_SCATTER_MODE(B, I, R, 2, MGPU_VALUES_SINGLE)                                    \
    MGPU_RANK_SCATTER_MODE(B, I, R, 3, MGPU_VALUES_MULTI)

// Packed device-wide passes of a segmented sort (atomic-add ranking), one
// chain across the device (_v) or chains per XCD (_x_v).
#define MGPU_PACKED_MODE(B, I, V, MODE)                                                       \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_multi_packed_b##B##_i##I##_v##V(mgpuMultiPassArgs a) {                           \
        packed_scatter_body<B, I, MODE, false>(a);                                            \
    }                                                                                         \
    extern "C" __global__ void __launch_bounds__(B)                                           \
        mgpu_multi_packed_b##B##_i##I##_x_v##V(mgpuMultiPassArgs a) {                         \
        packed_scatter_body<B, I, MODE, true>(a);                                             \
    }
#define MGPU_PACKED(B, I)                                                                     \
    MGPU_PACKED_MODE(B, I, 0, MGPU_VALUES_NONE)                                               \
    MGPU_PACKED_MODE(B, I, 1, MGPU_VALUES_INDEX)                                              \
    MGPU_PACKED_MODE(B, I, 2, MGPU_VALUES_SINGLE)                                             \
    MGPU_PACKED_MODE(B, I, 3, MGPU_VALUES_MULTI)

// Tile shapes: 1024 x 16 (16384 keys, one workgroup per CU: the fastest from
// 2^21 keys up), 512 x 16 (below that), 256 x 16 (kept for tests and sweeps);
// each with both ranking methods, with one lookback chain across the device
// (_v) or one per XCD segment (_x_v).
MGPU_RANK_SCATTER(256, 16, 0)
MGPU_RANK_SCATTER(512, 16, 0)
MGPU_RANK_SCATTER(1024, 16, 0)
MGPU_RANK_SCATTER(256, 16, 1)
MGPU_RANK_SCATTER(512, 16, 1)
MGPU_RANK_SCATTER(1024, 16, 1)
MGPU_RANK_SCATTER_EE(512, 16)
MGPU_RANK_SCATTER_EE(1024, 16)
MGPU_RANK_SCATTER_KM(512, 16)
MGPU_RANK_SCATTER_KM(1024, 16)
MGPU_PACKED(256, 16)
MGPU_PACKED(512, 16)
MGPU_PACKED(1024, 16)

// Check of the ordering RANK 1 relies on: wave w of one 1024-thread
// workgroup adds 1 per lane, for 16 rounds, to the count of digit
// mgpu_rank_probe_digit(w, round, lane) and writes back the old count it was
// given; the host compares with the counts in ascending lane order.  Digit
// patterns range from all lanes on one count to all on different ones.
extern "C" __global__ void __launch_bounds__(1024) mgpu_rank_probe(uint32_t* out) {
    __shared__ uint32_t cnt[16][RADIX];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (uint32_t i = tid; i < 16 * RADIX; i += 1024) (&cnt[0][0])[i] = 0;
    __syncthreads();
#pragma unroll 1
    for (uint32_t r = 0; r < 16; ++r) {
        const uint32_t d = mgpu_rank_probe_digit(wave, r, lane);
        out[r * 1024 + tid] = __hip_atomic_fetch_add(&cnt[wave][d], 1u, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// Check of what segments per XCD rely on, by one workgroup of 64 lanes per
// grid slot: lane 0 records the XCD it runs on (out[0 .. grid)), and the
// first two workgroups to arrive on each XCD ping-pong `rounds` times through
// a word of that XCD's area with the lookback's nontemporal loads and stores
// (every wait bounded); out[grid + 2x] = rounds completed on XCD x, out[grid +
// 2x + 1] = 1 if a wait gave up.  out: grid + 2 * MGPU_MAX_SEGS words; area:
// zeroed words, MGPU_MAX_SEGS * 128 (the arrival counter and the two flags of
// an XCD on separate cache lines).
extern "C" __global__ void __launch_bounds__(64) mgpu_xcd_probe(uint32_t* out, uint32_t* area,
                                                               uint32_t rounds) {
    if (threadIdx.x != 0) return;
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    xcc &= 0xFu;
    out[blockIdx.x] = xcc;
    uint32_t* base = area + xcc * 128;
    const uint32_t role =
        __hip_atomic_fetch_add(base, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (role > 1) return;
    uint32_t* ping = base + 32;
    uint32_t* pong = base + 64;
    uint32_t r = 1;
    bool failed = false;
    for (; r <= rounds && !failed; ++r) {
        uint32_t* wait_on = role == 0 ? pong : ping;
        if (role == 0) lookback_publish<true>(ping, r);
        uint32_t spins = 0;
        while (true) {
            asm volatile("" ::: "memory");
            if (lookback_word<true>(wait_on, 0, 0) == r) break;
            if (++spins > (1u << 20)) {
                failed = true;
                break;
            }
        }
        if (role == 1 && !failed) lookback_publish<true>(pong, r);
    }
    if (role == 0) out[gridDim.x + 2 * xcc] = failed ? r - 2 : rounds;
    if (failed) out[gridDim.x + 2 * xcc + 1] = 1;
}
