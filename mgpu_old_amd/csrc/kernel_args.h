// Argument blocks of the radix-sort code objects.  Each kernel takes one of
// these structs by value; the host engine fills the same struct and passes its
// address as the single entry of hipModuleLaunchKernel's kernelParams, so this
// header is the whole host/device contract.
#pragma once
#include <stdint.h>

#define MGPU_RADIX_BITS 8
#define MGPU_RADIX 256
#define MGPU_MAX_PASSES 4
#define MGPU_MAX_VALUES 6

// Lookback word: bits 31..30 = state, bits 29..0 = a count.
#define MGPU_LB_AGGREGATE 0x40000000u  // count = this tile's digit total
#define MGPU_LB_PREFIX 0x80000000u     // count = inclusive prefix over tiles
#define MGPU_LB_STATE 0xC0000000u
#define MGPU_LB_VALUE 0x3FFFFFFFu

// Value handling of one digit pass.
#define MGPU_VALUES_NONE 0    // keys only
#define MGPU_VALUES_INDEX 1   // write source indices to vals_out[0]
#define MGPU_VALUES_SINGLE 2  // carry vals_in[0]
#define MGPU_VALUES_MULTI 3   // carry vals_in[0 .. num_values)

// Digit that lane `lane` of wave `wave` uses in round `round` of
// mgpu_rank_probe (host and device share it): per wave a different pattern,
// from every lane on one digit to every lane on its own.
#if defined(__HIPCC__)
#define MGPU_HOST_DEVICE __host__ __device__
#else
#define MGPU_HOST_DEVICE
#endif
static inline MGPU_HOST_DEVICE uint32_t mgpu_rank_probe_digit(uint32_t wave, uint32_t round,
                                                               uint32_t lane) {
    uint32_t h = wave * 0x9E3779B9u + round * 0x85EBCA6Bu + lane * 0xC2B2AE35u + 1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    if (wave <= 8) return h & ((1u << wave) - 1u);  // 1, 2, 4, ... 256 distinct digits
    switch (wave) {
        case 9: return (h & (h >> 8) & (h >> 16)) & 255u;  // skewed toward 0
        case 10: return lane >> 3;                         // runs of 8 lanes
        case 11: return 63u - lane;                        // all distinct, descending
        case 12: return lane & 1u;                         // alternating
        case 13: return (lane * 37u) & 255u;               // all distinct, scattered
        case 14: return h & 3u;
        default: return 255u - (h & 1u);
    }
}

// A device-wide sort splits its tiles into num_segs contiguous segments, one
// per XCD (num_segs = 1: one segment, one lookback chain across the device).
// Segment x holds tiles [seg_first_tile(x), seg_first_tile(x + 1)) of
// num_tiles; the workgroups of XCD x sort it with a lookback chain of its own,
// which stays in that XCD's L2.
#define MGPU_MAX_SEGS 16
static inline MGPU_HOST_DEVICE uint32_t mgpu_seg_first_tile(uint32_t x, uint32_t num_tiles,
                                                             uint32_t num_segs) {
    return (uint32_t)(((uint64_t)x * num_tiles) / num_segs);
}

// digit_count: histograms of every pass's digit over all keys in one read,
// one set per segment.  Workgroup b counts chunk (b % chunks_per_seg) of
// segment (b / chunks_per_seg).
struct mgpuCountArgs {
    const uint32_t* keys;
    uint32_t* hist;          // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX], zero on entry
    uint32_t* clear;         // words zeroed on the way (the lookback words)
    uint64_t clear_words;    // multiple of 4
    uint32_t n;
    uint32_t chunk;          // keys per workgroup, multiple of 1024
    uint32_t num_passes;
    uint32_t vec_ok;         // keys pointer is 16-byte aligned
    uint32_t num_segs;
    uint32_t chunks_per_seg;
    uint32_t tile;           // keys per tile (segment bounds)
    uint32_t num_tiles;
    uint32_t shift[MGPU_MAX_PASSES];
    uint32_t mask[MGPU_MAX_PASSES];
    // Sortedness check (early exit): counts adjacent pairs whose bit field
    // (k >> check_shift) & check_mask decreases into unsorted[0], and those
    // whose digit of the first pass (shift[0], mask[0]) decreases into
    // unsorted[1].
    uint32_t* unsorted;      // null: no check
    uint32_t check_shift;
    uint32_t check_mask;
    // Signed or float keys are counted (and checked) as mgpu_key_map(k,
    // map_lo, map_hi), their unsigned order; (0, 0) counts the keys as stored.
    uint32_t map_lo;
    uint32_t map_hi;
};

// digit_scan: per pass, where each (segment, digit) run starts in the output
// (digits in order, segments in order within a digit); puts the histograms
// and the counters back to zero for the next sort.
struct mgpuScanArgs {
    uint32_t* hist;           // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX] -> zeroed
    uint32_t* digit_start;    // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX]
    uint32_t* tile_counters;  // [MGPU_LANE_COUNTERS] -> zeroed
    uint32_t num_passes;
    uint32_t num_segs;
    uint32_t* seg_count;      // per-segment counts the passes fill in -> zeroed
    uint32_t seg_count_words;
    uint32_t pad_;
    // Early exit (null: none): host-visible words the scan fills for the
    // host's decision: [0] descents of the bit field, [1] descents of pass
    // 0's digit (both from digit_count), [2 + p] the count of pass p's most
    // frequent digit.
    uint32_t* report;
};

// Counter words of a lane: pass p's tile counter of segment x, the
// workgroups of pass p that have finished, the early-exit descent count.
#define MGPU_TILE_COUNTER(p, x) ((p) * MGPU_MAX_SEGS + (x))
#define MGPU_FINISHED_SLOT (MGPU_MAX_PASSES * MGPU_MAX_SEGS)
// Early exit: adjacent pairs of the input whose sort bit field decreases
// (UNSORTED_SLOT), and whose digit of the first pass decreases
// (UNSORTED_SLOT + 1).
#define MGPU_UNSORTED_SLOT (MGPU_FINISHED_SLOT + MGPU_MAX_PASSES)
// Most frequent digit of each pass (segments per XCD only; digit_scan).
#define MGPU_MODE_SLOT (MGPU_UNSORTED_SLOT + 8)
#define MGPU_LANE_COUNTERS (MGPU_MODE_SLOT + MGPU_MAX_PASSES)

// iota: out[i] = i for i < n (index output of a sort that needs no pass).
struct mgpuIotaArgs {
    uint32_t* out;
    uint32_t n;
    uint32_t pad_;
};

// key_transform: int32 or float32 keys to and from the unsigned order the
// sort uses (in place).
#define MGPU_KEYS_INT32 1
#define MGPU_KEYS_FLOAT32 2
struct mgpuKeyTransformArgs {
    uint32_t* keys;
    uint32_t n;
    uint32_t type;       // MGPU_KEYS_INT32 or MGPU_KEYS_FLOAT32
    uint32_t to_radix;   // 1: to the sort's order, 0: back
    uint32_t pad_;
};

// The bijections of key_transform, shared with host code.  int32: flip the
// sign bit.  float32 (the mapping of the reference's FloatToUint /
// UintToFloat, common.cu:159-185): non-negative values get the sign bit set,
// negative ones have every bit flipped, so unsigned order is numeric order
// with -0 just below +0 (NaNs by their bits, at both ends).
static inline MGPU_HOST_DEVICE uint32_t mgpu_key_to_radix(uint32_t k, uint32_t type) {
    if (type == MGPU_KEYS_INT32) return k ^ 0x80000000u;
    return (k & 0x80000000u) ? ~k : (k | 0x80000000u);
}
static inline MGPU_HOST_DEVICE uint32_t mgpu_key_from_radix(uint32_t u, uint32_t type) {
    if (type == MGPU_KEYS_INT32) return u ^ 0x80000000u;
    return (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
}

// Both bijections (and the identity) in one form, for kernels that take the
// direction and type as two words: k ^ (top bit of k set ? hi : lo).
//   int32, either way:      lo = hi = 0x80000000
//   float32 to the order:   lo = 0x80000000, hi = 0xFFFFFFFF
//   float32 back:           lo = 0xFFFFFFFF, hi = 0x80000000
//   unsigned (identity):    lo = hi = 0
static inline MGPU_HOST_DEVICE uint32_t mgpu_key_map(uint32_t k, uint32_t lo, uint32_t hi) {
    return k ^ ((k & 0x80000000u) ? hi : lo);
}

// ---------------------------------------------------------------------------
// Segmented sort: many independent arrays in one persistent launch.

#define MGPU_SEG_MAX_PASSES 8

// One segment: its key and value buffers on both sides of the ping-pong.
// Side 0 holds the input; the result goes to side `result_side` of the
// launch (the side sortArray would leave it on).
struct mgpuSegDesc {
    uint32_t* keys[2];
    uint32_t* vals[MGPU_MAX_VALUES][2];
    uint32_t n;
    uint32_t pad_;
};

struct mgpuSegArgs {
    const mgpuSegDesc* segs;
    const uint32_t* big_list;    // segments for the workgroup routines, largest first
    const uint32_t* small_list;  // segments for the one-wave routine
    uint32_t* counters;          // [0] big cursor, [2] finished workgroups ([1],
                                 // [3] unused): zero on entry, left zero
    uint32_t* error_flag;
    uint32_t num_big;
    uint32_t num_small;
    uint32_t num_values;   // value arrays carried (index generation counts 1)
    uint32_t index_mode;   // generate source indices into vals[0]
    uint32_t result_side;  // 0 or 1
    // Passes over bits [first_bit, first_bit + num_bits): num_passes digits,
    // the narrower ones first, as sortPlanPasses splits the field (width
    // num_bits / num_passes, one more for the last num_bits % num_passes).
    // Derived in the kernel rather than read from per-pass arrays: indexing
    // an argument array by the pass put the whole argument block in scratch.
    uint32_t num_passes;   // <= MGPU_SEG_MAX_PASSES
    uint32_t first_bit;
    uint32_t num_bits;
    // 1: rank by LDS atomic adds, which the engine checked are served in
    // ascending lane order; 0: by match masks.
    uint32_t rank1;
};

// ---------------------------------------------------------------------------
// Device-wide passes over several arrays at once (large segments): the tiles
// of all arrays form one index space, each array's lookback chain starting at
// its own first tile.

// digit_count over chunks of the arrays: chunk c = keys [begin, end) of
// array `seg`, read from side 0.
struct mgpuMultiCountArgs {
    const mgpuSegDesc* segs;
    const uint32_t* chunks;  // [num_chunks][4]: seg, begin, end, unused
    uint32_t* hist;          // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX], zero on entry
    uint32_t* clear;         // lookback words to zero
    uint64_t clear_words;    // multiple of 4
    uint32_t num_chunks;
    uint32_t num_passes;
    uint32_t shift[MGPU_MAX_PASSES];
    uint32_t mask[MGPU_MAX_PASSES];
    // Lane-column kernel only: every chunk is a whole array (direct != 0),
    // so the kernel writes the digit starts and zeroes the tile counters
    // itself, in place of multi_scan; hist is not used.
    uint32_t* digit_start;    // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX]
    uint32_t* tile_counters;  // [MGPU_LANE_COUNTERS]
    uint32_t direct;
    uint32_t pad_;
};

// digit_scan, one workgroup per array; zeroes hist and the tile counters.
struct mgpuMultiScanArgs {
    uint32_t* hist;
    uint32_t* digit_start;    // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX]
    uint32_t* tile_counters;  // [MGPU_LANE_COUNTERS]
    uint32_t num_passes;
    uint32_t pad_;
};

// rank_scatter over the tiles of all arrays.
struct mgpuMultiPassArgs {
    const mgpuSegDesc* segs;
    const uint32_t* tile_begin;   // [num_segs + 1]: first global tile of each array
    const uint32_t* tile_seg;     // [num_tiles]: the array of each global tile
    const uint32_t* digit_start;  // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX]
    uint32_t* lookback;           // [num_tiles][MGPU_RADIX] for this pass, zero on entry
    uint32_t* tile_counter;       // [MGPU_MAX_SEGS] (one per XCD), zero on entry
    uint32_t* error_flag;
    uint32_t num_segs;
    uint32_t num_tiles;
    uint32_t pass;                // index into digit_start
    uint32_t in_side;             // buffers read: side in_side, written: in_side ^ 1
    uint32_t shift;
    uint32_t mask;
    uint32_t num_values;
    uint32_t spin_limit;          // polls of a lookback word before the pass gives up
    // Chains per XCD (kernels built for it only): the arrays are grouped by
    // XCD, XCD x holding global tiles [xcd_tile_begin[x], xcd_tile_begin[x+1])
    // of whole arrays, whose lookback chains its workgroups alone read and
    // write; finished counts the workgroups done (zero on entry).  The
    // kernels of one tile per workgroup read the same arrays, with
    // num_xcds > 1, as groups of arrays whose tiles come from a counter of
    // their own (tile_counter[g]).
    const uint32_t* xcd_tile_begin;  // [num_xcds + 1]
    uint32_t* finished;
    uint32_t num_xcds;
    uint32_t pad_;
    // Packed tiles (mgpu_multi_packed_* kernels only): the arrays are cut into
    // wave chunks of 64 * IPT keys, laid out back to back (an array's last
    // chunk may be partial), and a tile is BLOCK / 64 consecutive chunks, so
    // one tile can hold the end of one array, whole small arrays and the
    // start of another ("parts", at most MGPU_PACK_MAX_PARTS(W) per tile).
    const uint32_t* chunk_begin;  // [num_segs]: first chunk of each array
    // [num_tiles][2]: word 0 = the array of the tile's first chunk; word 1 =
    // bits 0-15: bit w set when chunk w starts a new array (bit 0 clear),
    // bits 16-23: chunks holding keys (the rest are padding), bit 24: the
    // first chunk continues an array begun in an earlier tile.
    const uint32_t* tile_info;
};

// Parts (arrays) per packed tile of W waves.
#define MGPU_PACK_MAX_PARTS(W) ((W) / 2)

// rank_scatter: one stable digit pass (rank in the tile, decoupled lookback
// across tiles, LDS-staged scatter).  With num_segs > 1 the workgroups of XCD
// x take the tiles of segment x (kernels built for it only).
struct mgpuPassArgs {
    const uint32_t* keys_in;
    uint32_t* keys_out;
    const uint32_t* vals_in[MGPU_MAX_VALUES];
    uint32_t* vals_out[MGPU_MAX_VALUES];
    const uint32_t* digit_start;  // [num_segs][MGPU_MAX_PASSES][MGPU_RADIX] + this pass
    uint32_t* lookback;           // [num_tiles][MGPU_RADIX], zero on entry
    uint32_t* tile_counter;       // this pass's [MGPU_MAX_SEGS] tile counters, zero on entry
    uint32_t* finished;           // this pass's finished-workgroup count, zero on entry
    uint32_t* error_flag;         // host-visible word: set when a self-check fails
    uint32_t n;
    uint32_t num_tiles;
    uint32_t shift;
    uint32_t mask;
    uint32_t num_values;          // MGPU_VALUES_MULTI only
    uint32_t num_segs;
    // Segments per XCD only.  A pass's segment starts are the device-wide
    // digit starts plus the keys of the digit in earlier segments, counted
    // over the keys in this pass's input order: from digit_count for the
    // pass that reads the input (seg_count_in null: digit_start already holds
    // per-segment starts), else from the pass before, which counts the next
    // pass's digit of every key it writes by destination segment into
    // seg_count_out ([MGPU_XSEG_COUNT_SEGS][MGPU_RADIX], zero on entry).
    const uint32_t* seg_count_in;
    uint32_t* seg_count_out;
    const uint32_t* next_mode;  // most frequent digit of the next pass (with seg_count_out)
    uint32_t next_shift;
    uint32_t next_mask;
    // Polls of one lookback word before the pass gives up and sets
    // *error_flag = 1 (0: every lookback gives up at once, which only tests
    // of the error path ask for).
    uint32_t spin_limit;
    // Signed or float keys (MGPU_KEYS_INT32 / MGPU_KEYS_FLOAT32; 0: unsigned,
    // left as they are): bits 0-7, the type of the keys the pass reads, which
    // it maps to their unsigned order (the sort's first pass); bits 8-15, the
    // type the keys it writes go back to (its last pass, which counts nothing
    // for a next one).  One word: the kernels run short of scalar registers.
    uint32_t key_types;
};

// Segments per XCD with counting for the next pass: at most this many.
#define MGPU_XSEG_COUNT_SEGS 8
