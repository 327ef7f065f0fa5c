"""The inputs behind the golden fixtures in tests/golden/ (shared by the
script that writes them and the tests that read them).  Every input is a pure
function of (kind, n, seed) through the SplitMix64 generators of
mgpu_old_amd.workloads, so the fixtures store digests and permutations, not
inputs."""
import hashlib

import numpy as np

from mgpu_old_amd import workloads

# name, n, key kind, seed, value count (-1: index generation, "iota": one
# value array 0..n-1), first bit, end bit.  Sizes and bit ranges follow the
# special cases the reference's code handles: empty and single inputs, the
# 2048-key block edges (mgpusort.cpp:663,696), partial ranges such as [7,25)
# with two value streams (simpletest2.cpp:62-71,104-105), index generation,
# 1..7-bit and top-bit ranges, six value arrays.
SMALL = [
    ("n0_kv", 0, "uniform", 11, 1, 0, 32),
    ("n1_k", 1, "uniform", 12, 0, 0, 32),
    ("n2_equal_idx", 2, "equal", 13, -1, 0, 32),
    ("n100_few_v2_2bit", 100, "few", 14, 2, 0, 2),
    ("n2047_kv", 2047, "uniform", 15, 1, 0, 32),
    ("n2048_idx", 2048, "uniform", 16, -1, 0, 32),
    ("n2049_and3_kv", 2049, "and3", 17, 1, 0, 32),
    ("n2049_v2_bits7_25", 2049, "uniform", 18, 2, 7, 25),
    ("n4096_v6", 4096, "uniform", 19, 6, 0, 32),
    ("n4096_sorted_idx", 4096, "sorted", 20, -1, 0, 32),
    ("n4096_reverse_kv", 4096, "reverse", 21, 1, 0, 32),
    ("n4096_k_1bit", 4096, "uniform", 22, 0, 0, 1),
    ("n3000_kv_6bit", 3000, "uniform", 23, 1, 0, 6),
    ("n3001_kv_7bit", 3001, "uniform", 24, 1, 0, 7),
    ("n4095_idx_top_bit", 4095, "uniform", 25, -1, 31, 32),
    ("n4096_equal_k", 4096, "equal", 26, 0, 0, 32),
    ("n1000_and3_kv_bits3_10", 1000, "and3", 27, 1, 3, 10),
    ("n4096_kv_bits8_32", 4096, "uniform", 28, 1, 8, 32),
]

# The benchmarked workloads at full size.
LARGE = [
    ("2^20 uniform keys, [0,32)", 1 << 20, "uniform", 1, 0, 0, 32),
    ("2^25 uniform keys + value 0..N-1, [0,32)", 1 << 25, "uniform", 2, "iota", 0, 32),
    ("2^28 AND3 keys, index generation, [0,32)", 1 << 28, "and3", 3, -1, 0, 32),
    ("2^28 AND3 keys & 0xFFFFFF, index generation, [0,24)", 1 << 28, "and3_24", 3, -1, 0, 24),
    ("2^28 uniform keys + value 0..N-1, [0,32)", 1 << 28, "uniform", 2028, "iota", 0, 32),
]


def keys(kind, n, seed):
    if kind == "uniform":
        return workloads.uniform_keys(n, seed)
    if kind == "and3":
        return workloads.and3_keys(n, seed)
    if kind == "and3_24":
        return workloads.and3_keys(n, seed, mask_bits=24)
    if kind == "equal":
        return np.full(n, 0x5A5A5A5A, dtype=np.uint32)
    if kind == "few":
        return workloads.uniform_keys(n, seed) & np.uint32(3)
    if kind == "sorted":
        return np.sort(workloads.uniform_keys(n, seed))
    if kind == "reverse":
        return np.sort(workloads.uniform_keys(n, seed))[::-1].copy()
    raise ValueError(kind)


def keys_torch(kind, n, seed, torch):
    """keys(kind, n, seed) generated on the GPU (int32 bits), for the full-size
    kinds."""
    if kind == "uniform":
        return workloads.uniform_keys_torch(n, seed, "cuda", torch=torch)
    if kind == "and3":
        return workloads.and3_keys_torch(n, seed, "cuda", torch=torch)
    if kind == "and3_24":
        return workloads.and3_keys_torch(n, seed, "cuda", mask_bits=24, torch=torch)
    raise ValueError(kind)


def values(n, seed, value_count):
    """The value arrays of a case: value j = uniform words of seed 1000*seed+j."""
    if value_count == "iota":
        return [np.arange(n, dtype=np.uint32)]
    return [workloads.uniform_keys(n, 1000 * seed + j) for j in range(max(0, value_count))]


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint32).view(np.uint8)).hexdigest()
