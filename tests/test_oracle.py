"""Pins the CPU oracle (oracle/oracle.cpp) before anything is checked against it.

Each check is independent of the oracle's own formula: numpy's stable argsort
(a library routine), brute force over every permutation of small inputs, the
invariants that define a stable sort, the published SplitMix64 output, and a
planted mistake that the invariant checker must catch.
"""
import itertools

import numpy as np
import pytest

from conftest import bitfield, check_stable_sort
from mgpu_old_amd import workloads


def _np_stable(keys, first_bit, end_bit):
    return np.argsort(bitfield(keys, first_bit, end_bit), kind="stable")


def test_splitmix64_published_value(oracle):
    # First SplitMix64 output for seed 0 (state 0): 0xE220A8397B1DCDAF.
    assert oracle.L.oracle_splitmix64_at(0, 0) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("seed", [0, 1, 12345, 2**63 + 7])
def test_workload_generators_agree(oracle, seed):
    n = 5000
    assert np.array_equal(workloads.uniform_keys(n, seed), oracle.fill_keys(n, seed, 0))
    assert np.array_equal(workloads.and3_keys(n, seed), oracle.fill_keys(n, seed, 1))
    assert np.array_equal(workloads.and3_keys(n, seed, 24), oracle.fill_keys(n, seed, 1, 24))
    w = workloads.splitmix64(seed, 10, 3)
    assert [int(x) for x in w] == [oracle.L.oracle_splitmix64_at(seed, i) for i in (10, 11, 12)]


@pytest.mark.parametrize("seed", [0, 77, 2**63 + 11])
def test_torch_generators_agree(oracle, seed):
    """The on-device (torch int64) generators give the oracle's words."""
    torch = pytest.importorskip("torch")
    n = 20000
    a = workloads.uniform_keys_torch(n, seed, "cpu", torch=torch).numpy().view(np.uint32)
    assert np.array_equal(a, oracle.fill_keys(n, seed, 0))
    b = workloads.and3_keys_torch(n, seed, "cpu", 24, torch=torch).numpy().view(np.uint32)
    assert np.array_equal(b, oracle.fill_keys(n, seed, 1, 24))


def test_and3_keys_bit_density():
    k = workloads.and3_keys(1 << 16, 99)
    bits = np.unpackbits(k.view(np.uint8)).mean()
    assert abs(bits - 1 / 8) < 0.005


@pytest.mark.parametrize("n", [0, 1, 2, 7, 100, 2047, 2048, 2049, 10000])
@pytest.mark.parametrize("bits", [(0, 32), (0, 1), (31, 32), (7, 25), (0, 24), (5, 6), (0, 16)])
def test_stable_sort_matches_numpy(oracle, n, bits):
    rng = np.random.default_rng(n * 131 + bits[0] * 7 + bits[1])
    # Few distinct keys as well as many, to exercise ties.
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        keys[: n // 2] = rng.integers(0, 8, n // 2).astype(np.uint32) << 7
    vals = [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    fb, eb = bits
    perm = _np_stable(keys, fb, eb)
    ok, ov = oracle.stable_sort(keys, vals, first_bit=fb, end_bit=eb)
    assert np.array_equal(ok, keys[perm])
    for v, o in zip(vals, ov):
        assert np.array_equal(o, v[perm])
    ok2, (idx,) = oracle.stable_sort(keys, value_count=-1, first_bit=fb, end_bit=eb)
    assert np.array_equal(ok2, keys[perm])
    assert np.array_equal(idx, perm.astype(np.uint32))
    check_stable_sort(keys, ok2, idx, fb, eb)


@pytest.mark.parametrize("multiset", [
    [3, 1, 2, 1, 3, 0], [5, 5, 5, 5], [0xFFFFFFFF, 0, 0xFFFFFFFF, 1, 0], [9, 8, 7, 6, 5, 4],
])
def test_stable_sort_brute_force(oracle, multiset):
    """The stable sort is the unique permutation whose keys are nondecreasing
    and whose ties keep source order: find it by trying every permutation."""
    keys = np.array(multiset, dtype=np.uint32)
    n = keys.size
    found = None
    for p in itertools.permutations(range(n)):
        k = keys[list(p)]
        if all(k[i] < k[i + 1] or (k[i] == k[i + 1] and p[i] < p[i + 1]) for i in range(n - 1)):
            assert found is None
            found = p
    ok, (idx,) = oracle.stable_sort(keys, value_count=-1)
    assert tuple(int(x) for x in idx) == found
    assert np.array_equal(ok, keys[list(found)])


def test_checker_catches_planted_instability(oracle):
    keys = np.array([4, 2, 4, 1, 2, 4], dtype=np.uint32)
    ok, (idx,) = oracle.stable_sort(keys, value_count=-1)
    check_stable_sort(keys, ok, idx, 0, 32)
    bad = idx.copy()
    bad[[3, 4]] = bad[[4, 3]]  # swap two equal keys' source indices
    with pytest.raises(AssertionError, match="unstable"):
        check_stable_sort(keys, ok, bad, 0, 32)
    bad_keys = ok.copy()
    bad_keys[[0, 1]] = bad_keys[[1, 0]]
    with pytest.raises(AssertionError):
        check_stable_sort(keys, bad_keys, idx, 0, 32)


@pytest.mark.parametrize("digit_bits", [1, 3, 4, 6, 8, 11])
@pytest.mark.parametrize("bits", [(0, 32), (7, 25), (3, 4)])
def test_sort1_sort2_match_stable_sort(oracle, digit_bits, bits):
    rng = np.random.default_rng(digit_bits * 100 + bits[1])
    n = 6000
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    keys[::3] &= 0xFFF
    vals = [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) for _ in range(2)]
    fb, eb = bits
    ek, ev = oracle.stable_sort(keys, vals, first_bit=fb, end_bit=eb)
    k1, v1 = oracle.sort1(keys, vals, first_bit=fb, end_bit=eb, digit_bits=digit_bits)
    assert np.array_equal(k1, ek) and all(np.array_equal(a, b) for a, b in zip(v1, ev))
    for block, threads in [(2048, 1), (64, 4), (1000, 3)]:
        k2, v2 = oracle.sort2(keys, vals, first_bit=fb, end_bit=eb, digit_bits=digit_bits,
                              block_size=block, threads=threads)
        assert np.array_equal(k2, ek) and all(np.array_equal(a, b) for a, b in zip(v2, ev))
    k3, (i3,) = oracle.sort2(keys, value_count=-1, first_bit=fb, end_bit=eb,
                             digit_bits=digit_bits, block_size=512, threads=2)
    _, (ei,) = oracle.stable_sort(keys, value_count=-1, first_bit=fb, end_bit=eb)
    assert np.array_equal(k3, ek) and np.array_equal(i3, ei)


def test_block_counts_and_offsets(oracle):
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 2**32, 10000, dtype=np.uint64).astype(np.uint32)
    for bit, nb, block in [(0, 8, 4096), (24, 8, 1000), (7, 5, 2048)]:
        counts = oracle.block_counts(keys, bit, nb, block)
        d = bitfield(keys, bit, bit + nb).astype(np.int64)
        for b in range(counts.shape[0]):
            ref = np.bincount(d[b * block:(b + 1) * block], minlength=1 << nb)
            assert np.array_equal(counts[b], ref)
        off = oracle.digit_offsets(counts)
        # Scattering every block's keys at its offsets reproduces one stable pass.
        out = np.empty_like(keys)
        cur = off.copy()
        for i, k in enumerate(keys):
            b, dd = i // block, d[i]
            out[cur[b, dd]] = k
            cur[b, dd] += 1
        assert np.array_equal(out, keys[np.argsort(d, kind="stable")])


def test_oracle_rejects_bad_ranges(oracle):
    keys = np.arange(10, dtype=np.uint32)
    with pytest.raises(ValueError):
        oracle.stable_sort(keys, first_bit=5, end_bit=4)
    with pytest.raises(ValueError):
        oracle.stable_sort(keys, first_bit=0, end_bit=33)
