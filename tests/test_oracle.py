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


def _special_floats():
    f = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.finfo(np.float32).max,
                  -np.finfo(np.float32).max, np.finfo(np.float32).tiny, -np.finfo(np.float32).tiny,
                  1e-45, -1e-45, 3.0, -3.0], dtype=np.float32)
    return f.view(np.uint32)


def test_key_transform_int32_order(oracle):
    """Unsigned order of the mapped keys is numpy's int32 order (a stable
    argsort by the mapped keys equals a stable argsort by the int32 values)."""
    rng = np.random.default_rng(11)
    v = rng.integers(-2**31, 2**31, 50000, dtype=np.int64).astype(np.int32)
    v[:6] = [-2**31, 2**31 - 1, -1, 0, 1, -2**31]
    u = oracle.key_transform(v.view(np.uint32), 1, True)
    assert np.array_equal(np.argsort(u, kind="stable"), np.argsort(v, kind="stable"))
    assert np.array_equal(oracle.key_transform(u, 1, False), v.view(np.uint32))


def test_key_transform_float32_order(oracle):
    """Unsigned order of the mapped floats is numeric order, with -0 below +0;
    NaNs sort at the ends (negative-signed below -inf, positive above +inf)."""
    rng = np.random.default_rng(12)
    f = np.concatenate([rng.standard_normal(30000).astype(np.float32) * 1e3,
                        rng.integers(0, 2**32, 30000, dtype=np.uint64).astype(np.uint32)
                        .view(np.float32), _special_floats().view(np.float32)])
    finite_or_inf = ~np.isnan(f)
    g = f[finite_or_inf]
    u = oracle.key_transform(g.view(np.uint32), 2, True)
    s = g[np.argsort(u, kind="stable")]
    assert (np.diff(s.astype(np.float64)) >= 0).all()
    # Among zeros every -0 comes before every +0.
    z = s[s == 0]
    signs = np.signbit(z)
    assert not (signs[1:] & ~signs[:-1]).any()
    # Stable among equal values: numpy's stable order after separating -0.
    key = g.astype(np.float64) - np.where((g == 0) & np.signbit(g), 1e-300, 0.0)
    assert np.array_equal(np.argsort(u, kind="stable"), np.argsort(key, kind="stable"))
    nan_pos = oracle.key_transform(np.array([0x7FC00000], dtype=np.uint32), 2, True)[0]
    nan_neg = oracle.key_transform(np.array([0xFFC00000], dtype=np.uint32), 2, True)[0]
    inf_pos = oracle.key_transform(np.array([0x7F800000], dtype=np.uint32), 2, True)[0]
    inf_neg = oracle.key_transform(np.array([0xFF800000], dtype=np.uint32), 2, True)[0]
    assert nan_neg < inf_neg < inf_pos < nan_pos


@pytest.mark.parametrize("key_type", [1, 2])
def test_key_transform_round_trip_and_bijection(oracle, key_type):
    rng = np.random.default_rng(13 + key_type)
    bits = np.concatenate([rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32),
                           np.arange(0, 1 << 16, dtype=np.uint32),
                           np.arange(2**32 - (1 << 16), 2**32, dtype=np.uint64).astype(np.uint32),
                           np.arange(2**31 - (1 << 15), 2**31 + (1 << 15), dtype=np.uint64)
                           .astype(np.uint32), _special_floats()])
    u = oracle.key_transform(bits, key_type, True)
    assert np.array_equal(oracle.key_transform(u, key_type, False), bits)
    assert np.unique(u).size == np.unique(bits).size


def test_key_transform_catches_a_planted_mistake(oracle):
    """The float order check fails for a mapping that forgets to flip the
    magnitude of negative values (a plausible slip)."""
    g = np.array([-2.0, -1.0, 0.5, 3.0], dtype=np.float32)
    wrong = g.view(np.uint32) ^ np.uint32(0x80000000)
    s = g[np.argsort(wrong, kind="stable")]
    assert not (np.diff(s.astype(np.float64)) >= 0).all()
    right = oracle.key_transform(g.view(np.uint32), 2, True)
    assert (np.diff(g[np.argsort(right, kind="stable")].astype(np.float64)) >= 0).all()
