"""Parity of the HIP sort with the oracle, through the C ABI (needs a gfx950 GPU).

Bit-exact: keys, every value array and generated indices must equal the
oracle's stable sort by the bit field [firstBit, endBit) on the same seeded
inputs.  Full-size workloads are checked by the size-independent properties
that define the stable sort (check_stable_sort), smaller ones element by
element against the oracle.
"""
import ctypes

import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from conftest import check_stable_sort
from gpu_helpers import GpuSort, copy_d2d, make_engine, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    eng = make_engine()
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def xcd_engine():
    """Segments per XCD at every size where each XCD gets a tile."""
    eng = make_engine({"MGPUSORT_XCD": "1"})
    yield eng
    eng.close()


def _check_vs_oracle(oracle, keys, g, values=(), value_count=None, fb=0, eb=32):
    vc = len(values) if value_count is None else value_count
    rk, rv = oracle.stable_sort(keys, list(values), value_count=vc, first_bit=fb, end_bit=eb)
    out_k = g.keys()
    assert np.array_equal(out_k, rk), "keys differ from the oracle"
    for j, (o, r) in enumerate(zip(g.values(), rv)):
        assert np.array_equal(o, r), f"value array {j} differs from the oracle"


def test_config1_keys_only_2p20(engine, oracle):
    """2^20 uniform u32 keys, [0, 32), keys only."""
    keys = workloads.uniform_keys(1 << 20, seed=1)
    g = GpuSort(keys).run(engine)
    _check_vs_oracle(oracle, keys, g)
    assert g.data.parity == 0


def test_config2_key_value_2p25(engine, oracle):
    """2^25 uniform keys + one value array holding 0..N-1."""
    n = 1 << 25
    keys = workloads.uniform_keys(n, seed=2)
    vals = np.arange(n, dtype=np.uint32)
    g = GpuSort(keys, [vals]).run(engine)
    out_k, (out_v,) = g.keys(), g.values()
    check_stable_sort(keys, out_k, out_v, 0, 32)
    rk, (rv,) = oracle.stable_sort(keys, [vals])
    assert np.array_equal(out_k, rk) and np.array_equal(out_v, rv)


@pytest.mark.parametrize("mask_bits,end_bit", [(32, 32), (24, 24)])
def test_config3_low_entropy_index_2p27(engine, mask_bits, end_bit):
    """Low-entropy keys (AND of 3 words) with index generation; full-size
    properties (the 2^28 run is in the benchmark's own parity check)."""
    n = 1 << 27
    keys = workloads.and3_keys(n, seed=3, mask_bits=mask_bits)
    g = GpuSort(keys, value_count=-1, end_bit=end_bit).run(engine)
    out_k, (idx,) = g.keys(), g.values()
    check_stable_sort(keys, out_k, idx, 0, end_bit)
    assert g.data.parity == len(mg.plan_passes(end_bit)) % 2


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 2047, 2048, 2049, 4095, 4096, 4097, 65537,
                               300_001])
@pytest.mark.parametrize("vc", [-1, 0, 1, 2])
def test_sizes_and_value_counts(engine, oracle, n, vc):
    keys = workloads.uniform_keys(n, seed=n)
    vals = [workloads.uniform_keys(n, seed=1000 + n + j) for j in range(max(vc, 0))]
    g = GpuSort(keys, vals, value_count=vc).run(engine)
    _check_vs_oracle(oracle, keys, g, vals, vc)


@pytest.mark.parametrize("n", [127 * 16384, 127 * 16384 + 1])
@pytest.mark.parametrize("vc", [-1, 1])
def test_tile_shape_switch(engine, oracle, n, vc):
    """The engine takes the 512 x 16 tiles below 128 tiles of 1024 x 16 keys
    (one per two CUs on 256 CUs) and the 1024 x 16 tiles from there on: both
    sides of the switch against the oracle."""
    keys = workloads.uniform_keys(n, seed=n)
    vals = [workloads.uniform_keys(n, seed=n + 1)] if vc == 1 else []
    g = GpuSort(keys, vals, value_count=vc).run(engine)
    engine.check_device()
    _check_vs_oracle(oracle, keys, g, vals, value_count=vc)


@pytest.mark.parametrize("vc", [3, 4, 5, 6])
def test_multi_value_arrays(engine, oracle, vc):
    n = 123_457
    keys = workloads.and3_keys(n, seed=vc)
    vals = [workloads.uniform_keys(n, seed=77 + j) for j in range(vc)]
    g = GpuSort(keys, vals).run(engine)
    _check_vs_oracle(oracle, keys, g, vals)


@pytest.mark.parametrize("fb,eb", [(0, 1), (31, 32), (7, 25), (0, 24), (4, 12), (3, 4), (0, 9),
                                   (16, 32), (1, 31)])
@pytest.mark.parametrize("vc", [0, 1, -1])
@pytest.mark.parametrize("which", ["default", "xcd"])
def test_bit_ranges(request, oracle, fb, eb, vc, which):
    """Digit widths below 8 bits (a narrow last pass: keys past the end take
    its top digit), with one lookback chain and with segments per XCD (nine
    tiles: the next pass's counts use its narrower digit)."""
    engine = request.getfixturevalue("engine" if which == "default" else "xcd_engine")
    n = 70_001
    keys = workloads.uniform_keys(n, seed=fb * 40 + eb)
    vals = [np.arange(n, dtype=np.uint32)] if vc == 1 else []
    g = GpuSort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb).run(engine)
    _check_vs_oracle(oracle, keys, g, vals, vc, fb, eb)
    assert g.data.parity == len(mg.plan_passes(eb - fb)) % 2


@pytest.mark.parametrize("kind", ["equal", "sorted", "reversed", "two_values", "skewed"])
@pytest.mark.parametrize("which", ["default", "xcd"])
def test_adversarial_inputs(request, oracle, kind, which):
    engine = request.getfixturevalue("engine" if which == "default" else "xcd_engine")
    n = 200_000
    if kind == "equal":
        keys = np.full(n, 0xDEADBEEF, dtype=np.uint32)
    elif kind == "sorted":
        keys = np.sort(workloads.uniform_keys(n, 5))
    elif kind == "reversed":
        keys = np.sort(workloads.uniform_keys(n, 6))[::-1].copy()
    elif kind == "two_values":
        keys = (workloads.uniform_keys(n, 7) & 1) * np.uint32(0xFFFFFFFF)
    else:
        keys = workloads.and3_keys(n, 8) & workloads.and3_keys(n, 9)
    g = GpuSort(keys, value_count=-1).run(engine)
    _check_vs_oracle(oracle, keys, g, value_count=-1)


def _typed_keys(n, key_type, seed):
    """int32 values, or float32 values mixing normal numbers, arbitrary bit
    patterns (NaNs, infinities, denormals) and signed zeros, as uint32 bits."""
    rng = np.random.default_rng(seed)
    if key_type == mg.KEYS_INT32:
        return rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32).view(np.uint32)
    f = (rng.standard_normal(n) * 100).astype(np.float32)
    pick = rng.integers(0, 4, n)
    f = np.where(pick == 1, rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
                 .view(np.float32), f)
    f = np.where(pick == 2, np.float32(0.0), f)
    f = np.where(pick == 3, np.float32(-0.0), f)
    return f.view(np.uint32)


@pytest.mark.parametrize("key_type", [mg.KEYS_INT32, mg.KEYS_FLOAT32])
@pytest.mark.parametrize("n", [1, 1000, 70_001, (1 << 22) + 3])
def test_signed_and_float_keys(engine, oracle, key_type, n):
    """sortTransformKeys to the sort's order, sortArray with index generation,
    sortTransformKeys back: keys and indices equal the oracle's map, stable
    sort and map back, and the device mapping alone equals the oracle's."""
    import torch
    bits = _typed_keys(n, key_type, seed=n + key_type)
    t = to_dev(bits)
    engine.transform_keys(t.data_ptr(), n, key_type, True)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(t, n), oracle.key_transform(bits, key_type, True))

    g = GpuSort(bits, value_count=-1)
    engine.transform_keys(g.data.keys[0], n, key_type, True)
    g.run(engine)
    engine.transform_keys(g.data.keys[0], n, key_type, False)
    torch.cuda.synchronize()
    engine.check_device()
    rk, (ri,) = oracle.stable_sort(oracle.key_transform(bits, key_type, True), value_count=-1)
    assert np.array_equal(g.keys(), oracle.key_transform(rk, key_type, False))
    assert np.array_equal(g.values()[0], ri)
    if key_type == mg.KEYS_INT32:
        assert np.array_equal(g.values()[0], np.argsort(bits.view(np.int32), kind="stable"))
    else:
        f = g.keys().view(np.float32)
        f = f[~np.isnan(f)]
        assert (np.diff(f.astype(np.float64)) >= 0).all()


def _typed_case(kind, n, key_type, seed):
    """Typed keys (uint32 bit patterns) of one early-exit shape."""
    rng = np.random.default_rng(seed)
    bits = _typed_keys(n, key_type, seed)
    if kind == "sorted_unsigned":  # in order as unsigned words, not as numbers
        return np.sort(bits)
    if kind == "sorted_typed":  # already in the sort's order: every pass skipped
        top = (bits >> np.uint32(31)).astype(bool)
        flip = np.uint32(0xFFFFFFFF) if key_type == mg.KEYS_FLOAT32 else np.uint32(0x80000000)
        return bits[np.argsort(bits ^ np.where(top, flip, np.uint32(0x80000000)), kind="stable")]
    if kind == "low_byte_constant":  # the first pass is skipped
        return (bits & np.uint32(0xFFFFFF00)) | np.uint32(0x5A)
    if kind == "top_byte_constant":  # the last pass is skipped
        return rng.integers(0, 1 << 24, n, dtype=np.int64).astype(np.uint32)
    return bits


@pytest.mark.parametrize("key_type", [mg.KEYS_INT32, mg.KEYS_FLOAT32])
@pytest.mark.parametrize("kind,n,vc,fb,eb,ee", [
    ("random", 1, -1, 0, 32, False),
    ("random", 70_001, -1, 0, 32, False),
    ("random", 70_001, 1, 0, 32, True),
    ("random", (1 << 22) + 3, -1, 0, 32, False),
    ("random", (1 << 22) + 3, 2, 0, 32, True),
    ("random", (1 << 24) + 5, 1, 0, 32, False),  # lookback chains per XCD
    ("random", 300_000, 0, 8, 32, False),
    ("random", 300_000, -1, 0, 8, False),  # one pass maps both ways
    ("random", 300_000, 1, 4, 21, True),
    ("sorted_unsigned", 200_000, -1, 0, 32, True),
    ("sorted_typed", 200_000, -1, 0, 32, True),
    ("low_byte_constant", 200_000, 1, 0, 32, True),
    ("top_byte_constant", 200_000, -1, 0, 32, True),
])
def test_typed_sort_maps_inside_the_passes(engine, oracle, key_type, kind, n, vc, fb, eb, ee):
    """sortArrayTyped (the order mapping inside the count and the first and
    last executed passes) against the oracle's map, stable sort and map back,
    which is what sortTransformKeys, sortArray, sortTransformKeys compute;
    with earlyExit, on inputs that skip every pass, the first or the last."""
    bits = _typed_case(kind, n, key_type, seed=n + 7 * key_type + len(kind))
    n = bits.size
    values = [workloads.uniform_keys(n, 500 + j) for j in range(max(vc, 0))]
    g = GpuSort(bits, values, value_count=vc, first_bit=fb, end_bit=eb)
    g.data.earlyExit = ee
    g.run(engine, key_type=key_type)
    engine.check_device()
    mapped = oracle.key_transform(bits, key_type, True)
    rk, rv = oracle.stable_sort(mapped, values, value_count=vc, first_bit=fb, end_bit=eb)
    assert np.array_equal(g.keys(), oracle.key_transform(rk, key_type, False))
    for o, r in zip(g.values(), rv):
        assert np.array_equal(o, r)
    if kind == "sorted_typed":
        assert g.data.parity == 0 and np.array_equal(g.keys(), bits)


@pytest.mark.parametrize("key_type,vc", [(mg.KEYS_FLOAT32, 1), (mg.KEYS_INT32, -1)])
def test_typed_sort_equals_three_calls_at_2p26(engine, key_type, vc):
    """At 2^26 keys (lookback chains per XCD), sortArrayTyped gives the same
    bits as sortTransformKeys, sortArray, sortTransformKeys on the same input
    (both compared on the device); the smaller typed tests tie both to the
    oracle."""
    import torch
    n = 1 << 26
    keys = workloads.uniform_keys_torch(n, 260 + key_type, "cuda", torch=torch)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    outs = []
    for fused in (True, False):
        k0, k1 = keys.clone(), torch.empty_like(keys)
        v0, v1 = (vals.clone() if vc == 1 else torch.empty_like(vals)), torch.empty_like(vals)
        d = mg.SortData()
        d.maxElements = d.numElements = n
        d.valueCount = vc
        d.firstBit, d.endBit = 0, 32
        d.keys[0], d.keys[1] = k0.data_ptr(), k1.data_ptr()
        d.values[0][0], d.values[0][1] = v0.data_ptr(), v1.data_ptr()
        torch.cuda.synchronize()
        if fused:
            engine.sort_array_typed(d, key_type)
        else:
            engine.transform_keys(d.keys[0], n, key_type, True)
            engine.sort_array(d)
            engine.transform_keys(d.keys[0], n, key_type, False)
        torch.cuda.synchronize()
        engine.check_device()
        by = {t.data_ptr(): t for t in (k0, k1, v0, v1)}
        outs.append((by[d.keys[0]], by[d.values[0][0]], d.parity))
    assert outs[0][2] == outs[1][2]
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    # and the result is in the type's order
    k = outs[0][0]
    if key_type == mg.KEYS_INT32:
        assert bool((k[1:] >= k[:-1]).all())
    else:
        f = k.view(torch.float32)
        f = f[~torch.isnan(f)]
        assert bool((f[1:] >= f[:-1]).all())


@pytest.mark.parametrize("env", [{"MGPUSORT_RANK": "0"}, {"MGPUSORT_TILE": "256x16"},
                                 {"MGPUSORT_XCD": "1"}])
@pytest.mark.parametrize("key_type", [mg.KEYS_INT32, mg.KEYS_FLOAT32])
def test_typed_sort_other_engines(oracle, env, key_type):
    """Engines whose tile shape has no key-mapping pass forms (match-mask
    ranking, 256 x 16 tiles) map the keys in place around the sort instead;
    chains per XCD at every size use the mapping forms of the _x kernels.
    Each gives the oracle's result, also with earlyExit skipping passes."""
    eng = make_engine(env)
    try:
        for kind, n, vc, ee in (("random", 100_003, 1, False), ("random", 1 << 20, -1, False),
                                ("low_byte_constant", 50_000, 2, True),
                                ("sorted_typed", 30_000, -1, True)):
            bits = _typed_case(kind, n, key_type, seed=n + key_type)
            values = [workloads.uniform_keys(n, 900 + j) for j in range(max(vc, 0))]
            g = GpuSort(bits, values, value_count=vc)
            g.data.earlyExit = ee
            g.run(eng, key_type=key_type)
            eng.check_device()
            rk, rv = oracle.stable_sort(oracle.key_transform(bits, key_type, True), values,
                                        value_count=vc)
            assert np.array_equal(g.keys(), oracle.key_transform(rk, key_type, False)), kind
            for o, r in zip(g.values(), rv):
                assert np.array_equal(o, r), kind
    finally:
        eng.close()


def test_typed_sort_rejects_bad_key_types(engine, oracle):
    g = GpuSort(workloads.uniform_keys(100, 3))
    for bad in (-1, 3, 7):
        with pytest.raises(mg.MgpuError) as e:
            g.run(engine, key_type=bad)
        assert e.value.status == mg.SortStatus.INVALID_VALUE
    # SORT_KEYS_UINT32 is sortArray itself.
    keys = workloads.uniform_keys(5000, 4)
    g = GpuSort(keys, value_count=-1).run(engine, key_type=mg.KEYS_UINT32)
    _check_vs_oracle(oracle, keys, g, value_count=-1)


def test_transform_keys_rejects_bad_arguments(engine):
    with pytest.raises(mg.MgpuError):
        engine.transform_keys(0, 10, mg.KEYS_INT32, True)
    t = to_dev(np.zeros(4, dtype=np.uint32))
    with pytest.raises(mg.MgpuError):
        engine.transform_keys(t.data_ptr(), 4, 7, True)
    with pytest.raises(mg.MgpuError):
        engine.transform_keys(t.data_ptr(), -1, mg.KEYS_FLOAT32, True)
    engine.transform_keys(t.data_ptr(), 0, mg.KEYS_FLOAT32, True)


def test_zero_bits_and_zero_elements_do_nothing(engine):
    keys = workloads.uniform_keys(5000, 11)
    g = GpuSort(keys, value_count=-1, first_bit=9, end_bit=9).run(engine)
    assert np.array_equal(g.keys(), keys) and g.data.parity == 0
    g = GpuSort(np.zeros(0, dtype=np.uint32), [np.zeros(0, dtype=np.uint32)], capacity=16)
    g.run(engine)
    assert g.data.parity == 0


@pytest.mark.parametrize("bit_pass", [1, 2, 3, 4, 5, 6, 0, 9])
def test_sort_array_ex(engine, oracle, bit_pass):
    n = 50_001
    keys = workloads.uniform_keys(n, seed=bit_pass + 50)
    fb, eb = 3, 27
    g = GpuSort(keys, value_count=-1, first_bit=fb, end_bit=eb)
    g.run(engine, ex=(256, 8, bit_pass, False))
    _check_vs_oracle(oracle, keys, g, value_count=-1, fb=fb, eb=eb)
    bp = 6 if bit_pass <= 0 or bit_pass > 6 else bit_pass
    passes = -(-(eb - fb) // bp)
    assert g.data.parity == passes % 2
    # The array at [0] before the sort is at [parity] after it.
    assert g.data.keys[g.data.parity] == g.start_keys


def test_sort_array_ex_rejects_reference_invalid_shapes(engine):
    g = GpuSort(workloads.uniform_keys(100, 1))
    for args in [(64, 8, 6, False), (256, 4, 6, False), (128, 16, 6, False)]:
        with pytest.raises(mg.MgpuError) as e:
            g.run(engine, ex=args)
        assert e.value.status == mg.SortStatus.INVALID_VALUE


def test_invalid_descriptors(engine):
    L = mg.lib()
    g = GpuSort(workloads.uniform_keys(100, 1))
    d = g.data
    d.numElements = d.maxElements + 1
    assert L.sortArray(engine.handle, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE
    d.numElements = 100
    d.endBit = 33
    assert L.sortArray(engine.handle, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE
    d.endBit, d.firstBit = 8, 9
    assert L.sortArray(engine.handle, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE
    d.firstBit, d.endBit, d.valueCount = 0, 32, 7
    assert L.sortArray(engine.handle, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE
    d.valueCount = 1  # value pointers missing
    assert L.sortArray(engine.handle, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE


def test_library_allocated_data_and_parity_chain(engine, oracle):
    """sortCreateData buffers; two sorts in a row continue the parity chain
    the way a multi-key client (bwt/src/mgpubwt/mgpubwt.cpp:262-292) uses it."""
    torch = pytest.importorskip("torch")
    L = mg.lib()
    n = 9000
    pd = ctypes.POINTER(mg.SortData)()
    assert L.sortCreateData(engine.handle, n, 1, ctypes.byref(pd)) == 0
    d = pd.contents
    assert d.maxElements == 10240 and d.numElements == n and d.endBit == 32
    assert all(d.keys[i] for i in range(2)) and all(d.values[0][i] for i in range(2))
    assert not d.values[1][0]
    lo = workloads.uniform_keys(n, 1)
    k0 = d.keys[0]
    torch.cuda.synchronize()
    copy_d2d(k0, to_dev(lo).data_ptr(), 4 * n)
    copy_d2d(d.values[0][0], to_dev(np.arange(n, dtype=np.uint32)).data_ptr(), 4 * n)
    d.firstBit, d.endBit = 0, 24  # 3 passes: odd
    assert L.sortArray(engine.handle, pd) == 0
    torch.cuda.synchronize()
    assert d.parity == 1 and d.keys[1] == k0
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    copy_d2d(out.data_ptr(), d.keys[0], 4 * n)
    rk, _ = oracle.stable_sort(lo, first_bit=0, end_bit=24)
    assert np.array_equal(to_host(out, n), rk)
    # Second sort (16 bits = 2 passes, even) continues from keys[0] and keeps
    # the chain: the original buffer stays at [parity].
    d.firstBit, d.endBit = 8, 24
    assert L.sortArray(engine.handle, pd) == 0
    torch.cuda.synchronize()
    assert d.parity == 1 and d.keys[1] == k0
    copy_d2d(out.data_ptr(), d.keys[0], 4 * n)
    rk2, _ = oracle.stable_sort(rk, first_bit=8, end_bit=24)
    assert np.array_equal(to_host(out, n), rk2)
    assert L.sortDestroyData(pd) == 0


def test_sort_host_and_device(engine, oracle):
    L = mg.lib()
    n = 33_333
    keys = workloads.uniform_keys(n, 31)
    vals = np.arange(n, dtype=np.uint32)
    hk, hv = keys.copy(), vals.copy()
    P = ctypes.POINTER(ctypes.c_uint32)
    assert L.sortHost(engine.handle, hk.ctypes.data_as(P), hv.ctypes.data_as(P), n, 24) == 0
    rk, (rv,) = oracle.stable_sort(keys, [vals], first_bit=0, end_bit=24)
    assert np.array_equal(hk, rk) and np.array_equal(hv, rv)
    # sortDevice with an odd pass count must still leave the result in the
    # caller's buffers.
    dk, dv = to_dev(keys), to_dev(vals)
    import torch
    torch.cuda.synchronize()
    assert L.sortDevice(engine.handle, dk.data_ptr(), dv.data_ptr(), n, 24) == 0
    assert np.array_equal(to_host(dk, n), rk) and np.array_equal(to_host(dv, n), rv)
    dk = to_dev(keys)
    torch.cuda.synchronize()
    assert L.sortDevice(engine.handle, dk.data_ptr(), 0, n, 32) == 0
    rk32, _ = oracle.stable_sort(keys)
    assert np.array_equal(to_host(dk, n), rk32)


@pytest.mark.parametrize("kind,executed", [("random", 4), ("sorted", 0), ("top_constant", 3),
                                           ("low_constant", 3), ("two_constant", 2),
                                           ("mid_constant", 3)])
@pytest.mark.parametrize("vc", [-1, 1])
@pytest.mark.parametrize("which", ["default", "xcd"])
def test_early_exit(request, oracle, kind, executed, vc, which):
    """earlyExit skips passes whose digit is the same for every key, and every
    pass of already-sorted input; the output is unchanged and parity follows
    the executed passes (with segments per XCD, each executed pass counts
    the segments of the next executed one)."""
    engine = request.getfixturevalue("engine" if which == "default" else "xcd_engine")
    n = 150_001
    keys = workloads.uniform_keys(n, seed=90)
    if kind == "sorted":
        keys = np.sort(keys)
    elif kind == "top_constant":
        keys = (keys & np.uint32(0xFFFFFF)) | np.uint32(0x5A000000)
    elif kind == "low_constant":
        keys = (keys & np.uint32(0xFFFFFF00)) | np.uint32(0x33)
    elif kind == "two_constant":
        keys = (keys & np.uint32(0x00FFFF00)) | np.uint32(0x77000011)
    elif kind == "mid_constant":  # pass 0 is followed by pass 2
        keys = (keys & np.uint32(0xFFFF00FF)) | np.uint32(0x00004400)
    vals = [np.arange(n, dtype=np.uint32)] if vc == 1 else []
    g = GpuSort(keys, vals, value_count=vc)
    g.data.earlyExit = True
    g.run(engine)
    engine.check_device()
    _check_vs_oracle(oracle, keys, g, vals, vc)
    assert g.data.parity == executed % 2
    assert g.data.keys[g.data.parity] == g.start_keys


def test_engine_info(engine):
    """The engine's creation-time choices on a gfx950 device: the atomic-add
    ranking and one lookback chain per XCD (MI355X: 8 XCDs), with no device
    inconsistency recorded."""
    info = engine.info()
    assert info["numCUs"] > 0
    assert info["rankMethod"] in (0, 1)
    assert info["xcdSegments"] == 1 or 2 <= info["xcdSegments"] <= 8
    assert info["smallTileKeys"] == 512 * 16 and info["largeTileKeys"] == 1024 * 16
    assert info["deviceError"] == 0
    print("engine info:", info, flush=True)


def test_keys_past_the_end_are_untouched(engine):
    n, cap = 5000, 8192
    keys = workloads.uniform_keys(n, 4)
    tail = workloads.uniform_keys(cap - n, 5)
    g = GpuSort(keys, value_count=-1, capacity=cap)
    g.tensors[0][n:].copy_(to_dev(tail)[: cap - n])
    g.tensors[1][n:].fill_(7)
    g.run(engine)
    # Both key buffers keep their contents past numElements.
    assert np.array_equal(to_host(g.tensors[0], cap)[n:], tail)
    assert (to_host(g.tensors[1], cap)[n:] == 7).all()


def test_repeat_determinism(engine):
    n = 1_000_003
    keys = workloads.and3_keys(n, 12)
    first = None
    for _ in range(10):
        g = GpuSort(keys, value_count=-1).run(engine)
        out = (g.keys(), g.values()[0])
        if first is None:
            first = out
        else:
            assert np.array_equal(out[0], first[0]) and np.array_equal(out[1], first[1])


@pytest.mark.parametrize("offset", [1, 2, 3])
@pytest.mark.parametrize("n", [100_003, (1 << 22) + 5, (1 << 24) + 3])
def test_sub_array_pointers(engine, oracle, offset, n):
    """Keys and values that start 4, 8 or 12 bytes into larger buffers (not
    16-byte aligned: the count reads them without vector loads), below and
    above the size where the count switches kernels and where passes split
    their tiles into one lookback chain per XCD; nothing outside
    [offset, offset + n) of either buffer changes."""
    import torch
    keys = workloads.uniform_keys(n, 700 + offset)
    vals = workloads.uniform_keys(n, 800 + offset)
    rk, (rv,) = oracle.stable_sort(keys, [vals])
    pad = 8
    bufs = []
    for a in (keys, vals):
        t = torch.full((n + 2 * pad,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
        t[offset:offset + n].copy_(torch.from_numpy(a.view(np.int32)))
        bufs.append(t)
    alt = [torch.full((n + 2 * pad,), 0x3C3C3C3C, dtype=torch.int32, device="cuda") for _ in range(2)]
    d = mg.SortData()
    d.maxElements = d.numElements = n
    d.valueCount, d.firstBit, d.endBit = 1, 0, 32
    d.keys[0], d.keys[1] = bufs[0].data_ptr() + 4 * offset, alt[0].data_ptr() + 4 * offset
    d.values[0][0], d.values[0][1] = bufs[1].data_ptr() + 4 * offset, alt[1].data_ptr() + 4 * offset
    torch.cuda.synchronize()
    engine.sort_array(d)
    torch.cuda.synchronize()
    # Four passes: the result is back in the caller's first buffers.
    assert d.parity == 0 and d.keys[0] == bufs[0].data_ptr() + 4 * offset
    out_k = bufs[0].cpu().numpy().view(np.uint32)
    out_v = bufs[1].cpu().numpy().view(np.uint32)
    assert np.array_equal(out_k[offset:offset + n], rk)
    assert np.array_equal(out_v[offset:offset + n], rv)
    for t, fill in ((bufs[0], 0x5A5A5A5A), (bufs[1], 0x5A5A5A5A), (alt[0], 0x3C3C3C3C),
                    (alt[1], 0x3C3C3C3C)):
        h = t.cpu().numpy().view(np.uint32)
        assert np.all(h[:offset] == fill) and np.all(h[offset + n:] == fill)


def test_stream_switch_waits_for_queued_work(oracle):
    """Two sorts on one engine, the second queued on another stream right
    after sortSetStream: it must wait for the first (both use the engine's
    scratch), and both results must be exact."""
    import torch
    s1 = torch.cuda.Stream()
    s2 = torch.cuda.Stream()
    eng = make_engine()
    try:
        k1 = workloads.uniform_keys(1 << 24, seed=400)
        k2 = workloads.uniform_keys(1 << 24, seed=401)
        g1, g2 = GpuSort(k1, [np.arange(k1.size, dtype=np.uint32)]), GpuSort(k2)
        torch.cuda.synchronize()
        eng.set_stream(s1.cuda_stream)
        eng.sort_array(g1.data)
        eng.set_stream(s2.cuda_stream)
        eng.sort_array(g2.data)
        torch.cuda.synchronize()
        eng.check_device()
        rk, (rv,) = oracle.stable_sort(k1, [np.arange(k1.size, dtype=np.uint32)])
        assert np.array_equal(g1.keys(), rk) and np.array_equal(g1.values()[0], rv)
        assert np.array_equal(g2.keys(), np.sort(k2))
    finally:
        eng.close()


def test_size_beyond_2p29(engine):
    """n in (2^29, 2^30), where byte offsets of an array pass 2^31: key +
    generated index, checked on the device by the properties that fix a
    stable sort (sorted keys, a permutation, keys gathered through it, ties in
    source order)."""
    import torch
    n = (1 << 29) + 12_345
    gen = torch.Generator(device="cuda")
    gen.manual_seed(29)
    k0 = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device="cuda", generator=gen)
    keys_in = k0.clone()
    k1, v0, v1 = (torch.empty_like(k0) for _ in range(3))
    d = mg.SortData()
    d.maxElements = d.numElements = n
    d.valueCount, d.firstBit, d.endBit = -1, 0, 32
    d.keys[0], d.keys[1] = k0.data_ptr(), k1.data_ptr()
    d.values[0][0], d.values[0][1] = v0.data_ptr(), v1.data_ptr()
    torch.cuda.synchronize()
    engine.sort_array(d)
    torch.cuda.synchronize()
    engine.check_device()
    by_ptr = {t.data_ptr(): t for t in (k0, k1, v0, v1)}
    out_k = by_ptr[d.keys[0]].long() & 0xFFFFFFFF
    idx = by_ptr[d.values[0][0]].long()
    assert bool((out_k[1:] >= out_k[:-1]).all()), "keys not sorted"
    assert int(idx.min()) >= 0 and int(idx.max()) < n
    seen = torch.zeros(n, dtype=torch.bool, device="cuda")
    seen[idx] = True
    assert bool(seen.all()), "indices are not a permutation"
    assert bool(((keys_in.long() & 0xFFFFFFFF)[idx] == out_k).all()), "keys do not follow indices"
    ties = out_k[1:] == out_k[:-1]
    assert bool((idx[1:][ties] > idx[:-1][ties]).all()), "equal keys out of source order"


@pytest.fixture(scope="module")
def ee_engine():
    """The early-exit kernel for the first pass that runs whenever earlyExit
    is set (MGPUSORT_EE=1), whatever the descent counts say."""
    eng = make_engine({"MGPUSORT_EE": "1"})
    yield eng
    eng.close()


def _ee_kernels_used(engine):
    return [k for k in engine.read_timing() if "_ee_" in k]


@pytest.mark.parametrize("case,fb,eb,executed,ee", [
    # sorted keys with 0.5% of their 1K-key blocks shuffled: the low digits
    # are disordered everywhere
    ("nearly_sorted", 0, 32, 4, False),
    # ... but bits [18, 26) and [20, 26) have few descents (4-6K over 2K
    # waves): the first pass takes the early-exit kernel
    ("nearly_sorted", 18, 26, 1, True),
    ("nearly_sorted", 20, 32, 2, True),
    # in order of the top byte only: the low passes scramble that order
    ("top_sorted", 0, 32, 4, False),
    # a sort by the top byte of input already in top-byte order: the identity
    ("top_sorted", 24, 32, 0, False),
    # by bits [16, 32) of input in order of its top 16 bits but for 64 keys:
    # pass [16, 24) runs with the early-exit kernel, then pass [24, 32)
    ("top_sorted_low_run", 16, 32, 2, True),
    # pass 0's digit is in order in the input (skipped), pass 1's is not
    ("mid_sorted", 8, 24, 1, False),
])
@pytest.mark.parametrize("vc", [-1, 0, 1, 3])
@pytest.mark.parametrize("which", ["default", "xcd"])
def test_early_exit_partial_order(request, oracle, case, fb, eb, executed, ee, vc, which):
    """earlyExit on partially ordered input: a leading pass whose digit is
    already in order in the input is skipped (the identity); the first pass
    that runs may take the early-exit kernel, which ranks waves and writes
    tiles already in digit order without the atomics and the staging.  The
    output equals the oracle's either way, and parity follows the executed
    passes."""
    engine = request.getfixturevalue("engine" if which == "default" else "xcd_engine")
    n = (1 << 21) + 77
    if case == "nearly_sorted":
        keys = workloads.nearly_sorted_keys(n, seed=91, shuffled_fraction=0.005, block=1024)
    elif case == "top_sorted":
        keys = workloads.top_sorted_keys(n, seed=92)
    elif case == "top_sorted_low_run":
        # top byte in order, bits [16, 24) in order within each top byte for
        # all but a few keys: pass [16, 24) has few descents
        keys = workloads.top_sorted_keys(n, seed=93, top_bits=16)
        k = workloads.words(94, 0, 64) % np.uint32(n)
        keys[k] = workloads.words(95, 0, 64)
    else:  # mid_sorted: bits [8, 16) nondecreasing, [16, 24) random
        keys = workloads.uniform_keys(n, seed=96)
        mid = np.sort((keys >> np.uint32(8)) & np.uint32(0xFF))
        keys = (keys & np.uint32(0xFFFF00FF)) | (mid << np.uint32(8))
    vals = [workloads.uniform_keys(n, seed=97 + j) for j in range(max(vc, 0))]
    g = GpuSort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb)
    g.data.earlyExit = True
    engine.enable_timing(True)
    engine.read_timing()
    g.run(engine)
    used = _ee_kernels_used(engine)
    engine.enable_timing(False)
    engine.check_device()
    _check_vs_oracle(oracle, keys, g, vals, vc, fb, eb)
    assert g.data.parity == executed % 2
    assert bool(used) == (ee and executed > 0), used


@pytest.mark.parametrize("kind", ["random", "nearly_sorted", "top_sorted", "waves_sorted"])
@pytest.mark.parametrize("vc", [-1, 0, 1, 2])
@pytest.mark.parametrize("which", ["ee", "ee_xcd"])
def test_early_exit_kernel_forced(request, oracle, kind, vc, which):
    """The early-exit kernel on every kind of input, also where it finds
    nothing in order (random) and where whole tiles are (waves_sorted: keys
    whose top byte is in order, sorted by that byte, so every pass-0 tile is
    ordered).  Bit-exact against the oracle."""
    env = {"MGPUSORT_EE": "1"}
    if which == "ee_xcd":
        env["MGPUSORT_XCD"] = "1"
    eng = make_engine(env)
    try:
        n = 3 * 16384 * 8 + 4097
        fb, eb = 0, 32
        if kind == "random":
            keys = workloads.uniform_keys(n, seed=98)
        elif kind == "nearly_sorted":
            keys = workloads.nearly_sorted_keys(n, seed=99, shuffled_fraction=0.3, block=1024)
            fb, eb = 20, 32
        elif kind == "top_sorted":
            keys = workloads.top_sorted_keys(n, seed=100)
            fb, eb = 16, 32
        else:
            keys = np.sort(workloads.uniform_keys(n, seed=101) >> np.uint32(24)) << np.uint32(24)
            keys |= workloads.uniform_keys(n, seed=102, mask_bits=24)
            fb, eb = 24, 32
            keys[5] ^= np.uint32(0xFF000000)  # one descent: the pass runs
        vals = [workloads.uniform_keys(n, seed=103 + j) for j in range(max(vc, 0))]
        g = GpuSort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb)
        g.data.earlyExit = True
        g.run(eng)
        eng.check_device()
        _check_vs_oracle(oracle, keys, g, vals, vc, fb, eb)
    finally:
        eng.close()
