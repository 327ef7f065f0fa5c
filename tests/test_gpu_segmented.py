"""Segmented sort parity (needs a gfx950 GPU): every segment must equal the
oracle's stable sort of that segment alone, bit for bit, including values and
generated indices, for every kernel variant and size class."""
import os

import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import GpuSort, empty_dev, make_engine, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    eng = make_engine()
    yield eng
    eng.close()


def _segments(sizes):
    starts = np.zeros(len(sizes) + 1, dtype=np.int64)
    starts[1:] = np.cumsum(sizes)
    return starts


def _check_segments(oracle, keys, vals, vc, starts, out_k, out_v, fb, eb):
    for i in range(len(starts) - 1):
        a, b = int(starts[i]), int(starts[i + 1])
        if a == b:
            continue
        rk, rv = oracle.stable_sort(keys[a:b], [v[a:b] for v in vals], value_count=vc,
                                    first_bit=fb, end_bit=eb)
        assert np.array_equal(out_k[a:b], rk), f"segment {i} (n={b - a}) keys differ"
        for j, (o, r) in enumerate(zip(out_v, rv)):
            assert np.array_equal(o[a:b], r), f"segment {i} (n={b - a}) values[{j}] differ"


def _run_segmented(engine, keys, vals, vc, starts, fb=0, eb=32):
    g = GpuSort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb)
    import torch
    torch.cuda.synchronize()
    engine.sort_segmented(g.data, starts)
    torch.cuda.synchronize()
    engine.check_device()
    return g


SIZES = [0, 1, 2, 3, 63, 64, 65, 127, 128, 129, 500, 1023, 1024, 1025, 2047, 2048, 2049, 4095,
         4096, 4097, 8191, 8192, 8193, 20000, 65535, 65536, 100_000, 300_000]


@pytest.mark.parametrize("variant", [None, "r8i16", "r8i8", "r4i8", "r5i16", "r6i8", "r7i16"])
@pytest.mark.parametrize("wide_min", [None, "65537"])
def test_size_classes_per_variant(engine, oracle, variant, wide_min):
    """Every size class with every kernel variant.  By default segments above
    4096 keys take the device-wide passes; MGPUSORT_SEG_WIDE_MIN=65537 keeps
    segments up to 2^16 keys in the persistent kernel, whose one-workgroup
    global routine then sorts those above its LDS capacity."""
    old = os.environ.get("MGPUSORT_SEG_VARIANT")
    old_wide = os.environ.get("MGPUSORT_SEG_WIDE_MIN")
    if variant:
        os.environ["MGPUSORT_SEG_VARIANT"] = variant
    if wide_min:
        os.environ["MGPUSORT_SEG_WIDE_MIN"] = wide_min
    try:
        starts = _segments(SIZES)
        n = int(starts[-1])
        keys = workloads.uniform_keys(n, seed=41)
        keys[: n // 3] &= np.uint32(0x0F0F)  # ties
        vals = [np.arange(n, dtype=np.uint32)]
        g = _run_segmented(engine, keys, vals, 1, starts)
        _check_segments(oracle, keys, vals, 1, starts, g.keys(), g.values(), 0, 32)
        assert g.data.parity == 0
    finally:
        if old is None:
            os.environ.pop("MGPUSORT_SEG_VARIANT", None)
        else:
            os.environ["MGPUSORT_SEG_VARIANT"] = old
        if old_wide is None:
            os.environ.pop("MGPUSORT_SEG_WIDE_MIN", None)
        else:
            os.environ["MGPUSORT_SEG_WIDE_MIN"] = old_wide


@pytest.mark.parametrize("vc", [-1, 0, 2, 6])
@pytest.mark.parametrize("fb,eb", [(0, 32), (0, 24), (7, 25), (16, 32), (0, 5)])
def test_value_counts_and_bit_ranges(engine, oracle, vc, fb, eb):
    sizes = [1, 17, 700, 1500, 3000, 5000, 9000, 70_000]
    starts = _segments(sizes)
    n = int(starts[-1])
    keys = workloads.and3_keys(n, seed=fb * 33 + eb)
    vals = [workloads.uniform_keys(n, seed=900 + j) for j in range(max(vc, 0))]
    g = _run_segmented(engine, keys, vals, vc, starts, fb, eb)
    _check_segments(oracle, keys, vals, vc, starts, g.keys(), g.values(), fb, eb)
    assert g.data.parity == len(mg.plan_passes(eb - fb)) % 2


def test_config4_4096_log_uniform_segments(engine, oracle):
    """4096 segments, sizes log-uniform in [1, 2^16], key + one value array."""
    sizes = workloads.log_uniform_sizes(4096, 0, 16, seed=4)
    starts = _segments(sizes)
    n = int(starts[-1])
    keys = workloads.uniform_keys(n, seed=44)
    vals = [workloads.uniform_keys(n, seed=45)]
    g = _run_segmented(engine, keys, vals, 1, starts)
    _check_segments(oracle, keys, vals, 1, starts, g.keys(), g.values(), 0, 32)


@pytest.mark.parametrize("eb", [32, 24])
def test_plan_replay(engine, oracle, eb):
    """A captured plan replayed on fresh inputs, in both buffer orientations."""
    import torch
    sizes = [5, 900, 3000, 12_000, 40_000, 1]
    starts = _segments(sizes)
    n = int(starts[-1])
    g = GpuSort(np.zeros(n, dtype=np.uint32), [np.zeros(n, dtype=np.uint32)], end_bit=eb)
    plan = engine.segmented_plan(g.data, starts)
    for rep in range(4):
        keys = workloads.uniform_keys(n, seed=100 + rep)
        vals = [np.arange(n, dtype=np.uint32) * np.uint32(3)]
        g.by_ptr[g.data.keys[0]][:n].copy_(to_dev(keys)[:n])
        g.by_ptr[g.data.values[0][0]][:n].copy_(to_dev(vals[0])[:n])
        torch.cuda.synchronize()
        plan.execute(g.data)
        torch.cuda.synchronize()
        engine.check_device()
        _check_segments(oracle, keys, vals, 1, starts, g.keys(), g.values(), 0, eb)
        assert g.data.parity == (rep + 1) % 2 if eb == 24 else g.data.parity == 0
    plan.close()


def test_batch_of_separate_buffers_alone_vs_mix(engine, oracle):
    """sortSegmentedBatch over arrays with buffers of their own; each result
    equals sortArray on that array alone."""
    import torch
    sizes = [3, 1000, 5000, 33_000, 300_000, 64]
    segs, bufs, inputs = [], [], []
    for i, n in enumerate(sizes):
        k = workloads.uniform_keys(n, seed=200 + i)
        v = np.arange(n, dtype=np.uint32)
        t = [to_dev(k), empty_dev(n), to_dev(v), empty_dev(n)]
        s = mg.SortSegment()
        s.keys[0], s.keys[1] = t[0].data_ptr(), t[1].data_ptr()
        s.values[0][0], s.values[0][1] = t[2].data_ptr(), t[3].data_ptr()
        s.numElements = n
        segs.append(s)
        bufs.append(t)
        inputs.append((k, v))
    torch.cuda.synchronize()
    engine.sort_batch(segs, 1, 0, 24)
    engine.check_device()
    side = len(mg.plan_passes(24)) % 2
    for (k, v), t, n in zip(inputs, bufs, sizes):
        alone = GpuSort(k, [v], end_bit=24).run(engine)
        mix_k, mix_v = to_host(t[side], n), to_host(t[2 + side], n)
        assert np.array_equal(mix_k, alone.keys()) and np.array_equal(mix_v, alone.values()[0])
        rk, (rv,) = oracle.stable_sort(k, [v], first_bit=0, end_bit=24)
        assert np.array_equal(mix_k, rk) and np.array_equal(mix_v, rv)
