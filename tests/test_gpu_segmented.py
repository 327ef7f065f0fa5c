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
    global routine then sorts those above its LDS capacity (the kernel built
    without that routine, _ng, takes launches that need none)."""
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
        engine.enable_timing(True)
        engine.read_timing()
        g = _run_segmented(engine, keys, vals, 1, starts)
        names = {k for k in engine.read_timing() if k.startswith("mgpu_segsort")}
        engine.enable_timing(False)
        _check_segments(oracle, keys, vals, 1, starts, g.keys(), g.values(), 0, 32)
        assert g.data.parity == 0
        # Segments above a workgroup's tile need the kernel with the global
        # routine; without any (the shipped table's variants: 512 threads,
        # every tile at least 4096 keys), every launch takes the one without.
        if wide_min:
            assert any(not k.endswith("_ng") for k in names), names
        elif variant is None:
            assert names and all(k.endswith("_ng") for k in names), names
    finally:
        engine.enable_timing(False)
        if old is None:
            os.environ.pop("MGPUSORT_SEG_VARIANT", None)
        else:
            os.environ["MGPUSORT_SEG_VARIANT"] = old
        if old_wide is None:
            os.environ.pop("MGPUSORT_SEG_WIDE_MIN", None)
        else:
            os.environ["MGPUSORT_SEG_WIDE_MIN"] = old_wide


@pytest.mark.parametrize("wide_min", [None, "65537"])
def test_size_classes_match_mask_ranking(oracle, wide_min):
    """The persistent kernels' ranking by match masks (an engine made with
    MGPUSORT_RANK=0), whose per-wave mask tables share LDS with the staging
    of keys and payloads, over every size class; with MGPUSORT_SEG_WIDE_MIN
    =65537 also the one-workgroup global routine.  One value array (carried
    through the passes) and two (gathered at the end)."""
    saved = os.environ.get("MGPUSORT_SEG_WIDE_MIN")
    eng = make_engine({"MGPUSORT_RANK": "0"})
    try:
        if wide_min:
            os.environ["MGPUSORT_SEG_WIDE_MIN"] = wide_min
        starts = _segments(SIZES)
        n = int(starts[-1])
        keys = workloads.uniform_keys(n, seed=43)
        keys[: n // 3] &= np.uint32(0x0F0F)  # ties
        for vc in (1, 2):
            vals = [workloads.uniform_keys(n, seed=50 + j) for j in range(vc)]
            g = _run_segmented(eng, keys, vals, vc, starts)
            _check_segments(oracle, keys, vals, vc, starts, g.keys(), g.values(), 0, 32)
    finally:
        eng.close()
        if saved is None:
            os.environ.pop("MGPUSORT_SEG_WIDE_MIN", None)
        else:
            os.environ["MGPUSORT_SEG_WIDE_MIN"] = saved


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


def test_plan_replay_determinism(engine, oracle):
    """A race check: one captured plan over a mixed batch (tiny segments on
    the register network, wave and workgroup sorts, device-wide segments with
    the count pieces and grouped tile counters), replayed 10 times from the
    same input, gives the same bits every time, and the first result is the
    oracle's."""
    import torch
    sizes = workloads.log_uniform_sizes(1500, 0, 16, seed=21)
    starts = _segments(sizes)
    n = int(starts[-1])
    keys = workloads.and3_keys(n, seed=22)  # skewed digits: many equal ranks
    vals = [np.arange(n, dtype=np.uint32)]
    g = GpuSort(np.zeros(n, dtype=np.uint32), [np.zeros(n, dtype=np.uint32)])
    plan = engine.segmented_plan(g.data, starts)
    first = None
    for rep in range(10):
        g.by_ptr[g.data.keys[0]][:n].copy_(to_dev(keys)[:n])
        g.by_ptr[g.data.values[0][0]][:n].copy_(to_dev(vals[0])[:n])
        torch.cuda.synchronize()
        plan.execute(g.data)
        torch.cuda.synchronize()
        engine.check_device()
        out = (g.keys().copy(), g.values()[0].copy())
        if first is None:
            first = out
            _check_segments(oracle, keys, vals, 1, starts, out[0], [out[1]], 0, 32)
        else:
            assert np.array_equal(out[0], first[0]) and np.array_equal(out[1], first[1]), rep
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


def test_plan_keeps_its_scratch_when_the_lane_grows(engine, oracle):
    """A captured plan addresses scratch of its own: a larger sortArray and a
    segmented sort with more device-wide arrays on the same engine (both grow
    the engine's scratch, freeing the old) must not disturb its replay."""
    import torch
    sizes = [6000, 50_000, 9000, 120_000]
    starts = _segments(sizes)
    n = int(starts[-1])
    g = GpuSort(np.zeros(n, dtype=np.uint32), [np.zeros(n, dtype=np.uint32)])
    plan = engine.segmented_plan(g.data, starts)
    big = GpuSort(workloads.uniform_keys(1 << 26, seed=300))
    big.run(engine)
    engine.check_device()
    many = [5000 + 37 * i for i in range(64)]
    ms = _segments(many)
    mk = workloads.uniform_keys(int(ms[-1]), seed=301)
    _run_segmented(engine, mk, [np.arange(mk.size, dtype=np.uint32)], 1, ms)
    keys = workloads.uniform_keys(n, seed=302)
    vals = [workloads.uniform_keys(n, seed=303)]
    g.by_ptr[g.data.keys[0]][:n].copy_(to_dev(keys)[:n])
    g.by_ptr[g.data.values[0][0]][:n].copy_(to_dev(vals[0])[:n])
    torch.cuda.synchronize()
    plan.execute(g.data)
    torch.cuda.synchronize()
    engine.check_device()
    _check_segments(oracle, keys, vals, 1, starts, g.keys(), g.values(), 0, 32)
    plan.close()
    del big


@pytest.mark.parametrize("eb", [32, 24, 8])
def test_index_mode_needs_both_value_sides(engine, eb):
    """Index generation writes through values[0] on both sides, so a null
    side 0 is rejected before anything runs (sortArray accepts it only for a
    one-pass sort; the segmented calls always need it)."""
    import ctypes
    g = GpuSort(workloads.uniform_keys(10_000, seed=310), value_count=-1, end_bit=eb)
    g.data.values[0][0] = 0
    starts = (ctypes.c_int * 3)(0, 5000, 10_000)
    L = mg.lib()
    assert L.sortSegmentedArray(engine.handle, ctypes.byref(g.data), 2, starts) == \
        mg.SortStatus.INVALID_VALUE
    seg = mg.SortSegment()
    seg.keys[0], seg.keys[1] = g.data.keys[0], g.data.keys[1]
    seg.values[0][1] = g.data.values[0][1]
    seg.numElements = 10_000
    assert L.sortSegmentedBatch(engine.handle, ctypes.byref(seg), 1, -1, 0, eb) == \
        mg.SortStatus.INVALID_VALUE
    plan = ctypes.c_void_p()
    assert L.sortSegmentedPlanCreate(engine.handle, ctypes.byref(g.data), 2, starts,
                                     ctypes.byref(plan)) == mg.SortStatus.INVALID_VALUE


@pytest.mark.parametrize("layout", ["xcd", "one_chain", "small_pieces", "wave_count",
                                    "dominant_segment", "unpacked", "packed_xcd", "one_counter"])
def test_device_wide_layouts(oracle, layout):
    """The device-wide passes of a segmented sort in each layout: lookback
    chains per XCD with whole segments grouped by XCD (MGPUSORT_SEG_XCD=1),
    one chain per segment across the device (MGPUSORT_XCD=0; or a segment
    holding more than a quarter of an XCD's share of the tiles, even when
    asked for chains per XCD), count pieces smaller than a segment
    (MGPUSORT_SEG_WIDE_CHUNK=4096, one chain per segment:
    MGPUSORT_SEG_XCD=0; a histogram in global memory and a scan launch
    after the count), the per-wave count kernel instead of the lane-column
    one (MGPUSORT_SEG_COUNT=wave).  Packed tiles (segments back to back,
    MGPUSORT_SEG_PACK=1) with one chain per segment or chains per XCD, and
    tiles of each segment's own (MGPUSORT_SEG_PACK=0); with one chain per
    segment the tiles come from a counter per group of segments, or from
    one counter (MGPUSORT_SEG_GROUPS=1).  Every segment bit-exact against
    the oracle."""
    import torch
    env = {"MGPUSORT_XCD": "0"} if layout == "one_chain" else {}
    eng = make_engine(env)
    knobs = {"xcd": {"MGPUSORT_SEG_XCD": "1"},
             "one_chain": {"MGPUSORT_SEG_PACK": "1"},
             "dominant_segment": {"MGPUSORT_SEG_XCD": "1", "MGPUSORT_SEG_PACK": "1"},
             "small_pieces": {"MGPUSORT_SEG_WIDE_CHUNK": "4096", "MGPUSORT_SEG_XCD": "0",
                              "MGPUSORT_SEG_PACK": "1"},
             "wave_count": {"MGPUSORT_SEG_COUNT": "wave"},
             "unpacked": {"MGPUSORT_SEG_PACK": "0", "MGPUSORT_SEG_XCD": "0"},
             "packed_xcd": {"MGPUSORT_SEG_PACK": "1", "MGPUSORT_SEG_XCD": "1"},
             "one_counter": {"MGPUSORT_SEG_GROUPS": "1", "MGPUSORT_SEG_XCD": "0"}}.get(layout, {})
    saved = {k: os.environ.get(k) for k in knobs}
    try:
        os.environ.update(knobs)
        sizes = list(workloads.log_uniform_sizes(300, 12, 17, seed=320))
        if layout == "dominant_segment":
            sizes[7] = 1 << 22
        starts = _segments(sizes)
        n = int(starts[-1])
        keys = workloads.uniform_keys(n, seed=321)
        vals = [workloads.uniform_keys(n, seed=322)]
        for vc, fb, eb in [(1, 0, 32), (-1, 3, 27)]:
            vv = vals if vc == 1 else []
            eng.enable_timing(True)
            eng.read_timing()
            g = _run_segmented(eng, keys, vv, vc, starts, fb, eb)
            names = set(eng.read_timing())
            eng.enable_timing(False)
            _check_segments(oracle, keys, vv, vc, starts, g.keys(), g.values(), fb, eb)
            scatter = {k for k in names if k.startswith("mgpu_multi_")
                       and not k.startswith("mgpu_multi_count") and k != "mgpu_multi_scan"}
            assert scatter, names
            per_xcd = any("_x_" in k for k in scatter)
            xcds = eng.info()["xcdSegments"] > 1
            packed = {"one_chain": True, "small_pieces": True, "dominant_segment": True,
                      "unpacked": False, "packed_xcd": True}.get(layout)
            if packed is not None:
                assert all(k.startswith("mgpu_multi_packed" if packed else "mgpu_multi_scatter")
                           for k in scatter), names
            if layout in ("xcd", "packed_xcd") and xcds:
                assert per_xcd, names
            if layout in ("one_chain", "dominant_segment", "small_pieces", "unpacked"):
                assert not per_xcd, names
            assert ("mgpu_multi_count_cols" in names) == (layout != "wave_count"), names
            # Segments of at most 2^18 keys, no piece size asked for: the count
            # writes the digit starts itself.
            assert ("mgpu_multi_scan" in names) == (
                layout in ("small_pieces", "wave_count", "dominant_segment")), names
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        eng.close()


def test_seg_table_dimensions_honoured(tmp_path, oracle):
    """Every dimension of a segmented tuning-table row is honoured: digit
    width, keys per thread and workgroup size per size bucket (read from the
    names of the kernels that ran), and the wide row's routing threshold and
    tile shape (its count piece of 1024 keys cuts segments into many pieces);
    malformed rows and a 128-thread 8-bit variant (not built) are ignored.
    Every segment bit-exact against the oracle."""
    import shutil
    import torch
    for name in ("mgpusort_radix.hsaco", "mgpusort_seg.hsaco"):
        shutil.copy(os.path.join(mg.LIB_DIR, name), tmp_path / name)
    (tmp_path / "mgpusort_seg_tuning.txt").write_text(
        "# test table\n6 5 8 128\n8 8 8 128\n10 8 16 512\nnonsense\n31 7 16 256\n"
        "wide 2049 1024 large\n")
    eng = mg.Engine(str(tmp_path) + "/")
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        sizes = [1, 40, 64, 65, 700, 1024, 1025, 2048, 2049, 5000, 30_000, 3]
        starts = _segments(sizes)
        n = int(starts[-1])
        keys = workloads.uniform_keys(n, seed=330)
        vals = [workloads.uniform_keys(n, seed=331)]
        eng.enable_timing(True)
        eng.read_timing()
        g = _run_segmented(eng, keys, vals, 1, starts)
        names = set(eng.read_timing())
        eng.enable_timing(False)
        _check_segments(oracle, keys, vals, 1, starts, g.keys(), g.values(), 0, 32)
        # Kernels with or without the global-memory routine (_ng) count alike.
        seg_kernels = {k.removesuffix("_ng") for k in names if k.startswith("mgpu_segsort")}
        assert seg_kernels == {"mgpu_segsort_r5_i8_b128", "mgpu_segsort_r8_i16_b512",
                               "mgpu_segsort_r7_i16_b256"}, names
        scatter = {k for k in names if k.startswith("mgpu_multi_packed")}
        assert scatter and all("_b1024_" in k for k in scatter), names
    finally:
        eng.close()


@pytest.mark.parametrize("tile", ["small", "large"])
@pytest.mark.parametrize("xcd", ["0", "1"])
def test_packed_tiles_many_parts(tmp_path, oracle, tile, xcd):
    """Packed device-wide tiles holding many segments: with every segment
    routed device-wide and packed by a table row `wide 1 16384 <tile>
    packed`, segments of 1 to ~3000 keys fill tiles with as many parts as a
    tile may hold, so the layout pads tiles, starts parts mid-tile and ends
    segments on partial chunks; chains per XCD too (MGPUSORT_SEG_XCD=1).
    Keys only, index generation, one and six value arrays, over several bit
    ranges; every segment bit-exact against the oracle."""
    import shutil
    import torch
    for name in ("mgpusort_radix.hsaco", "mgpusort_seg.hsaco"):
        shutil.copy(os.path.join(mg.LIB_DIR, name), tmp_path / name)
    (tmp_path / "mgpusort_seg_tuning.txt").write_text(f"wide 1 16384 {tile} packed\n")
    saved = {k: os.environ.get(k) for k in ("MGPUSORT_SEG_XCD", "MGPUSORT_SEG_PACK")}
    os.environ["MGPUSORT_SEG_XCD"] = xcd
    os.environ.pop("MGPUSORT_SEG_PACK", None)
    eng = mg.Engine(str(tmp_path) + "/")
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        rng = np.random.default_rng(340 + int(xcd) + (tile == "large") * 2)
        sizes = [int(x) for x in rng.choice([1, 2, 5, 63, 64, 65, 1023, 1024, 1025, 2047, 2049, 3000],
                                            size=400)]
        sizes[10:14] = [0, 0, 9_000, 0]
        starts = _segments(sizes)
        n = int(starts[-1])
        keys = workloads.uniform_keys(n, seed=341)
        vals = [workloads.uniform_keys(n, seed=342 + j) for j in range(6)]
        for vc, fb, eb in [(0, 0, 32), (-1, 4, 20), (1, 0, 32), (6, 8, 32), (1, 0, 7)]:
            vv = [] if vc in (0, -1) else vals[:vc]
            eng.enable_timing(True)
            eng.read_timing()
            g = _run_segmented(eng, keys, vv, vc, starts, fb, eb)
            names = set(eng.read_timing())
            eng.enable_timing(False)
            _check_segments(oracle, keys, vv, vc, starts, g.keys(), g.values(), fb, eb)
            packed = {k for k in names if k.startswith("mgpu_multi_packed")}
            assert packed, names
            assert not any(k.startswith("mgpu_segsort") for k in names), names
            if xcd == "1" and eng.info()["xcdSegments"] > 1:
                assert all("_x_" in k for k in packed), names
            if xcd == "0":
                assert not any("_x_" in k for k in packed), names
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        eng.close()
