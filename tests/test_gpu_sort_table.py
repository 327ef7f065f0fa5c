"""sortArray's tuning table (kernelPath + "mgpusort_sort_tuning.txt"): an
engine created on a directory holding a table runs the tile shape and
lookback layout the table names for each value class, ignores malformed
lines, and sorts bit-exactly either way.  Needs a gfx950 GPU."""
import os
import shutil

import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import GpuSort

pytestmark = pytest.mark.gpu


def _engine_with_table(tmp_path, text):
    lib = mg.LIB_DIR
    for name in ("mgpusort_radix.hsaco", "mgpusort_seg.hsaco"):
        shutil.copy(os.path.join(lib, name), tmp_path / name)
    (tmp_path / "mgpusort_sort_tuning.txt").write_text(text)
    import torch
    eng = mg.Engine(str(tmp_path) + "/")
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    return eng


def _scatter_kernels(eng, keys, vals, vc):
    eng.enable_timing(True)
    g = GpuSort(keys, vals, value_count=vc).run(eng)
    names = [k for k in eng.read_timing() if k.startswith("mgpu_rank_scatter")]
    eng.enable_timing(False)
    return g, names


def test_table_picks_shape_and_chains(tmp_path, oracle):
    eng = _engine_with_table(tmp_path, "\n".join([
        "# comment",
        "31 keys large one",
        "18 kv large xcd",
        "31 kv small one",
        "garbage line",
        "20 index huge one",      # bad tile: ignored
        "31 index large xcd",
    ]) + "\n")
    try:
        if eng.info()["xcdSegments"] < 2:
            pytest.skip("no per-XCD lookback chains on this device")
        n = 1 << 20
        keys = workloads.uniform_keys(n, 61)
        vals = workloads.uniform_keys(n, 62)
        rk, (rv,) = oracle.stable_sort(keys, [vals])
        # keys only: the large tile and one chain, where the built-in rules
        # give the small tile (1024 x 16 from 2^21 keys).
        g, names = _scatter_kernels(eng, keys, [], 0)
        assert names and all("_b1024_" in k and "_x_" not in k for k in names), names
        assert np.array_equal(g.keys(), rk)
        # one value, 2^20 > 2^18: the second kv row (small, one chain).
        g, names = _scatter_kernels(eng, keys, [vals], 1)
        assert names and all("_b512_" in k and "_x_" not in k for k in names), names
        assert np.array_equal(g.keys(), rk) and np.array_equal(g.values()[0], rv)
        # one value, 2^18: the first kv row (large tile, chains per XCD).
        m = 1 << 18
        rk2, (rv2,) = oracle.stable_sort(keys[:m], [vals[:m]])
        g, names = _scatter_kernels(eng, keys[:m], [vals[:m]], 1)
        assert names and all("_b1024_" in k and "_x_" in k for k in names), names
        assert np.array_equal(g.keys(), rk2) and np.array_equal(g.values()[0], rv2)
        # index generation: the malformed row is skipped, the valid one applies.
        g, names = _scatter_kernels(eng, keys, None, -1)
        assert names and all("_b1024_" in k and "_x_" in k for k in names), names
        assert np.array_equal(g.keys(), rk)
        assert np.array_equal(g.values()[0], np.argsort(keys, kind="stable").astype(np.uint32))
    finally:
        eng.close()


def test_unreadable_table_keeps_builtin_rules(tmp_path, oracle):
    eng = _engine_with_table(tmp_path, "nothing valid here\n7 kv\n")
    try:
        n = 1 << 20
        keys = workloads.uniform_keys(n, 63)
        rk, _ = oracle.stable_sort(keys)
        g, names = _scatter_kernels(eng, keys, [], 0)
        # Built-in rule at 2^20 keys: 1024 x 16 from 2^21, so the small tile.
        assert names and all("_b512_" in k for k in names), names
        assert np.array_equal(g.keys(), rk)
    finally:
        eng.close()
