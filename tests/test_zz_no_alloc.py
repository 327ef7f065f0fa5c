"""sortArrayNoAlloc through the C ABI (needs a gfx950 GPU).

The entry point allocates nothing: its scratch is a fixed array in the
library's device code.  Its output must equal sortArray's on the same inputs
bit for bit (keys, every value array or generated indices, parity and buffer
orientation), on consecutive calls whose sizes shrink (the array keeps the
last call's contents), and sizes above SORT_NOALLOC_MAX_ELEMENTS are refused
before anything runs.
"""
import ctypes

import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import GpuSort, make_engine

pytestmark = pytest.mark.gpu

MAX_ELEMENTS = 1 << 24  # SORT_NOALLOC_MAX_ELEMENTS


@pytest.fixture(scope="module")
def engine():
    eng = make_engine()
    yield eng
    eng.close()


def _no_alloc(engine, data):
    L = mg.lib()
    L.sortArrayNoAlloc.restype = ctypes.c_int
    L.sortArrayNoAlloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(mg.SortData)]
    return L.sortArrayNoAlloc(engine.handle, ctypes.byref(data))


def _inputs(n, vc, seed):
    keys = workloads.uniform_keys(n, seed=seed)
    vals = [workloads.uniform_keys(n, seed=seed + 1 + j) for j in range(max(vc, 0))]
    return keys, vals


@pytest.mark.parametrize("cases", [
    # consecutive calls, each smaller than the one before
    [(MAX_ELEMENTS, 1, 0, 32), ((1 << 20) + 3, 0, 0, 32), (300_001, -1, 0, 24),
     (5000, 2, 3, 20), (1, 1, 0, 32)],
    [((1 << 22) + 5, -1, 0, 32), (70_000, 6, 5, 29)],
])
def test_no_alloc_matches_sort_array(engine, oracle, cases):
    import torch
    for i, (n, vc, fb, eb) in enumerate(cases):
        keys, vals = _inputs(n, vc, seed=500 + 10 * i)
        ref = GpuSort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb).run(engine)
        got = GpuSort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb)
        torch.cuda.synchronize()
        assert _no_alloc(engine, got.data) == mg.SortStatus.SUCCESS, f"case {i}: n={n}"
        torch.cuda.synchronize()
        engine.check_device()
        assert got.data.parity == ref.data.parity
        assert (got.data.keys[0] == got.start_keys) == (ref.data.keys[0] == ref.start_keys)
        assert np.array_equal(got.keys(), ref.keys()), f"case {i}: keys differ from sortArray"
        for j, (a, b) in enumerate(zip(got.values(), ref.values())):
            assert np.array_equal(a, b), f"case {i}: values[{j}] differ from sortArray"
        if n <= (1 << 20) + 3:
            rk, rv = oracle.stable_sort(keys, vals, value_count=vc, first_bit=fb, end_bit=eb)
            assert np.array_equal(got.keys(), rk)
            for a, b in zip(got.values(), rv):
                assert np.array_equal(a, b)


def test_no_alloc_refuses_beyond_maximum(engine):
    import torch
    n = MAX_ELEMENTS + 1
    keys = workloads.uniform_keys(n, seed=600)
    g = GpuSort(keys)
    before = (g.data.keys[0], g.data.keys[1], g.data.parity)
    torch.cuda.synchronize()
    assert _no_alloc(engine, g.data) == mg.SortStatus.CONFIG_NOT_SUPPORTED
    assert (g.data.keys[0], g.data.keys[1], g.data.parity) == before
    assert np.array_equal(g.keys(), keys), "a refused call must leave the keys alone"
    # and the entry point still works afterwards
    small = workloads.uniform_keys(4097, seed=601)
    g2 = GpuSort(small)
    assert _no_alloc(engine, g2.data) == mg.SortStatus.SUCCESS
    torch.cuda.synchronize()
    assert np.array_equal(g2.keys(), np.sort(small))
