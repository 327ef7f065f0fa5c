"""Every built rank_scatter tile shape, with either in-wave ranking method and
either lookback layout, gives the same bit-exact result as the oracle, for
plain and multi-array (segmented large) sorts.  Read at engine creation:
MGPUSORT_TILE forces one shape for every size; MGPUSORT_RANK one ranking
method (0: LDS match masks; 1: one LDS atomic add per key, which the engine
otherwise takes only after checking that the device orders same-word lanes);
MGPUSORT_XCD=0 keeps one lookback chain across the device, MGPUSORT_XCD=1
takes one per XCD segment at every size (the engine otherwise takes them,
after its own check, for large sorts only)."""
import os

import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import GpuSort, make_engine

pytestmark = pytest.mark.gpu

VARIANTS = [(t, r, x) for x in ("1", "0") for r in ("0", "1")
            for t in ("256x16", "512x16", "1024x16")]
ENV = ("MGPUSORT_TILE", "MGPUSORT_RANK", "MGPUSORT_XCD")


@pytest.fixture(params=VARIANTS, ids=[f"{t}_r{r}_xcd{x}" for t, r, x in VARIANTS])
def variant_engine(request):
    tile, rank, xcd = request.param
    saved = {k: os.environ.get(k) for k in ENV}
    os.environ.update(MGPUSORT_TILE=tile, MGPUSORT_RANK=rank, MGPUSORT_XCD=xcd)
    try:
        eng = make_engine()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    yield eng
    eng.close()


@pytest.mark.parametrize("n", [1, 5000, 16384, 16385, 1_000_003, 4_194_305])
@pytest.mark.parametrize("vc", [-1, 1, 3])
def test_variant_matches_oracle(variant_engine, oracle, n, vc):
    keys = workloads.and3_keys(n, seed=n + vc)
    if n > 5000:  # a run of equal digits and a stretch of one repeated key
        keys[100:1100] &= 0xFFFF00FF
        keys[2000:2600] = keys[2000]
    vals = [workloads.uniform_keys(n, seed=300 + j) for j in range(max(vc, 0))]
    g = GpuSort(keys, vals, value_count=vc).run(variant_engine)
    variant_engine.check_device()
    rk, rv = oracle.stable_sort(keys, vals, value_count=vc)
    assert np.array_equal(g.keys(), rk)
    for o, r in zip(g.values(), rv):
        assert np.array_equal(o, r)


def test_variant_multi_array_path(variant_engine, oracle):
    """Segments above 2^16 keys take the multi-array device-wide passes."""
    import torch
    sizes = [70_000, 3, 200_001, 65_537, 1000]
    starts = np.zeros(len(sizes) + 1, dtype=np.int64)
    starts[1:] = np.cumsum(sizes)
    n = int(starts[-1])
    keys = workloads.uniform_keys(n, seed=9)
    vals = [np.arange(n, dtype=np.uint32)]
    g = GpuSort(keys, vals, end_bit=24)
    torch.cuda.synchronize()
    variant_engine.sort_segmented(g.data, starts)
    variant_engine.check_device()
    out_k, (out_v,) = g.keys(), g.values()
    for i in range(len(sizes)):
        a, b = int(starts[i]), int(starts[i + 1])
        rk, (rv,) = oracle.stable_sort(keys[a:b], [vals[0][a:b]], first_bit=0, end_bit=24)
        assert np.array_equal(out_k[a:b], rk) and np.array_equal(out_v[a:b], rv)
    assert g.data.parity == len(mg.plan_passes(24)) % 2
