"""The GPU sort against the committed golden fixtures (tests/golden/, from the
oracle): every small case bit-exact, and the full-size benchmark workloads
(2^20 .. 2^28 keys) by the SHA-256 of their sorted keys and values and their
first and last 64 elements.  Needs a gfx950 GPU."""
import json
import os

import numpy as np
import pytest

import golden_cases as gc
import mgpu_old_amd as mg
from gpu_helpers import GpuSort, make_engine, to_host

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def engine():
    eng = make_engine()
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def manifest():
    with open(os.path.join(HERE, "manifest.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def perms():
    with np.load(os.path.join(HERE, "sort_small.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", gc.SMALL, ids=[c[0] for c in gc.SMALL])
def test_small_case(case, engine, perms):
    name, n, kind, seed, vc, fb, eb = case
    k = gc.keys(kind, n, seed)
    vals = gc.values(n, seed, vc)
    perm = perms[name].astype(np.int64)
    g = GpuSort(k, vals, value_count=vc, first_bit=fb, end_bit=eb).run(engine)
    assert np.array_equal(g.keys(), k[perm])
    if vc == -1:
        assert np.array_equal(g.values()[0], perm.astype(np.uint32))
    for got, v in zip(g.values(), vals):
        assert np.array_equal(got, v[perm])


def _sha(t, n):
    return gc.sha256(to_host(t, n))


@pytest.mark.parametrize("case", gc.LARGE, ids=[c[0] for c in gc.LARGE])
def test_full_size_workload(case, engine, manifest):
    import torch
    name, n, kind, seed, vc, fb, eb = case
    rec = next(c for c in manifest["large"] if c["name"] == name)
    print(f"[golden] {name}", flush=True)
    k0 = gc.keys_torch(kind, n, seed, torch)
    assert _sha(k0, n) == rec["input_keys_sha256"]
    k1 = torch.empty_like(k0)
    d = mg.SortData()
    d.maxElements = d.numElements = n
    d.firstBit, d.endBit = fb, eb
    d.keys[0], d.keys[1] = k0.data_ptr(), k1.data_ptr()
    tensors = [k0, k1]
    if vc != 0:
        v0 = (torch.arange(n, dtype=torch.int32, device="cuda") if vc == "iota"
              else torch.empty(n, dtype=torch.int32, device="cuda"))
        v1 = torch.empty_like(v0)
        d.values[0][0], d.values[0][1] = v0.data_ptr(), v1.data_ptr()
        tensors += [v0, v1]
    d.valueCount = 1 if vc == "iota" else vc
    by_ptr = {t.data_ptr(): t for t in tensors}
    torch.cuda.synchronize()
    engine.sort_array(d)
    torch.cuda.synchronize()
    out_k = to_host(by_ptr[d.keys[0]], n)
    assert [int(x) for x in out_k[:64]] == rec["output_keys_head"]
    assert [int(x) for x in out_k[-64:]] == rec["output_keys_tail"]
    assert gc.sha256(out_k) == rec["output_keys_sha256"]
    del out_k
    if vc != 0:
        out_v = to_host(by_ptr[d.values[0][0]], n)
        assert [int(x) for x in out_v[:64]] == rec["output_values_head"]
        assert [int(x) for x in out_v[-64:]] == rec["output_values_tail"]
        assert gc.sha256(out_v) == rec["output_values_sha256"]
    del tensors, by_ptr, k0, k1
    torch.cuda.empty_cache()
