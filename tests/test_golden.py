"""The committed golden fixtures (tests/golden/, written by
tests/golden/make_fixtures.py) against this machine: the generators rebuild
the same inputs, and the oracle and numpy give the same stable permutations.
The full-size 2^28 workloads are checked on the GPU only."""
import json
import os

import numpy as np
import pytest

import golden_cases as gc
from conftest import bitfield

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def manifest():
    with open(os.path.join(HERE, "manifest.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def perms():
    with np.load(os.path.join(HERE, "sort_small.npz")) as z:
        return {k: z[k] for k in z.files}


def test_manifest_lists_every_case(manifest, perms):
    assert [c["name"] for c in manifest["small"]] == [c[0] for c in gc.SMALL]
    assert [c["name"] for c in manifest["large"]] == [c[0] for c in gc.LARGE]
    assert sorted(perms) == sorted(c[0] for c in gc.SMALL)


@pytest.mark.parametrize("case", gc.SMALL, ids=[c[0] for c in gc.SMALL])
def test_small_case(case, manifest, perms, oracle):
    name, n, kind, seed, vc, fb, eb = case
    rec = next(c for c in manifest["small"] if c["name"] == name)
    k = gc.keys(kind, n, seed)
    assert gc.sha256(k) == rec["input_keys_sha256"]
    perm = perms[name].astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(n))
    _, (got,) = oracle.stable_sort(k, [], value_count=-1, first_bit=fb, end_bit=eb)
    assert np.array_equal(got, perm)
    assert np.array_equal(np.argsort(bitfield(k, fb, eb), kind="stable"), perm)


@pytest.mark.parametrize("case", [c for c in gc.LARGE if c[1] <= 1 << 25],
                         ids=lambda c: c[0])
def test_large_case_up_to_2p25(case, manifest):
    name, n, kind, seed, vc, fb, eb = case
    rec = next(c for c in manifest["large"] if c["name"] == name)
    k = gc.keys(kind, n, seed)
    assert gc.sha256(k) == rec["input_keys_sha256"]
    assert [int(x) for x in k[:8]] == rec["input_keys_head"]
    perm = np.argsort(bitfield(k, fb, eb), kind="stable")
    out = k[perm]
    assert gc.sha256(out) == rec["output_keys_sha256"]
    assert [int(x) for x in out[:64]] == rec["output_keys_head"]
    if vc != 0:
        assert gc.sha256(perm.astype(np.uint32)) == rec["output_values_sha256"]
