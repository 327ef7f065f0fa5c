"""Writes the sort's golden fixtures from the oracle (oracle/oracle.cpp,
std::stable_sort of (bit field, index) pairs), each result cross-checked
against numpy's stable argsort of the bit field:

  tests/golden/sort_small.npz   the stable permutation of every small case
                                (uint16; outputs = inputs gathered through it)
  tests/golden/manifest.json    per case: parameters and the SHA-256 of the
                                input keys; per full-size workload also the
                                SHA-256 of the sorted keys and values and
                                their first and last 64 elements

Inputs are regenerated from (kind, n, seed) by tests/golden_cases.py.
Run from the repository root: python tests/golden/make_fixtures.py"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import golden_cases as gc  # noqa: E402
from conftest import Oracle, bitfield  # noqa: E402


def stable_perm(oracle, k, first_bit, end_bit):
    """The oracle's index-generation sort, checked against numpy."""
    _, (perm,) = oracle.stable_sort(k, [], value_count=-1, first_bit=first_bit, end_bit=end_bit)
    ref = np.argsort(bitfield(k, first_bit, end_bit), kind="stable")
    assert np.array_equal(perm, ref.astype(np.uint32)), "oracle and numpy disagree"
    return perm


def main():
    oracle = Oracle()
    manifest = {"generator": "tests/golden/make_fixtures.py", "small": [], "large": []}
    perms = {}
    for name, n, kind, seed, vc, fb, eb in gc.SMALL:
        k = gc.keys(kind, n, seed)
        perm = stable_perm(oracle, k, fb, eb)
        vals = gc.values(n, seed, vc)
        if vals:  # the oracle's own value path agrees with the permutation
            ok, outs = oracle.stable_sort(k, vals, value_count=len(vals), first_bit=fb, end_bit=eb)
            assert np.array_equal(ok, k[perm]) and all(
                np.array_equal(o, v[perm]) for o, v in zip(outs, vals))
        perms[name] = perm.astype(np.uint16)
        manifest["small"].append({"name": name, "n": n, "kind": kind, "seed": seed,
                                  "value_count": vc, "first_bit": fb, "end_bit": eb,
                                  "input_keys_sha256": gc.sha256(k)})
    np.savez_compressed(os.path.join(HERE, "sort_small.npz"), **perms)
    for name, n, kind, seed, vc, fb, eb in gc.LARGE:
        t = time.time()
        k = gc.keys(kind, n, seed)
        perm = stable_perm(oracle, k, fb, eb)
        out_k = k[perm]
        rec = {"name": name, "n": n, "kind": kind, "seed": seed, "value_count": vc,
               "first_bit": fb, "end_bit": eb, "input_keys_sha256": gc.sha256(k),
               "input_keys_head": [int(x) for x in k[:8]],
               "output_keys_sha256": gc.sha256(out_k),
               "output_keys_head": [int(x) for x in out_k[:64]],
               "output_keys_tail": [int(x) for x in out_k[-64:]]}
        if vc != 0:  # index generation and 0..n-1 values both come out as perm
            rec["output_values_sha256"] = gc.sha256(perm)
            rec["output_values_head"] = [int(x) for x in perm[:64]]
            rec["output_values_tail"] = [int(x) for x in perm[-64:]]
        manifest["large"].append(rec)
        del k, perm, out_k
        print(f"{name}: {time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
