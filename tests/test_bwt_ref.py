"""Pins the BWT checker (tests/bwt_ref.py) before the GPU BWT test relies on
it: published transforms, brute force over every rotation, the inverse."""
import numpy as np
import pytest

import bwt_ref
from mgpu_old_amd import workloads


def brute_bwt(block: bytes):
    n = len(block)
    rows = sorted(range(n), key=lambda i: (block[i:] + block[:i], i))
    return bytes(block[(i - 1) % n] for i in rows), rows.index(0), rows


@pytest.mark.parametrize("block, transform, primary", [
    (b"abraca", b"caraab", 1),            # Burrows & Wheeler 1994, the worked example
    (b"banana", b"nnbaaa", 3),
    (b"mississippi", b"pssmipissii", 4),
    (b"abracadabra", b"rdarcaaaabb", 2),
])
def test_published_transforms(block, transform, primary):
    assert bwt_ref.bwt(block) == (transform, primary)
    assert bwt_ref.inverse_bwt(transform, primary) == block


def _blocks():
    out = [b"a", b"ab", b"ba", b"aaaa", b"abab" * 20, b"abcabcabd" * 9, b"zyx" * 33 + b"w"]
    for seed, (n, alphabet) in enumerate([(50, 1), (97, 2), (200, 2), (256, 4), (300, 26),
                                          (128, 256), (255, 3)]):
        w = workloads.words(900 + seed, 0, n) % alphabet
        out.append((w + 97).astype(np.uint8).tobytes() if alphabet <= 26 else
                   w.astype(np.uint8).tobytes())
    return out


@pytest.mark.parametrize("block", _blocks(), ids=lambda b: f"n{len(b)}")
def test_matches_brute_force(block):
    transform, primary, rows = brute_bwt(block)
    assert bwt_ref.bwt(block) == (transform, primary)
    rank = bwt_ref.rotation_ranks(block)
    # Equal ranks exactly for equal rotations, order as the rotations.
    rot = [block[i:] + block[:i] for i in range(len(block))]
    for a, b in zip(rows, rows[1:]):
        assert (rank[a] == rank[b]) == (rot[a] == rot[b]) and rank[a] <= rank[b]
    assert list(bwt_ref.rotation_order(block)) == rows
    assert bwt_ref.inverse_bwt(transform, primary) == block


def test_inverse_catches_a_wrong_transform():
    block = b"the whale, the whale!"
    transform, primary = bwt_ref.bwt(block)
    swapped = bytearray(transform)
    i = next(k for k in range(len(swapped) - 1) if swapped[k] != swapped[k + 1])
    swapped[i], swapped[i + 1] = swapped[i + 1], swapped[i]
    assert bwt_ref.inverse_bwt(bytes(swapped), primary) != block
