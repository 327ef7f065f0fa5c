"""CPU checker for the BWT client test (test infrastructure): the
Burrows-Wheeler transform of a block over its cyclic rotations, the output of
bwtSortBlock (bwt/src/mgpubwt/mgpubwt.cpp:226-371: rows = rotations of the
block in lexicographic order, transform[r] = last symbol of row r), from a
cyclic suffix array built by prefix doubling, and the inverse transform."""
import numpy as np


def rotation_ranks(block: bytes) -> np.ndarray:
    """rank[i]: rank of rotation i of `block` among all its rotations, equal
    rotations sharing a rank.  Prefix doubling: ranks by the first `length`
    symbols, then by pairs (rank of i, rank of i + length) for 2 * length,
    until the prefixes cover a whole rotation or every rank is distinct."""
    s = np.frombuffer(block, dtype=np.uint8).astype(np.int64)
    n = s.size
    if n == 0:
        return s
    idx = np.arange(n)
    _, rank = np.unique(s, return_inverse=True)
    length = 1
    while length < n and rank.max() < n - 1:
        second = rank[(idx + length) % n]
        order = np.lexsort((second, rank))
        r, q = rank[order], second[order]
        step = np.concatenate(([0], ((r[1:] != r[:-1]) | (q[1:] != q[:-1])).astype(np.int64)))
        rank = np.empty(n, dtype=np.int64)
        rank[order] = np.cumsum(step)
        length *= 2
    return rank


def rotation_order(block: bytes) -> np.ndarray:
    """Rotation start positions in sorted order (equal rotations by position)."""
    rank = rotation_ranks(block)
    return np.lexsort((np.arange(rank.size), rank))


def bwt(block: bytes):
    """(transform, primary): the last symbols of the sorted rotations and the
    row that holds the block itself."""
    order = rotation_order(block)
    s = np.frombuffer(block, dtype=np.uint8)
    last = s[(order + s.size - 1) % s.size]
    return last.tobytes(), int(np.flatnonzero(order == 0)[0])


def inverse_bwt(transform: bytes, primary: int) -> bytes:
    """The block whose rotation sits at row `primary`.  Row r's last symbol
    precedes its rotation's first; the row of that preceding rotation is
    LF(r) = the position of r in a stable sort of the last column."""
    last = np.frombuffer(transform, dtype=np.uint8)
    n = last.size
    if n == 0:
        return b""
    lf = np.empty(n, dtype=np.int64)
    lf[np.argsort(last, kind="stable")] = np.arange(n)
    out = np.empty(n, dtype=np.uint8)
    r = primary
    for k in range(n - 1, -1, -1):
        out[k] = last[r]
        r = lf[r]
    return out.tobytes()
