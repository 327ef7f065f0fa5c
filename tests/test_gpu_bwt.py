"""The BWT block sort of the reference's BWT library, driven through the C ABI
as that library drives sortArray (bwt/src/mgpubwt/mgpubwt.cpp:226-371,
bwtSortBlock; its test bwt/src/bwttest/bwttest.cpp:46-81 sorts the Moby Dick
text at every key size 1..24).  Needs a gfx950 GPU.

The client gathers, for every position of the block, its next gpuKeySize
symbols of the doubled block as big-endian 32-bit words (GatherBWTKeys,
bwt/src/kernels/bwt.cu:24-114; here on the host), sorts the word streams
least significant first, each sortArray carrying the words still unsorted
and the position index as values (up to 6) over bits [32 - 8 * bytes, 32),
with the client keeping the ping-pong parity itself; then positions whose
key bytes tie are ordered on the host by the rest of their rotation (as the
reference's quicksort does) and transform[r] = the symbol before row r's
rotation.

Checks: after the GPU sorts the positions are exactly the stable order by
the first gpuKeySize symbols; the transform equals the CPU checker's BWT;
adjacent rows are in rotation order (the reference test's own check); the
inverse transform gives back the block.  Blocks below the reference's
TinyBlockCutoff (3000) never reach the GPU there and are not used here."""
import ctypes

import numpy as np
import pytest

import bwt_ref
import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import make_engine, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    eng = make_engine()
    yield eng
    eng.close()


def _moby_block():
    """The excerpt joined as bwttest.cpp:33-44 joins the whole book: each
    non-empty line plus a space, a newline for each empty line."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "mobydick_excerpt.txt")
    out = bytearray()
    with open(path, "rb") as f:
        for line in f.read().split(b"\n")[:-1]:
            out += line + b" " if line else b"\n"
    return bytes(out)


def _gather(doubled, count, num_streams):
    b = np.frombuffer(doubled, dtype=np.uint8).astype(np.uint32)
    pos = np.arange(count)
    return [(b[pos + 4 * w] << 24) | (b[pos + 4 * w + 1] << 16) | (b[pos + 4 * w + 2] << 8) |
            b[pos + 4 * w + 3] for w in range(num_streams)]


def _tie_sort(run, doubled, start, end_offset):
    """Orders the positions of `run` by symbols [p + start, p + end_offset)
    of the doubled block, 64 at a time; equal rotations keep their order."""
    out = sorted(run, key=lambda p: doubled[p + start:p + min(start + 64, end_offset)])
    if start + 64 >= end_offset:
        return out
    res, i = [], 0
    while i < len(out):
        j = i + 1
        key = doubled[out[i] + start:out[i] + start + 64]
        while j < len(out) and doubled[out[j] + start:out[j] + start + 64] == key:
            j += 1
        res += _tie_sort(out[i:j], doubled, start + 64, end_offset) if j - i > 1 else out[i:j]
        i = j
    return res


def bwt_sort_block(engine, block, gpu_key_size):
    """bwtSortBlock: (transform, indices, segCount, GPU-sorted positions)."""
    import torch
    count = len(block)
    num_streams = -(-gpu_key_size // 4)
    doubled = block + block
    words = _gather(doubled + bytes(32), count, num_streams)
    cap = -(-count // 2048) * 2048  # AllocSortMem: buffers rounded to 2048
    streams = [[torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(2)]
               for _ in range(num_streams + 1)]
    for w in range(num_streams):
        streams[w][0][:count].copy_(torch.from_numpy(words[w].view(np.int32)))
    streams[num_streams][0][:count].copy_(torch.arange(count, dtype=torch.int32))
    index = streams[num_streams]
    by_ptr = {t.data_ptr(): t for pair in streams for t in pair}

    d = mg.SortData()
    d.maxElements = d.numElements = count
    d.preserveEndKeys = d.earlyExit = False
    d.parity = 0
    torch.cuda.synchronize()
    for cur in range(num_streams - 1, -1, -1):
        p = d.parity
        d.keys[0], d.keys[1] = streams[cur][p].data_ptr(), streams[cur][1 - p].data_ptr()
        for i in range(cur):
            d.values[i][0], d.values[i][1] = streams[i][p].data_ptr(), streams[i][1 - p].data_ptr()
        d.values[cur][0], d.values[cur][1] = index[p].data_ptr(), index[1 - p].data_ptr()
        d.valueCount = cur + 1
        cur_bytes = min(gpu_key_size - 4 * cur, 4)
        d.firstBit, d.endBit = 32 - 8 * cur_bytes, 32
        st = mg.lib().sortArray(engine.handle, ctypes.byref(d))
        assert st == 0, f"sortArray on stream {cur}: {mg.status_string(st)}"
    torch.cuda.synchronize()
    gpu_order = to_host(by_ptr[d.values[0][0]], count).astype(np.int64)
    first_words = to_host(by_ptr[d.keys[0]], count)
    assert d.values[0][0] == index[d.parity].data_ptr()

    # Head flags: the first gpuKeySize symbols differ from the row before.
    last_mask = np.uint32(0xFFFFFFFF)
    if gpu_key_size & 3:
        last_mask = np.uint32((0xFFFFFFFF << (8 * (4 - (gpu_key_size & 3)))) & 0xFFFFFFFF)
    masked = [w[gpu_order] for w in words]
    masked[-1] = masked[-1] & last_mask
    head = np.ones(count + 1, dtype=bool)
    head[1:count] = np.any([m[1:] != m[:-1] for m in masked], axis=0)
    indices = gpu_order.copy()
    seg_count = 0
    starts = np.flatnonzero(head)
    for a, b in zip(starts[:-1], starts[1:]):
        if b - a >= 2:
            seg_count += 1
            indices[a:b] = _tie_sort(list(gpu_order[a:b]), doubled, gpu_key_size, count)
    src = np.frombuffer(doubled, dtype=np.uint8)
    transform = src[indices + count - 1].tobytes()
    return transform, indices, seg_count, gpu_order, first_words, words


def _check(engine, block, key_sizes):
    count = len(block)
    rank = bwt_ref.rotation_ranks(block)
    want, primary = bwt_ref.bwt(block)
    for ks in key_sizes:
        transform, indices, segs, gpu_order, first_words, words = bwt_sort_block(engine, block, ks)
        # The GPU part: positions stable-sorted by their first ks symbols.
        keys = [w.copy() for w in words]
        if ks & 3:
            keys[-1] &= np.uint32((0xFFFFFFFF << (8 * (4 - (ks & 3)))) & 0xFFFFFFFF)
        expect = np.lexsort(tuple(reversed(keys)))
        assert np.array_equal(gpu_order, expect), f"key size {ks}: GPU order differs"
        assert np.array_equal(first_words, words[0][expect]), f"key size {ks}: word 0 differs"
        # The whole block sort.
        assert np.array_equal(np.sort(indices), np.arange(count))
        assert np.all(rank[indices[1:]] >= rank[indices[:-1]]), f"key size {ks}: rows out of order"
        assert transform == want, f"key size {ks} ({segs} tie segments): transform differs"
        row0 = int(np.flatnonzero(indices == 0)[0])
        assert bwt_ref.inverse_bwt(transform, row0) == block


def test_bwt_moby_dick_every_key_size(engine):
    block = _moby_block()
    assert len(block) > 40_000
    _check(engine, block, range(1, 25))


def test_bwt_binary_alphabet(engine):
    w = workloads.words(77, 0, 20_000) & 1
    _check(engine, (w + ord("a")).astype(np.uint8).tobytes(), [4, 13, 24])


def test_bwt_periodic_block(engine):
    """Identical rotations: tie segments that no key size resolves."""
    _check(engine, b"abracadabra" * 400, [3, 11, 24])
