import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class Oracle:
    """ctypes view of oracle/liboracle.so (the CPU checker)."""

    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.POINTER
        u32pp = P(P(ctypes.c_uint32))
        L.oracle_splitmix64_at.restype = ctypes.c_uint64
        L.oracle_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_key_transform.restype = ctypes.c_int
        L.oracle_key_transform.argtypes = [P(ctypes.c_uint32), ctypes.c_uint64, ctypes.c_int,
                                           ctypes.c_int, P(ctypes.c_uint32)]
        L.oracle_fill_keys.restype = None
        L.oracle_fill_keys.argtypes = [P(ctypes.c_uint32), ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_int, ctypes.c_int]
        L.oracle_stable_sort.restype = ctypes.c_int
        L.oracle_stable_sort.argtypes = [P(ctypes.c_uint32), u32pp, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.c_int, ctypes.c_int, P(ctypes.c_uint32), u32pp]
        L.oracle_sort1.restype = ctypes.c_int
        L.oracle_sort1.argtypes = [P(ctypes.c_uint32), u32pp, ctypes.c_int, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_uint32),
                                   u32pp]
        L.oracle_sort2.restype = ctypes.c_int
        L.oracle_sort2.argtypes = [P(ctypes.c_uint32), u32pp, ctypes.c_int, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                   ctypes.c_int, P(ctypes.c_uint32), u32pp]
        L.oracle_block_counts.restype = ctypes.c_int
        L.oracle_block_counts.argtypes = [P(ctypes.c_uint32), ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_uint64, P(ctypes.c_uint32)]
        L.oracle_digit_offsets.restype = ctypes.c_int
        L.oracle_digit_offsets.argtypes = [P(ctypes.c_uint32), ctypes.c_uint64, ctypes.c_int,
                                           P(ctypes.c_uint64)]
        self.L = L

    @staticmethod
    def _ptr_array(arrays):
        arr = (ctypes.POINTER(ctypes.c_uint32) * max(1, len(arrays)))()
        for i, a in enumerate(arrays):
            arr[i] = _u32p(a)
        return arr

    def _run(self, fn, keys, values, value_count, first_bit, end_bit, *extra):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        values = [np.ascontiguousarray(v, dtype=np.uint32) for v in (values or [])]
        n = keys.size
        out_k = np.empty(n, dtype=np.uint32)
        nout = 1 if value_count == -1 else max(0, value_count)
        outs = [np.empty(n, dtype=np.uint32) for _ in range(nout)]
        rc = fn(_u32p(keys), self._ptr_array(values), value_count, n, first_bit, end_bit,
                *extra, _u32p(out_k), self._ptr_array(outs))
        if rc != 0:
            raise ValueError("oracle rejected the arguments")
        return out_k, outs

    def stable_sort(self, keys, values=(), value_count=None, first_bit=0, end_bit=32):
        vc = len(values) if value_count is None else value_count
        return self._run(self.L.oracle_stable_sort, keys, values, vc, first_bit, end_bit)

    def sort1(self, keys, values=(), value_count=None, first_bit=0, end_bit=32, digit_bits=4):
        vc = len(values) if value_count is None else value_count
        return self._run(self.L.oracle_sort1, keys, values, vc, first_bit, end_bit, digit_bits)

    def sort2(self, keys, values=(), value_count=None, first_bit=0, end_bit=32, digit_bits=4,
              block_size=2048, threads=1):
        vc = len(values) if value_count is None else value_count
        return self._run(self.L.oracle_sort2, keys, values, vc, first_bit, end_bit, digit_bits,
                         block_size, threads)

    def key_transform(self, keys, key_type, to_radix):
        """keys: uint32 bit patterns; key_type 1 = int32, 2 = float32."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.empty_like(keys)
        assert self.L.oracle_key_transform(_u32p(keys), keys.size, key_type, int(to_radix),
                                           _u32p(out)) == 0
        return out

    def fill_keys(self, n, seed, kind=0, mask_bits=32):
        out = np.empty(n, dtype=np.uint32)
        self.L.oracle_fill_keys(_u32p(out), n, seed, kind, mask_bits)
        return out

    def block_counts(self, keys, bit, num_bits, block_size):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        nb = (keys.size + block_size - 1) // block_size
        out = np.zeros(nb * (1 << num_bits), dtype=np.uint32)
        assert self.L.oracle_block_counts(_u32p(keys), keys.size, bit, num_bits, block_size,
                                          _u32p(out)) == 0
        return out.reshape(nb, 1 << num_bits)

    def digit_offsets(self, counts):
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        nb, nd = counts.shape
        out = np.zeros(nb * nd, dtype=np.uint64)
        assert self.L.oracle_digit_offsets(_u32p(counts.ravel()), nb, int(nd).bit_length() - 1,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))) == 0
        return out.reshape(nb, nd)


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


def bitfield(keys, first_bit, end_bit):
    keys = np.asarray(keys, dtype=np.uint64)
    nb = end_bit - first_bit
    return (keys >> np.uint64(first_bit)) & np.uint64((1 << nb) - 1)


def check_stable_sort(in_keys, out_keys, perm, first_bit, end_bit):
    """Size-independent check that (out_keys, perm) is THE stable sort of
    in_keys by the bit field: perm is a permutation, out_keys gathers in_keys
    through it, the bit field is nondecreasing, and ties keep source order.
    These properties determine the stable sort uniquely."""
    n = in_keys.size
    perm = np.asarray(perm, dtype=np.int64)
    assert out_keys.size == n and perm.size == n
    seen = np.zeros(n, dtype=bool)
    assert perm.min(initial=0) >= 0 and perm.max(initial=-1) < n
    seen[perm] = True
    assert seen.all(), "values are not a permutation of the source indices"
    assert np.array_equal(out_keys, in_keys[perm]), "keys do not follow the permutation"
    bf = bitfield(out_keys, first_bit, end_bit)
    if n > 1:
        d = np.diff(bf.astype(np.int64))
        assert (d >= 0).all(), "bit field not sorted"
        ties = d == 0
        assert (np.diff(perm)[ties] > 0).all(), "equal keys out of source order (unstable)"
