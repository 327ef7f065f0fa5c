"""Device-side helpers for the GPU tests: run the sort through the C ABI on
torch-owned buffers and read the results back."""
import numpy as np

import mgpu_old_amd as mg


def torch_mod():
    import torch
    return torch


def to_dev(arr, capacity=None):
    torch = torch_mod()
    arr = np.ascontiguousarray(arr, dtype=np.uint32)
    cap = max(1, arr.size if capacity is None else capacity)
    t = torch.empty(cap, dtype=torch.int32, device="cuda")
    if arr.size:
        t[: arr.size].copy_(torch.from_numpy(arr.view(np.int32)))
    return t


def empty_dev(n):
    torch = torch_mod()
    return torch.empty(max(1, n), dtype=torch.int32, device="cuda")


def to_host(t, n):
    return t[:n].cpu().numpy().view(np.uint32).copy()


class GpuSort:
    """Owns the device buffers of one sortData_d filled from numpy inputs."""

    def __init__(self, keys, values=None, value_count=None, first_bit=0, end_bit=32,
                 capacity=None):
        values = list(values or [])
        self.n = keys.size
        self.value_count = len(values) if value_count is None else value_count
        cap = max(self.n, 1) if capacity is None else capacity
        self.tensors = []
        d = mg.SortData()
        d.maxElements = cap
        d.numElements = self.n
        d.valueCount = self.value_count
        d.firstBit = first_bit
        d.endBit = end_bit
        k0, k1 = to_dev(keys, cap), empty_dev(cap)
        d.keys[0], d.keys[1] = k0.data_ptr(), k1.data_ptr()
        self.tensors += [k0, k1]
        self.start_keys = k0.data_ptr()
        carried = 1 if self.value_count == -1 else self.value_count
        for j in range(carried):
            v0 = to_dev(values[j], cap) if self.value_count != -1 else empty_dev(cap)
            v1 = empty_dev(cap)
            d.values[j][0], d.values[j][1] = v0.data_ptr(), v1.data_ptr()
            self.tensors += [v0, v1]
        self.data = d
        self.by_ptr = {t.data_ptr(): t for t in self.tensors}

    def run(self, engine, ex=None, key_type=None):
        torch = torch_mod()
        torch.cuda.synchronize()
        if key_type is not None:
            engine.sort_array_typed(self.data, key_type)
        elif ex is None:
            engine.sort_array(self.data)
        else:
            engine.sort_array_ex(self.data, *ex)
        torch.cuda.synchronize()
        return self

    def keys(self):
        return to_host(self.by_ptr[self.data.keys[0]], self.n)

    def values(self):
        carried = 1 if self.value_count == -1 else self.value_count
        return [to_host(self.by_ptr[self.data.values[j][0]], self.n) for j in range(carried)]


_hip = None


def copy_d2d(dst, src, nbytes):
    """Synchronous device-to-device copy between raw device addresses."""
    import ctypes
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    assert _hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3) == 0


def make_engine(env=None):
    """An engine on torch's current stream; `env` holds MGPUSORT_* settings
    that the engine reads at creation (restored afterwards)."""
    import os
    torch = torch_mod()
    env = env or {}
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        eng = mg.Engine()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    return eng
