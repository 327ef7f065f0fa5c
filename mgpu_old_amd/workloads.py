"""Seeded synthetic inputs for the sort workloads (numpy, host side).

word(seed, i) = high 32 bits of the i-th SplitMix64 output for `seed`
(counter form: mix(seed + (i + 1) * 0x9e3779b97f4a7c15)), so any slice of a
stream can be generated independently and in chunks.

Workloads (sizes as benchmarked):
  uniform keys        key[i] = word(seed, i)
  low-entropy keys    key[i] = word(3i) & word(3i+1) & word(3i+2)  (P(bit)=1/8)
  log-uniform sizes   floor(2 ** (lo + (hi - lo) * u)), u from word / 2**32
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_CHUNK = 1 << 22


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed, start, count):
    """Outputs start .. start+count-1 of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + 1 + count, dtype=np.uint64)
        return _mix(np.uint64(seed) + i * GAMMA)


def words(seed, start, count):
    return (splitmix64(seed, start, count) >> np.uint64(32)).astype(np.uint32)


def uniform_keys(n, seed, mask_bits=32, out=None):
    out = np.empty(n, dtype=np.uint32) if out is None else out
    mask = np.uint32(0xFFFFFFFF if mask_bits >= 32 else (1 << mask_bits) - 1)
    for s in range(0, n, _CHUNK):
        c = min(_CHUNK, n - s)
        out[s:s + c] = words(seed, s, c) & mask
    return out


def and3_keys(n, seed, mask_bits=32, out=None):
    """Low-entropy keys: AND of three independent uniform words."""
    out = np.empty(n, dtype=np.uint32) if out is None else out
    mask = np.uint32(0xFFFFFFFF if mask_bits >= 32 else (1 << mask_bits) - 1)
    for s in range(0, n, _CHUNK):
        c = min(_CHUNK, n - s)
        w = words(seed, 3 * s, 3 * c).reshape(c, 3)
        out[s:s + c] = w[:, 0] & w[:, 1] & w[:, 2] & mask
    return out


def _i64(x):
    """uint64 constant as the int64 with the same bits."""
    return x - (1 << 64) if x >= 1 << 63 else x


def _torch_mix(z, torch):
    """SplitMix64's mix on int64 tensors (two's-complement wraparound; logical
    shifts done as arithmetic shift + mask)."""
    def srl(x, k):
        return (x >> k) & ((1 << (64 - k)) - 1)
    z = (z ^ srl(z, 30)) * _i64(0xBF58476D1CE4E5B9)
    z = (z ^ srl(z, 27)) * _i64(0x94D049BB133111EB)
    return z ^ srl(z, 31)


def words_torch(seed, start, count, device, torch=None):
    """words(seed, start, count) computed on a torch device (int32 tensor with
    the same bits as the uint32 words)."""
    if torch is None:
        import torch
    i = torch.arange(start + 1, start + 1 + count, dtype=torch.int64, device=device)
    z = _torch_mix(_i64(seed % (1 << 64)) + i * _i64(0x9E3779B97F4A7C15), torch)
    # High 32 bits; the int32 cast keeps the low 32 bits of the shifted value.
    return (z >> 32).to(torch.int32)


def uniform_keys_torch(n, seed, device, mask_bits=32, torch=None):
    """uniform_keys(n, seed, mask_bits) generated on the device (int32 bits)."""
    if torch is None:
        import torch
    out = torch.empty(n, dtype=torch.int32, device=device)
    mask = -1 if mask_bits >= 32 else (1 << mask_bits) - 1
    for s in range(0, n, 1 << 26):
        c = min(1 << 26, n - s)
        out[s:s + c] = words_torch(seed, s, c, device, torch) & mask
    return out


def and3_keys_torch(n, seed, device, mask_bits=32, torch=None):
    """and3_keys(n, seed, mask_bits) generated on the device (int32 bits)."""
    if torch is None:
        import torch
    out = torch.empty(n, dtype=torch.int32, device=device)
    mask = -1 if mask_bits >= 32 else (1 << mask_bits) - 1
    for s in range(0, n, 1 << 24):
        c = min(1 << 24, n - s)
        w = words_torch(seed, 3 * s, 3 * c, device, torch).view(c, 3)
        out[s:s + c] = w[:, 0] & w[:, 1] & w[:, 2] & mask
    return out


def nearly_sorted_keys(n, seed, shuffled_fraction=0.01, block=16384):
    """Sorted uniform keys in which about `shuffled_fraction` of the blocks
    of `block` keys (chosen by the seed) hold their keys in a seeded random
    order."""
    keys = np.sort(uniform_keys(n, seed))
    nb = (n + block - 1) // block
    pick = words(seed + 1, 0, nb).astype(np.float64) / 4294967296.0 < shuffled_fraction
    for b in np.flatnonzero(pick):
        s, e = int(b) * block, min(n, (int(b) + 1) * block)
        keys[s:e] = keys[s:e][np.argsort(words(seed + 2, s, e - s), kind="stable")]
    return keys


def top_sorted_keys(n, seed, top_bits=8):
    """Uniform keys in stable order of their top `top_bits` bits only (the
    bits below stay in random order)."""
    keys = uniform_keys(n, seed)
    return keys[np.argsort(keys >> np.uint32(32 - top_bits), kind="stable")]


def log_uniform_sizes(count, lo_log2, hi_log2, seed, lo=None, hi=None):
    """count sizes floor(2 ** (lo_log2 + (hi_log2 - lo_log2) * u)), clamped to
    [lo, hi] (default [2**lo_log2, 2**hi_log2])."""
    u = words(seed, 0, count).astype(np.float64) / 4294967296.0
    s = np.floor(np.exp2(lo_log2 + (hi_log2 - lo_log2) * u)).astype(np.int64)
    lo = 1 << lo_log2 if lo is None else lo
    hi = 1 << hi_log2 if hi is None else hi
    return np.clip(s, lo, hi)
