"""Benchmark of the gfx950 LSD radix sort through the mgpusort C ABI.

Headline (the `value` of the JSON line): sortArray of 2^28 uint32 keys uniform
over 32 bits plus one uint32 value array (0..N-1), bits [0, 32), Gkeys/s on
one GPU.  Every timed step sorts its own pre-made copy of the input that is
already resident in HBM, so the timed region holds exactly `steps` sortArray
calls and nothing else.

Also reported on the same line:
  roofline      the dominant kernel (rank_scatter) from per-launch HIP events:
                algorithmic bytes 8*N*(1+v) per launch / average launch time
  sort_roofline whole-sort algorithmic bytes sum_p 4N(3+2v) (index pass 16N)
                / sort time, against 8 TB/s
  cpu_baseline  the oracle's std::stable_sort (1 core) on a bounded sample
  extra         the other sort workloads (2^20 keys, 2^25 key+value, 2^28
                low-entropy with index generation at 32 and 24 bits)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-log2n", type=int, default=25)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sort_bytes(n, value_count, passes):
    """Algorithmic bytes of one sort: 4N(3+2v) per pass, 16N for an
    index-generating first pass (v counts the index as one array)."""
    v = 1 if value_count == -1 else value_count
    total = 0
    for p in range(passes):
        total += 16 * n if (value_count == -1 and p == 0) else 4 * n * (3 + 2 * v)
    return total


class Workload:
    """Device buffers for `copies` independent inputs of one sort workload."""

    def __init__(self, torch, keys_np, values_np, value_count, first_bit, end_bit, copies):
        self.torch = torch
        self.n = keys_np.size
        self.value_count = value_count
        self.first_bit, self.end_bit = first_bit, end_bit
        dev = torch.device("cuda")
        self.src_k = torch.from_numpy(keys_np.view(np.int32)).to(dev)
        self.src_v = [torch.from_numpy(v.view(np.int32)).to(dev) for v in values_np]
        carried = 1 if value_count == -1 else value_count
        self.scratch_k = torch.empty_like(self.src_k)
        self.scratch_v = [torch.empty_like(self.src_k) for _ in range(carried)]
        self.copies = []
        for _ in range(copies):
            k = self.src_k.clone()
            vs = [v.clone() for v in self.src_v] if value_count != -1 else \
                [torch.empty_like(self.src_k)]
            self.copies.append((k, vs))
        torch.cuda.synchronize()

    def data(self, i):
        import mgpu_old_amd as mg
        k, vs = self.copies[i]
        d = mg.SortData()
        d.maxElements = self.n
        d.numElements = self.n
        d.valueCount = self.value_count
        d.firstBit, d.endBit = self.first_bit, self.end_bit
        d.keys[0], d.keys[1] = k.data_ptr(), self.scratch_k.data_ptr()
        for j, v in enumerate(vs):
            d.values[j][0], d.values[j][1] = v.data_ptr(), self.scratch_v[j].data_ptr()
        return d

    def result(self, d):
        by = {t.data_ptr(): t for t in [self.scratch_k, *self.scratch_v]}
        for k, vs in self.copies:
            by[k.data_ptr()] = k
            for v in vs:
                by[v.data_ptr()] = v
        return by[d.keys[0]], (by[d.values[0][0]] if self.value_count else None)


def verify_on_device(torch, keys_in, keys_out, perm, first_bit, end_bit):
    """The stable-sort properties on the GPU (torch): perm is a permutation,
    keys_out = keys_in[perm], bit field nondecreasing, ties in source order."""
    mask = (1 << (end_bit - first_bit)) - 1
    kin = keys_in.long() & 0xFFFFFFFF
    kout = keys_out.long() & 0xFFFFFFFF
    p = perm.long() & 0xFFFFFFFF
    n = kin.numel()
    ok = bool((p < n).all())
    ok = ok and bool((torch.bincount(p, minlength=n) == 1).all())
    ok = ok and bool((kin[p] == kout).all())
    bf = (kout >> first_bit) & mask
    if n > 1:
        d = bf[1:] - bf[:-1]
        ok = ok and bool((d >= 0).all())
        ties = d == 0
        ok = ok and bool(((p[1:] - p[:-1]) > 0)[ties].all())
    return ok


def time_workload(torch, eng, wl, steps, warmup, timing=False):
    """Wall time of `steps` sorts (barrier + sync on both sides); optional
    per-kernel event timing of the same steps."""
    for i in range(warmup):
        eng.sort_array(wl.data(i))
    torch.cuda.synchronize()
    eng.check_device()
    eng.enable_timing(timing)
    datas = [wl.data(warmup + i) for i in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for d in datas:
        eng.sort_array(d)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = eng.read_timing() if timing else {}
    eng.enable_timing(False)
    eng.check_device()
    return (t1 - t0) / steps, datas, kt


def run_workload(torch, eng, name, keys_np, values_np, value_count, first_bit, end_bit, steps,
                 warmup, timing=False, verify=True):
    import mgpu_old_amd as mg
    n = keys_np.size
    log(f"[bench] {name}: n={n} vc={value_count} bits=[{first_bit},{end_bit}) "
        f"copies={steps + warmup}")
    wl = Workload(torch, keys_np, values_np, value_count, first_bit, end_bit, steps + warmup)
    sec, datas, kt = time_workload(torch, eng, wl, steps, warmup, timing)
    passes = len(mg.plan_passes(end_bit - first_bit))
    out = {
        "workload": name, "n": n, "value_count": value_count, "bits": [first_bit, end_bit],
        "passes": passes, "ms_per_sort": sec * 1e3, "gkeys_per_s": n / sec / 1e9,
    }
    b = sort_bytes(n, value_count, passes)
    out["sort_roofline"] = {"bytes_alg": b, "achieved_GBps": b / sec / 1e9,
                            "frac": b / sec / HBM_PEAK,
                            "floor_bytes": 8 * n * (1 + (1 if value_count == -1 else value_count)),
                            }
    if verify:
        k_out, v_out = wl.result(datas[0])
        if value_count in (1, -1):
            perm = v_out if value_count == -1 else v_out  # values hold 0..N-1 / indices
            ok = verify_on_device(torch, wl.src_k, k_out, perm, first_bit, end_bit)
        else:
            kin = wl.src_k.long() & 0xFFFFFFFF
            kout = k_out.long() & 0xFFFFFFFF
            ok = bool((torch.sort(kin).values == kout).all())
        out["bit_exact_properties"] = ok
        log(f"[bench] {name}: verify={'ok' if ok else 'FAILED'}")
    out["kernels"] = kt
    del wl
    torch.cuda.empty_cache()
    log(f"[bench] {name}: {out['gkeys_per_s']:.2f} Gkeys/s, {out['ms_per_sort']:.3f} ms, "
        f"sort roofline {out['sort_roofline']['frac']:.3f}")
    return out


def cpu_baseline(log2n):
    """Oracle std::stable_sort of (key, value) on 2^log2n keys, one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    from mgpu_old_amd import workloads
    n = 1 << log2n
    keys = workloads.uniform_keys(n, seed=28)
    vals = np.arange(n, dtype=np.uint32)
    o = Oracle()
    t0 = time.perf_counter()
    o.stable_sort(keys, [vals])
    dt = time.perf_counter() - t0
    return {"value": n / dt / 1e9, "unit": "Gkeys/s", "cores": 1, "kind": "port",
            "sample": f"oracle std::stable_sort of 2^{log2n} uniform u32 keys + 1 u32 value "
                      f"(same workload family as the headline, {dt:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                          "note only", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads

    torch.cuda.set_device(0)
    eng = mg.Engine()
    eng.set_stream(torch.cuda.current_stream().cuda_stream)

    n = 1 << args.log2n
    log(f"[bench] generating 2^{args.log2n} keys")
    keys = workloads.uniform_keys(n, seed=2028)
    vals = np.arange(n, dtype=np.uint32)
    head = run_workload(torch, eng, f"sortArray 2^{args.log2n} u32 key + 1 u32 value (0..N-1), "
                        "uniform keys, bits [0,32)", keys, [vals], 1, 0, 32, args.steps,
                        args.warmup, timing=False)
    # Same workload again with per-kernel events for the dominant kernel.
    timed = run_workload(torch, eng, "timed", keys, [vals], 1, 0, 32, min(args.steps, 5), 1,
                         timing=True, verify=False)
    del keys, vals
    kt = timed["kernels"]
    scatter = {k: v for k, v in kt.items() if k.startswith("mgpu_rank_scatter")}
    name, (launches, total_ms, mn, mx) = max(scatter.items(), key=lambda kv: kv[1][1])
    avg_s = total_ms / launches / 1e3
    per_launch = 8 * n * 2  # read + write of keys and one value array
    roof = {"bound": "hbm", "kernel": name, "achieved": per_launch / avg_s / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": per_launch / avg_s / HBM_PEAK,
            "traffic": None, "avg_launch_ms": avg_s * 1e3, "launches": launches,
            "bytes_per_launch": per_launch}
    kernel_ms = {k: v[1] / v[0] for k, v in kt.items()}

    extra = []
    if not args.no_extra:
        k1 = workloads.uniform_keys(1 << 20, seed=1)
        extra.append(run_workload(torch, eng, "2^20 u32 keys only, [0,32)", k1, [], 0, 0, 32,
                                  20, 3))
        k2 = workloads.uniform_keys(1 << 25, seed=2)
        extra.append(run_workload(torch, eng, "2^25 u32 key + value, [0,32)", k2,
                                  [np.arange(1 << 25, dtype=np.uint32)], 1, 0, 32, 10, 3))
        del k1, k2
        log("[bench] generating 2^28 low-entropy keys")
        k3 = workloads.and3_keys(1 << 28, seed=3)
        extra.append(run_workload(torch, eng, "2^28 AND3 keys, index generation, [0,32)", k3,
                                  [], -1, 0, 32, 5, 2))
        k3 &= np.uint32(0xFFFFFF)
        extra.append(run_workload(torch, eng, "2^28 AND3 keys & 0xFFFFFF, index generation, "
                                  "[0,24)", k3, [], -1, 0, 24, 5, 2))
        del k3
    for e in extra:
        e.pop("kernels", None)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(args.cpu_log2n)
        cpu["cores_available"] = os.cpu_count()

    ms = head["ms_per_sort"]
    line = {
        "metric": "Gkeys/s, 2^28 u32 key+value sort; %HBM roofline; batched/served p99 latency",
        "value": head["gkeys_per_s"],
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64-seeded uniform keys, values 0..N-1)",
        "config": {"workload": head["workload"], "n": n, "value_count": 1, "bits": [0, 32],
                   "passes": head["passes"], "bit_exact_properties": head["bit_exact_properties"]},
        "roofline": roof,
        "sort_roofline": head["sort_roofline"],
        "kernel_avg_ms": kernel_ms,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
