"""Benchmark of the gfx950 LSD radix sort through the mgpusort C ABI.

Headline (`value` of the JSON line): sortArray of 2^28 uint32 keys uniform
over 32 bits plus one uint32 value array (0..N-1), bits [0, 32), Gkeys/s on
one GPU.  Every timed step sorts its own pre-made copy of the input, already
resident in HBM, so the timed region holds exactly `steps` sortArray calls.

Also on the same line:
  roofline      the dominant kernel (rank_scatter) from per-launch HIP events:
                8*N*(1+v) algorithmic bytes per launch / average launch time
  sort_roofline the whole sort: sum over passes of 4N(3+2v) (16N for an
                index-generating pass) / sort time, against 8 TB/s
  cpu_baseline  the fastest of the oracle's CPU sorts (std::stable_sort on 1
                core; cpusort.cpp Sort2 on 1 core and on every host thread) on a
                bounded sample; all three in cpu_baselines
  extra         2^20 keys only, 2^25 key+value, 2^28 low-entropy keys with
                index generation at 32 and 24 bits
  segmented     4096 segments (sizes log-uniform in [1, 2^16], key + value)
                in one graph-replayed launch vs a loop of sortArray calls
  served        Poisson stream through the sort server: p50/p99 latency and
                Gkeys/s at offered loads of 50% and 80% of measured capacity
  early_exit    2^28 key + value with earlyExit on and off: random input (the
                cost of the check), sorted, nearly sorted (0.5% of the 1K-key
                blocks of sorted keys shuffled) and in top-byte order
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s
# HBM bytes per launch of the dominant kernel, measured by rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE passes over this script and scaled by the same
# counters' reading of a plain 1 GiB copy (the profile summary beside it
# gives the commands); used only when the kernel and size match.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rank_scatter_2p28_kv_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-segmented", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-early-exit", action="store_true")
    ap.add_argument("--no-typed", action="store_true")
    ap.add_argument("--segmented-direct", action="store_true",
                    help="segmented leg through sortSegmentedArray instead of plan replays")
    ap.add_argument("--no-headline", action="store_true",
                    help="skip the headline sort (profiling the other legs alone); value is null")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-log2n", type=int, default=25)
    ap.add_argument("--served-requests", type=int, default=600)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one plain device copy of this size (calibrates the HBM "
                         "byte counters of a profiler run)")
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
    """Device buffers for `copies` independent inputs of one sort workload
    (keys and values given as int32 device tensors)."""

    def __init__(self, torch, keys, values, value_count, first_bit, end_bit, copies):
        self.torch = torch
        self.n = keys.numel()
        self.value_count = value_count
        self.first_bit, self.end_bit = first_bit, end_bit
        self.src_k = keys
        self.src_v = values
        carried = 1 if value_count == -1 else value_count
        self.scratch_k = torch.empty_like(keys)
        self.scratch_v = [torch.empty_like(keys) for _ in range(carried)]
        self.copies = []
        for _ in range(copies):
            k = keys.clone()
            vs = [v.clone() for v in values] if value_count != -1 else [torch.empty_like(keys)]
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
    """The stable-sort properties on the GPU: perm is a permutation,
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


def time_workload(torch, eng, wl, steps, warmup, timing=False, early_exit=False):
    """Wall time per sort of `steps` sorts (synchronized on both sides), and
    optionally the per-kernel event timings of the same sorts."""
    for i in range(warmup):
        d = wl.data(i)
        d.earlyExit = early_exit
        eng.sort_array(d)
    torch.cuda.synchronize()
    eng.check_device()
    eng.enable_timing(timing)
    datas = [wl.data(warmup + i) for i in range(steps)]
    for d in datas:
        d.earlyExit = early_exit
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


def run_workload(torch, eng, name, keys, values, value_count, first_bit, end_bit, steps,
                 warmup, timing=False, verify=True):
    import mgpu_old_amd as mg
    n = keys.numel()
    log(f"[bench] {name}: n={n} vc={value_count} bits=[{first_bit},{end_bit}) "
        f"copies={steps + warmup}")
    wl = Workload(torch, keys, values, value_count, first_bit, end_bit, steps + warmup)
    sec, datas, kt = time_workload(torch, eng, wl, steps, warmup, timing)
    passes = len(mg.plan_passes(end_bit - first_bit))
    v = 1 if value_count == -1 else value_count
    out = {
        "workload": name, "n": n, "value_count": value_count, "bits": [first_bit, end_bit],
        "passes": passes, "ms_per_sort": sec * 1e3, "gkeys_per_s": n / sec / 1e9,
    }
    b = sort_bytes(n, value_count, passes)
    out["sort_roofline"] = {"bytes_alg": b, "achieved_GBps": b / sec / 1e9,
                            "frac": b / sec / HBM_PEAK, "floor_bytes": 8 * n * (1 + v),
                            "floor_frac": 8 * n * (1 + v) / sec / HBM_PEAK}
    if verify:
        k_out, v_out = wl.result(datas[0])
        if value_count in (1, -1):
            ok = verify_on_device(torch, wl.src_k, k_out, v_out, first_bit, end_bit)
        else:
            ok = bool((torch.sort(wl.src_k.long() & 0xFFFFFFFF).values ==
                       (k_out.long() & 0xFFFFFFFF)).all())
        out["bit_exact_properties"] = ok
        log(f"[bench] {name}: verify={'ok' if ok else 'FAILED'}")
    out["kernels"] = kt
    del wl
    torch.cuda.empty_cache()
    log(f"[bench] {name}: {out['gkeys_per_s']:.2f} Gkeys/s, {out['ms_per_sort']:.3f} ms, "
        f"sort roofline {out['sort_roofline']['frac']:.3f}")
    return out


def typed_bench(torch, eng, log2n, steps):
    """float32 keys (uniform 32-bit patterns: NaNs, infinities and denormals
    included) + one value at 2^log2n: sortArrayTyped, which maps the keys
    inside the count and the first and last passes, against
    sortTransformKeys + sortArray + sortTransformKeys; wall time per sort over
    `steps` fresh inputs each, results compared bit for bit."""
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads
    n = 1 << log2n
    keys = workloads.uniform_keys_torch(n, 2029, "cuda", torch=torch)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    f32 = mg.KEYS_FLOAT32

    def fused(d):
        eng.sort_array_typed(d, f32)

    def three_calls(d):
        eng.transform_keys(d.keys[0], n, f32, True)
        eng.sort_array(d)
        eng.transform_keys(d.keys[0], n, f32, False)

    out, results = {"workload": f"float32 2^{log2n} key + value, [0,32)", "n": n}, []
    for name, fn in (("fused", fused), ("three_calls", three_calls)):
        log(f"[bench] typed keys: {name}")
        wl = Workload(torch, keys, [vals], 1, 0, 32, steps + 1)
        fn(wl.data(0))
        datas = [wl.data(1 + i) for i in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for d in datas:
            fn(d)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        eng.check_device()
        k_out, v_out = wl.result(datas[0])
        results.append((k_out.clone(), v_out.clone()))
        out[f"{name}_ms"] = sec * 1e3
        out[f"{name}_gkeys_per_s"] = n / sec / 1e9
        del wl, datas
        torch.cuda.empty_cache()
    out["same_result"] = bool(torch.equal(results[0][0], results[1][0]) and
                              torch.equal(results[0][1], results[1][1]))
    log(f"[bench] typed keys: fused {out['fused_ms']:.3f} ms, three calls "
        f"{out['three_calls_ms']:.3f} ms, same result {out['same_result']}")
    del keys, vals, results
    torch.cuda.empty_cache()
    return out


def early_exit_bench(torch, eng, log2n, steps):
    """earlyExit on 2^log2n key + value sorts: each input timed with the
    flag off and on, alternating (same copies, every round verified)."""
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads
    n = 1 << log2n
    keys = workloads.uniform_keys_torch(n, 2029, "cuda", torch=torch)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")

    def sorted_by(k, first_bit, end_bit):
        # Input generation with the sort itself (checked by the runs below).
        wl = Workload(torch, k, [], 0, first_bit, end_bit, 1)
        d = wl.data(0)
        eng.sort_array(d)
        torch.cuda.synchronize()
        out = wl.result(d)[0].clone()
        del wl
        return out

    srt = sorted_by(keys, 0, 32)
    nearly = srt.clone()
    blocks = nearly[: n // 1024 * 1024].view(-1, 1024)
    u = workloads.words_torch(2030, 0, blocks.shape[0], "cuda", torch).long() & 0xFFFFFFFF
    pick = torch.nonzero(u < int(0.005 * 4294967296.0)).flatten()
    rnd = workloads.words_torch(2031, 0, pick.numel() * 1024, "cuda", torch).view(-1, 1024)
    order = torch.argsort(rnd, dim=1)
    blocks[pick] = torch.gather(blocks[pick], 1, order)
    top = sorted_by(keys, 24, 32)
    cases = [("random", keys, 0, 32), ("sorted", srt, 0, 32),
             ("nearly sorted (0.5% of 1K blocks shuffled)", nearly, 0, 32),
             ("nearly sorted, bits [18,26)", nearly, 18, 26),
             ("top-byte order", top, 0, 32)]
    out = {"n": n, "value_count": 1, "shuffled_blocks": int(pick.numel()), "cases": []}
    rounds = 3
    for name, k, fb, eb in cases:
        row = {"input": name, "bits": [fb, eb]}
        # Off and on alternate over `rounds` rounds of `steps` sorts each (the
        # same copies, restored from the input before every round), and each
        # takes its median round: the clock drifts between rounds by more
        # than the check costs.
        wl = Workload(torch, k, [vals], 1, fb, eb, steps)

        def round_of(flag):
            for kc, vcs in wl.copies:
                kc.copy_(k)
                vcs[0].copy_(vals)
            torch.cuda.synchronize()
            datas = [wl.data(i) for i in range(steps)]
            for d in datas:
                d.earlyExit = flag
            t0 = time.perf_counter()
            for d in datas:
                eng.sort_array(d)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / steps, datas

        round_of(False)
        round_of(True)
        times = {False: [], True: []}
        for _ in range(rounds):
            for flag in (False, True):
                sec, datas = round_of(flag)
                times[flag].append(sec)
                k_out, v_out = wl.result(datas[0])
                ok = verify_on_device(torch, wl.src_k, k_out, v_out, fb, eb)
                prev = row.get("on" if flag else "off", {}).get("bit_exact_properties", True)
                row["on" if flag else "off"] = {"parity": datas[0].parity,
                                                "bit_exact_properties": prev and ok}
        eng.check_device()
        for flag in (False, True):
            sec = sorted(times[flag])[rounds // 2]
            row["on" if flag else "off"].update({"ms_per_sort": sec * 1e3,
                                                 "gkeys_per_s": n / sec / 1e9,
                                                 "ms_rounds": [t * 1e3 for t in times[flag]]})
        del wl
        torch.cuda.empty_cache()
        row["speedup_on_vs_off"] = row["off"]["ms_per_sort"] / row["on"]["ms_per_sort"]
        out["cases"].append(row)
        log(f"[bench] earlyExit {name} [{fb},{eb}): off {row['off']['ms_per_sort']:.3f} ms, "
            f"on {row['on']['ms_per_sort']:.3f} ms (parity {row['on']['parity']}), "
            f"verify {row['off']['bit_exact_properties'] and row['on']['bit_exact_properties']}")
    rnd_row = out["cases"][0]
    out["cost_on_random"] = rnd_row["on"]["ms_per_sort"] / rnd_row["off"]["ms_per_sort"] - 1
    del keys, vals, srt, nearly, top
    torch.cuda.empty_cache()
    return out


def segmented_bench(torch, eng, steps, direct=False):
    """4096 segments, sizes log-uniform in [1, 2^16], key + one value array:
    graph-replayed segmented plan vs a loop of sortArray per segment.  With
    direct, sortSegmentedArray calls (each synchronous, no graph) take the
    plan's place and the loop is left out: the form for rocprofv3's counter
    mode, which follows no graph replays and crashed inside the loop's
    4096 sortArray launches."""
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads
    sizes = workloads.log_uniform_sizes(4096, 0, 16, seed=4)
    starts = np.zeros(len(sizes) + 1, dtype=np.int64)
    starts[1:] = np.cumsum(sizes)
    n = int(starts[-1])
    log(f"[bench] segmented: 4096 segments, {n} keys")
    keys = workloads.uniform_keys_torch(n, 44, "cuda", torch=torch)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    copies = []
    for _ in range(steps + 1):
        k, v = keys.clone(), vals.clone()
        k1, v1 = torch.empty_like(k), torch.empty_like(v)
        d = mg.SortData()
        d.maxElements = d.numElements = n
        d.valueCount, d.firstBit, d.endBit = 1, 0, 32
        d.keys[0], d.keys[1] = k.data_ptr(), k1.data_ptr()
        d.values[0][0], d.values[0][1] = v.data_ptr(), v1.data_ptr()
        copies.append((d, k, v, k1, v1, None if direct else eng.segmented_plan(d, starts)))
    torch.cuda.synchronize()
    if direct:
        eng.sort_segmented(copies[0][0], starts)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for d, *_rest in copies[1:]:
            eng.sort_segmented(d, starts)
        torch.cuda.synchronize()
    else:
        copies[0][5].execute(copies[0][0])  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for d, *_rest, plan in copies[1:]:
            plan.execute(d)
        torch.cuda.synchronize()
    seg_s = (time.perf_counter() - t0) / steps
    eng.check_device()
    seg_k, seg_v = copies[1][1], copies[1][2]  # 4 passes: result back in side 0

    if direct:
        out = {"segments": 4096, "n": n, "form": "direct", "ms_segmented": seg_s * 1e3,
               "gkeys_per_s_segmented": n / seg_s / 1e9}
        log(f"[bench] segmented (direct): {out['ms_segmented']:.3f} ms")
        del copies
        torch.cuda.empty_cache()
        return out

    # Baseline: one sortArray per segment on the same input.
    k, v = keys.clone(), vals.clone()
    k1, v1 = torch.empty_like(k), torch.empty_like(v)
    datas = []
    for i in range(len(sizes)):
        a, b = int(starts[i]), int(starts[i + 1])
        d = mg.SortData()
        d.maxElements = d.numElements = b - a
        d.valueCount, d.firstBit, d.endBit = 1, 0, 32
        d.keys[0], d.keys[1] = k[a:].data_ptr(), k1[a:].data_ptr()
        d.values[0][0], d.values[0][1] = v[a:].data_ptr(), v1[a:].data_ptr()
        datas.append(d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for d in datas:
        eng.sort_array(d)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    eng.check_device()
    same = bool(torch.equal(seg_k, k)) and bool(torch.equal(seg_v, v))
    for c in copies:
        if c[5] is not None:
            c[5].close()
    out = {"segments": 4096, "n": n, "form": "plan replay",
           "ms_segmented": seg_s * 1e3,
           "gkeys_per_s_segmented": n / seg_s / 1e9, "ms_sortarray_loop": loop_s * 1e3,
           "gkeys_per_s_sortarray_loop": n / loop_s / 1e9, "speedup_vs_loop": loop_s / seg_s,
           "bit_exact_vs_loop": same}
    log(f"[bench] segmented: {out['ms_segmented']:.3f} ms vs loop {out['ms_sortarray_loop']:.1f} ms "
        f"= {out['speedup_vs_loop']:.1f}x, bit-exact={same}")
    del copies
    torch.cuda.empty_cache()
    return out


def served_bench(torch, eng, count):
    """Poisson stream of sort requests through the sort server."""
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads
    sizes = workloads.log_uniform_sizes(count, 10, 24, seed=5)
    w = workloads.words(6, 0, 2 * count)
    widths = np.array([16, 24, 32])[w[:count] % 3]
    u = w[count:] % 100
    vcs = np.where(u < 50, 0, np.where(u < 85, 1, 2))
    total = int(sizes.sum())
    log(f"[bench] served: {count} requests, {total} keys")
    pristine, live, reqs = [], [], []
    for i in range(count):
        n, wd, vc = int(sizes[i]), int(widths[i]), int(vcs[i])
        k = workloads.uniform_keys_torch(n, 10_000 + i, "cuda", mask_bits=wd, torch=torch)
        vs = [workloads.uniform_keys_torch(n, 20_000 + 7 * i + j, "cuda", torch=torch)
              for j in range(vc)]
        pristine.append((k, vs))
        lk, lvs = k.clone(), [x.clone() for x in vs]
        live.append((lk, lvs))
        r = mg.SortRequest()
        r.keys = lk.data_ptr()
        for j, x in enumerate(lvs):
            r.values[j] = x.data_ptr()
        r.numElements, r.valueCount, r.firstBit, r.endBit = n, vc, 0, wd
        reqs.append(r)
    budget = 16 << 30
    # Six in-flight slots: the same capacity as three, lower latency under
    # load (p99 at 80% load 1.9 against 2.3 ms).
    slots = 6

    def reset():
        for (k, vs), (lk, lvs) in zip(pristine, live):
            lk.copy_(k)
            for a, b in zip(lvs, vs):
                a.copy_(b)
        torch.cuda.synchronize()

    def run(arrivals):
        reset()
        with mg.Server(eng, budget, max_in_flight=slots, max_elements=1 << 24) as srv:
            comps = srv.replay(reqs, arrivals)
        assert all(c.status == 0 for c in comps)
        lat = np.array([c.doneMs - c.submitMs for c in comps])
        span = max(c.doneMs for c in comps) - min(c.submitMs for c in comps)
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "gkeys_per_s": total / span / 1e6,
                "batches": len({c.batch for c in comps}), "makespan_ms": span}

    # Two warm-up runs: the first runs of a process, on a freshly allocated
    # slab, are 30-40% slower.
    run(np.zeros(count))
    run(np.zeros(count))
    # Capacity: every request at t=0; the median of three such runs (one run
    # spans ~20 ms, so one slow batch moves it by several Gkeys/s).
    bursts = sorted((run(np.zeros(count)) for _ in range(3)), key=lambda r: r["gkeys_per_s"])
    cap = bursts[1]
    cap["gkeys_per_s_runs"] = [b["gkeys_per_s"] for b in bursts]
    capacity = cap["gkeys_per_s"] * 1e9
    mean_keys = total / count
    out = {"requests": count, "keys": total, "hbm_budget_bytes": budget, "max_in_flight": slots,
           "capacity": cap, "loads": {}}
    for frac in (0.5, 0.8):
        rate = frac * capacity / mean_keys  # requests per second
        uu = (workloads.words(77, 0, count).astype(np.float64) + 0.5) / 4294967296.0
        arrivals = np.cumsum(-np.log(uu) / rate * 1e3)
        res = run(arrivals)
        res["offered_gkeys_per_s"] = frac * capacity / 1e9
        res["offered_requests_per_s"] = rate
        out["loads"][f"{int(frac * 100)}%"] = res
        log(f"[bench] served at {int(frac * 100)}% load: p50 {res['p50_ms']:.3f} ms, "
            f"p99 {res['p99_ms']:.3f} ms, {res['gkeys_per_s']:.2f} Gkeys/s")
    # Each served result equals the same request sorted alone by sortArray.
    same = True
    for i in range(0, count, max(1, count // 12)):
        (k, vs), (lk, lvs) = pristine[i], live[i]
        n, vc, wd = reqs[i].numElements, reqs[i].valueCount, reqs[i].endBit
        wl = Workload(torch, k, vs, vc, 0, wd, 1)
        d = wl.data(0)
        eng.sort_array(d)
        torch.cuda.synchronize()
        ak, av = wl.result(d)
        same = same and bool(torch.equal(ak, lk))
        if vc:
            same = same and bool(torch.equal(av, lvs[0]))
    out["alone_equals_served"] = same
    del pristine, live
    torch.cuda.empty_cache()
    return out


def cpu_baseline(log2n):
    """The oracle's CPU sorts of (key, value) on 2^log2n keys: std::stable_sort
    on one core, the restated cpusort.cpp Sort2 (4-bit digits, blocks of 2048
    keys) on one core and with its parallel parts on every host thread.  The
    fastest is the reported baseline; all three are listed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    from mgpu_old_amd import workloads
    n = 1 << log2n
    keys = workloads.uniform_keys(n, seed=28)
    vals = np.arange(n, dtype=np.uint32)
    o = Oracle()
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    size = f"2^{log2n} uniform u32 keys + 1 u32 value (the headline workload at 1/{1 << (28 - log2n)} size"
    runs = []
    t0 = time.perf_counter()
    rk, rv = o.stable_sort(keys, [vals])
    dt = time.perf_counter() - t0
    runs.append({"value": n / dt / 1e9, "unit": "Gkeys/s", "cores": 1, "kind": "port",
                 "sample": f"oracle std::stable_sort of {size}, {dt:.1f} s)"})
    for th in (1, threads):
        t0 = time.perf_counter()
        k2, v2 = o.sort2(keys, [vals], digit_bits=4, block_size=2048, threads=th)
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(k2, rk) and np.array_equal(v2[0], rv[0]))
        runs.append({"value": n / dt / 1e9, "unit": "Gkeys/s", "cores": th, "kind": "port",
                     "sample": f"oracle cpusort.cpp Sort2 (4-bit digits, 2048-key blocks, {th} "
                               f"thread{'s' if th > 1 else ''}) of {size}, {dt:.1f} s; equals "
                               f"std::stable_sort: {same})"})
    best = dict(max(runs, key=lambda r: r["value"]))
    return best, runs


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; how the sort would shard over "
                          "several is a design note only", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads

    torch.cuda.set_device(0)
    eng = mg.Engine()
    eng.set_stream(torch.cuda.current_stream().cuda_stream)

    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib << 18, dtype=torch.int32, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"[bench] copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")

    n = 1 << args.log2n
    head = {"gkeys_per_s": None, "ms_per_sort": None, "workload": None, "passes": None,
            "bit_exact_properties": None, "sort_roofline": None}
    roof, kernel_ms = None, {}
    if not args.no_headline:
        keys = workloads.uniform_keys_torch(n, 2028, "cuda", torch=torch)
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        head = run_workload(torch, eng, f"sortArray 2^{args.log2n} u32 key + 1 u32 value (0..N-1), "
                            "uniform keys, bits [0,32)", keys, [vals], 1, 0, 32, args.steps,
                            args.warmup)
        timed = run_workload(torch, eng, "per-kernel timing", keys, [vals], 1, 0, 32,
                             min(args.steps, 5), 1, timing=True, verify=False)
        del keys, vals
        kt = timed["kernels"]
        scatter = {k: v for k, v in kt.items() if k.startswith("mgpu_rank_scatter")}
        name, (launches, total_ms, mn, mx) = max(scatter.items(), key=lambda kv: kv[1][1])
        avg_s = total_ms / launches / 1e3
        per_launch = 8 * n * 2  # read + write of the keys and one value array
        roof = {"bound": "hbm", "kernel": name, "achieved": per_launch / avg_s / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": per_launch / avg_s / HBM_PEAK,
                "traffic": None, "avg_launch_ms": avg_s * 1e3, "launches": launches,
                "bytes_per_launch": per_launch}
        try:
            with open(TRAFFIC_FILE) as f:
                measured = json.load(f)
            if measured.get("kernel") == name and measured.get("n") == n:
                roof["traffic"] = measured["traffic_bytes"]
                roof["traffic_source"] = measured["source"]
        except (OSError, ValueError, KeyError):
            pass
        kernel_ms = {k: {"avg_ms": v[1] / v[0], "launches": v[0]} for k, v in kt.items()}

    extra = []
    if not args.no_extra:
        k1 = workloads.uniform_keys_torch(1 << 20, 1, "cuda", torch=torch)
        extra.append(run_workload(torch, eng, "2^20 u32 keys only, [0,32)", k1, [], 0, 0, 32,
                                  20, 3))
        k2 = workloads.uniform_keys_torch(1 << 25, 2, "cuda", torch=torch)
        v2 = torch.arange(1 << 25, dtype=torch.int32, device="cuda")
        extra.append(run_workload(torch, eng, "2^25 u32 key + value (0..N-1), [0,32)", k2, [v2],
                                  1, 0, 32, 10, 3))
        del k1, k2, v2
        k3 = workloads.and3_keys_torch(1 << 28, 3, "cuda", torch=torch)
        extra.append(run_workload(torch, eng, "2^28 AND3 keys, index generation, [0,32)", k3,
                                  [], -1, 0, 32, 5, 2))
        k3 &= 0xFFFFFF
        extra.append(run_workload(torch, eng, "2^28 AND3 keys & 0xFFFFFF, index generation, "
                                  "[0,24)", k3, [], -1, 0, 24, 5, 2))
        del k3
        torch.cuda.empty_cache()
    for e in extra:
        e.pop("kernels", None)

    ee = None if args.no_early_exit else early_exit_bench(torch, eng, args.log2n, 5)
    typed = None if args.no_typed else typed_bench(torch, eng, args.log2n, 5)
    seg = None if args.no_segmented else segmented_bench(torch, eng, 10, args.segmented_direct)
    served = None if args.no_served else served_bench(torch, eng, args.served_requests)

    cpu, cpu_runs = None, None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu, cpu_runs = cpu_baseline(args.cpu_log2n)
        cpu["cores_available"] = os.cpu_count()

    line = {
        "metric": "Gkeys/s, 2^28 u32 key+value sort; %HBM roofline; batched/served p99 latency",
        "value": head["gkeys_per_s"],
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_sort"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64-seeded uniform keys, values 0..N-1)",
        "config": {"workload": head["workload"], "n": n, "value_count": 1, "bits": [0, 32],
                   "passes": head["passes"],
                   "bit_exact_properties": head["bit_exact_properties"]},
        "roofline": roof,
        "sort_roofline": head["sort_roofline"],
        "kernel_avg_ms": kernel_ms,
        "cpu_baseline": cpu,
        "cpu_baselines": cpu_runs,
        "extra": extra,
        "early_exit": ee,
        "typed_keys": typed,
        "segmented": seg,
        "served": served,
    }
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
