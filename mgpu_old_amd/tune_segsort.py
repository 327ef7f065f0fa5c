"""Autotuner of the segmented sort: picks, per segment-size bucket, the kernel
variant (digit width 4..8 bits x 8 or 16 keys per thread; block size 256) with
the shortest replay time, and writes the table the engine loads at creation
(kernelPath + "mgpusort_seg_tuning.txt").

It plays the role of the reference's offline pass-table generator
(sort/src/tablegen/tablegen.cpp:9-64 over measured rates in throughputs.txt),
but measures every candidate directly on the GPU instead of modelling them.

Usage (on a gfx950 box):  python -m mgpu_old_amd.tune_segsort [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# Buckets by ceil(log2 n): (low, high] in log2, inclusive of high.
BUCKETS = [(0, 6), (6, 8), (8, 10), (10, 12), (12, 14), (14, 16), (16, 18)]
VARIANTS = [f"r{rb}i{ipt}" for rb in range(4, 9) for ipt in (8, 16)]


def bucket_sizes(lo, hi, total, seed):
    from mgpu_old_amd import workloads
    sizes = []
    s = 0
    i = 0
    while s < total:
        u = workloads.words(seed, i, 1)[0] / 4294967296.0
        n = int(2 ** (lo + (hi - lo) * u))
        n = max(n, (1 << lo) + 1 if lo else 1)
        n = min(n, 1 << hi)
        sizes.append(n)
        s += n
        i += 1
    return np.array(sizes, dtype=np.int64)


def time_variant(torch, engine, variant, keys, vals, starts, reps):
    import mgpu_old_amd as mg
    os.environ["MGPUSORT_SEG_VARIANT"] = variant
    n = keys.numel()
    k1, v1 = torch.empty_like(keys), torch.empty_like(vals)
    d = mg.SortData()
    d.maxElements = d.numElements = n
    d.valueCount = 1
    d.firstBit, d.endBit = 0, 32
    d.keys[0], d.keys[1] = keys.data_ptr(), k1.data_ptr()
    d.values[0][0], d.values[0][1] = vals.data_ptr(), v1.data_ptr()
    plan = engine.segmented_plan(d, starts)
    plan.execute(d)  # warm-up (also builds nothing new: 32-bit keys keep orientation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.execute(d)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    plan.close()
    engine.check_device()
    return dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "lib", "mgpusort_seg_tuning.txt"))
    ap.add_argument("--elements", type=int, default=1 << 22, help="keys per bucket")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()

    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads

    # The table being written must not steer the measurement.
    if os.path.exists(args.out):
        os.remove(args.out)
    engine = mg.Engine()
    engine.set_stream(torch.cuda.current_stream().cuda_stream)
    rows, report = [], {}
    for bi, (lo, hi) in enumerate(BUCKETS):
        sizes = bucket_sizes(lo, hi, args.elements, seed=500 + bi)
        starts = np.zeros(len(sizes) + 1, dtype=np.int64)
        starts[1:] = np.cumsum(sizes)
        n = int(starts[-1])
        keys = torch.from_numpy(workloads.uniform_keys(n, 7).view(np.int32)).cuda()
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        times = {}
        for v in VARIANTS:
            times[v] = time_variant(torch, engine, v, keys, vals, starts, args.reps)
            print(f"[tune] bucket (2^{lo}, 2^{hi}] {len(sizes)} segs {n} keys {v}: "
                  f"{times[v] * 1e3:.3f} ms ({n / times[v] / 1e9:.2f} Gkeys/s)", flush=True)
        best = min(times, key=times.get)
        rb, ipt = int(best[1]), int(best[3:])
        rows.append((hi, rb, ipt))
        report[f"{lo}-{hi}"] = {"best": best, "ms": {k: t * 1e3 for k, t in times.items()},
                                "segments": len(sizes), "keys": n}
    os.environ.pop("MGPUSORT_SEG_VARIANT", None)
    with open(args.out, "w") as f:
        f.write("# segmented-sort tuning table: max ceil(log2 n), radix bits, keys per thread\n")
        f.write("# written by mgpu_old_amd/tune_segsort.py (key + 1 value, bits [0,32))\n")
        for hi, rb, ipt in rows:
            f.write(f"{hi} {rb} {ipt}\n")
    print(json.dumps({"table": rows, "report": report}), flush=True)
    engine.close()


if __name__ == "__main__":
    main()
