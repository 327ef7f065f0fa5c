"""Autotuner of the segmented sort.  It writes the table the engine loads at
creation (kernelPath + "mgpusort_seg_tuning.txt"):

  1. per segment-size bucket, the persistent kernel variant with the
     shortest replay: digit width 4..8 bits x keys per thread (8, 16) x
     workgroup size (128 threads for digits up to 7 bits, 256, 512), timed
     with every segment kept in the persistent kernel;
  2. the device-wide passes, on the batch they serve (4096 segments with
     sizes log-uniform in [1, 2^16], key + one value): the size from which a
     segment takes them, the tile shape and whether tiles are packed
     (segments back to back) or one segment's own, timed with the bucket
     rows of step 1 in force (the count piece size only matters for segments
     above 2^18 keys, which this batch does not have: it stays at 2^16, the
     better of 2^14 and 2^16 on the sort server's batches);
  3. on the same batch, for each threshold of step 2, one variant for every
     segment below it in place of the bucket rows of step 1, when faster: each bucket row
     is a persistent launch of its own, and in a batch those launches run one
     after another (each grid fills the device), so one launch whose tile
     holds the largest such segment can beat rows that win bucket by bucket
     (measured: the 3,089 segments of up to 4096 keys of that batch 0.116 ms
     as three launches, 0.070 ms as one of r8i8b512).

It plays the role of the reference's offline pass-table generator
(sort/src/tablegen/tablegen.cpp:9-64 over measured rates in throughputs.txt),
but measures every candidate directly on the GPU instead of modelling them.

Usage (on a gfx950 box):  python -m mgpu_old_amd.tune_segsort [--out PATH]
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# Buckets by ceil(log2 n): (low, high] in log2, inclusive of high.
BUCKETS = [(0, 6), (6, 8), (8, 10), (10, 12), (12, 13)]
VARIANTS = [f"r{rb}i{ipt}b{b}" for rb in range(4, 9) for ipt in (8, 16) for b in (128, 256, 512)
            if b >= (1 << rb)]
WIDE_MIN = [1025, 2049, 4097]
WIDE_CHUNK = [1 << 16]
# "auto" picks the small tile for this batch's segments (under four large
# tiles on average) and the large one for batches of larger segments, which
# a fixed "small" would slow down (the sort server's batches: 47 against 61
# Gkeys/s); "large" is the alternative measured here.
WIDE_TILE = ["auto", "large"]
WIDE_PACK = ["packed", "unpacked"]


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


def time_plan(torch, engine, keys, vals, starts, reps):
    """Mean time of `reps` replays of a captured segmented plan (key + one
    value, bits [0, 32)), after one warm-up replay: the median of three
    such runs."""
    import mgpu_old_amd as mg
    n = keys.numel()
    k1, v1 = torch.empty_like(keys), torch.empty_like(vals)
    d = mg.SortData()
    d.maxElements = d.numElements = n
    d.valueCount = 1
    d.firstBit, d.endBit = 0, 32
    d.keys[0], d.keys[1] = keys.data_ptr(), k1.data_ptr()
    d.values[0][0], d.values[0][1] = vals.data_ptr(), v1.data_ptr()
    plan = engine.segmented_plan(d, starts)
    plan.execute(d)  # warm-up (32-bit keys: 4 passes, the buffers keep their orientation)
    torch.cuda.synchronize()
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.execute(d)
        torch.cuda.synchronize()
        runs.append((time.perf_counter() - t0) / reps)
    plan.close()
    engine.check_device()
    return sorted(runs)[1]


def engine_on(torch, dirpath, table_text):
    """An engine over a copy of the code objects with the given table."""
    import mgpu_old_amd as mg
    for name in ("mgpusort_radix.hsaco", "mgpusort_seg.hsaco"):
        shutil.copy(os.path.join(mg.LIB_DIR, name), os.path.join(dirpath, name))
    with open(os.path.join(dirpath, "mgpusort_seg_tuning.txt"), "w") as f:
        f.write(table_text)
    eng = mg.Engine(dirpath + os.sep)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    return eng


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "lib", "mgpusort_seg_tuning.txt"))
    ap.add_argument("--elements", type=int, default=1 << 22, help="keys per bucket")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()

    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    from mgpu_old_amd import workloads

    tmp = tempfile.mkdtemp(prefix="segtune")
    report = {"buckets": {}, "wide": {}}
    try:
        # 1. Persistent kernel per bucket: every segment stays in it.
        os.environ["MGPUSORT_SEG_WIDE_MIN"] = str(1 << 30)
        engine = engine_on(torch, tmp, "# no rows: built-in defaults\n")
        rows = []
        for bi, (lo, hi) in enumerate(BUCKETS):
            sizes = bucket_sizes(lo, hi, args.elements, seed=500 + bi)
            starts = np.zeros(len(sizes) + 1, dtype=np.int64)
            starts[1:] = np.cumsum(sizes)
            n = int(starts[-1])
            keys = torch.from_numpy(workloads.uniform_keys(n, 7).view(np.int32)).cuda()
            vals = torch.arange(n, dtype=torch.int32, device="cuda")
            times = {}
            for v in VARIANTS:
                os.environ["MGPUSORT_SEG_VARIANT"] = v
                times[v] = time_plan(torch, engine, keys, vals, starts, args.reps)
                print(f"[tune] bucket (2^{lo}, 2^{hi}] {len(sizes)} segs {n} keys {v}: "
                      f"{times[v] * 1e3:.3f} ms ({n / times[v] / 1e9:.2f} Gkeys/s)", flush=True)
            os.environ.pop("MGPUSORT_SEG_VARIANT", None)
            best = min(times, key=times.get)
            rb, rest = best[1:].split("i")
            ipt, block = rest.split("b")
            rows.append((hi, int(rb), int(ipt), int(block)))
            report["buckets"][f"{lo}-{hi}"] = {"best": best, "segments": len(sizes), "keys": n,
                                               "ms": {k: t * 1e3 for k, t in times.items()}}
        engine.close()
        os.environ.pop("MGPUSORT_SEG_WIDE_MIN", None)
        bucket_text = "".join(f"{hi} {rb} {ipt} {b}\n" for hi, rb, ipt, b in rows)

        # 2. Device-wide passes on the 4096-segment batch, bucket rows in force.
        sizes = workloads.log_uniform_sizes(4096, 0, 16, seed=4)
        starts = np.zeros(len(sizes) + 1, dtype=np.int64)
        starts[1:] = np.cumsum(sizes)
        n = int(starts[-1])
        keys = workloads.uniform_keys_torch(n, 44, "cuda", torch=torch)
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        wide = {}
        for m in WIDE_MIN:
            for c in WIDE_CHUNK:
                for t in WIDE_TILE:
                    for pk in WIDE_PACK:
                        line = f"wide {m} {c} {t} {pk}"
                        eng = engine_on(torch, tmp, bucket_text + line + "\n")
                        wide[line] = time_plan(torch, eng, keys, vals, starts, args.reps)
                        eng.close()
                        print(f"[tune] 4096 segments ({n} keys) {line}: {wide[line] * 1e3:.3f} ms",
                              flush=True)
        best_wide = min(wide, key=wide.get)
        report["wide"] = {"best": best_wide, "keys": n, "ms": {k: t * 1e3 for k, t in wide.items()}}

        # 3. One launch for every size below the device-wide threshold, for
        #    each threshold (with its best device-wide line of step 2).
        single = {("buckets", best_wide): wide[best_wide]}
        for m in WIDE_MIN:
            line_m = min((line for line in wide if int(line.split()[1]) == m), key=wide.get)
            top = max(0, (m - 2).bit_length())  # ceil(log2(m - 1))
            above = "".join(f"{hi} {rb} {ipt} {b}\n" for hi, rb, ipt, b in rows if hi > top)
            for v in VARIANTS:
                rb, rest = v[1:].split("i")
                ipt, block = (int(x) for x in rest.split("b"))
                if ipt * block < (1 << top):
                    continue  # the largest such segment would not fit the tile
                text = f"{top} {rb} {ipt} {block}\n" + above
                eng = engine_on(torch, tmp, text + line_m + "\n")
                single[(text, line_m)] = time_plan(torch, eng, keys, vals, starts, args.reps)
                eng.close()
                print(f"[tune] 4096 segments ({n} keys) {line_m}, one launch of {v} up to 2^{top}: "
                      f"{single[(text, line_m)] * 1e3:.3f} ms", flush=True)
        best_rows, best_wide = min(single, key=single.get)
        if best_rows != "buckets":
            bucket_text = best_rows
            rows = [tuple(int(x) for x in line.split()) for line in best_rows.splitlines()]
        report["single"] = {"best": [best_rows, best_wide],
                            "ms": {f"{k[0]}|{k[1]}": t * 1e3 for k, t in single.items()}}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)

    with open(args.out, "w") as f:
        f.write("# segmented-sort tuning table, written by mgpu_old_amd/tune_segsort.py (key + 1 value,\n"
                "# bits [0,32)): <max ceil(log2 n)> <radix bits> <keys per thread> <threads> per\n"
                "# size bucket of the persistent kernel; wide <min keys> <count piece> <tile>\n"
                "# <packed|unpacked>\n")
        f.write(bucket_text)
        f.write(best_wide + "\n")
    print(json.dumps({"table": rows, "wide": best_wide, "report": report}), flush=True)


if __name__ == "__main__":
    main()
