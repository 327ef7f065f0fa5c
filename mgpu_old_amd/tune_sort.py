"""Autotuner of sortArray: picks, per size class (ceil(log2 n)) and value
class (keys only, index generation, one value array, several), the tile shape
(512 x 16 or 1024 x 16 keys) and the lookback layout (one chain across the
device, or one per XCD) with the shortest sort, and writes the table the
engine loads at creation (kernelPath + "mgpusort_sort_tuning.txt").

It plays the role of the reference's pass tables (sort/src/mgpusort/
sorttables.cpp:16-1465, generated offline by sort/src/tablegen/tablegen.cpp:
9-64 from measured rates), but times every candidate directly on the GPU.
A candidate replaces the engine's built-in choice only when it is faster by
more than --margin, so timing noise does not flip choices.

Usage (on a gfx950 box):  python -m mgpu_old_amd.tune_sort [--out PATH]
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CLASSES = {"keys": 0, "index": -1, "kv": 1, "multi": 2}  # value class -> valueCount
CANDIDATES = {("small", "one"): {"MGPUSORT_TILE": "512x16", "MGPUSORT_XCD": "0"},
              ("large", "one"): {"MGPUSORT_TILE": "1024x16", "MGPUSORT_XCD": "0"},
              ("small", "xcd"): {"MGPUSORT_TILE": "512x16", "MGPUSORT_XCD": "1"},
              ("large", "xcd"): {"MGPUSORT_TILE": "1024x16", "MGPUSORT_XCD": "1"}}


def choose(times, builtin, margin):
    """times: {candidate: seconds} plus "builtin"; builtin: the candidate the
    built-in rules run.  The fastest candidate if it beats the built-in time
    by more than `margin`, else the built-in choice."""
    best = min((c for c in times if c != "builtin"), key=times.get)
    return best if times[best] < times["builtin"] * (1.0 - margin) else builtin


def table_rows(choices):
    """choices: {class: [(log2 size, (tile, chains)), ...] ascending} ->
    rows (max log2, class, tile, chains), one per run of equal choices; a
    class's last row covers every larger size."""
    rows = []
    for cls, seq in choices.items():
        for i, (lg, c) in enumerate(seq):
            if i + 1 == len(seq) or seq[i + 1][1] != c:
                rows.append((31 if i + 1 == len(seq) else lg, cls, c[0], c[1]))
    return rows


def make_engine(mg, torch, env):
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


def time_sort(torch, mg, eng, n, value_count, src, reps):
    """Median of `reps` event-timed sortArray calls (input copied back before
    each, outside the timed region)."""
    carried = 1 if value_count == -1 else value_count
    k = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(2)]
    v = [[torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(2)]
         for _ in range(carried)]
    d = mg.SortData()
    ts = []
    for r in range(reps + 1):
        k[0].copy_(src)
        for j in range(max(0, value_count)):
            v[j][0].copy_(src)
        d.maxElements = d.numElements = n
        d.valueCount, d.firstBit, d.endBit, d.parity = value_count, 0, 32, 0
        d.keys[0], d.keys[1] = k[0].data_ptr(), k[1].data_ptr()
        for j in range(carried):
            d.values[j][0], d.values[j][1] = v[j][0].data_ptr(), v[j][1].data_ptr()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        eng.sort_array(d)
        b.record()
        b.synchronize()
        if r:  # the first is a warm-up
            ts.append(a.elapsed_time(b) / 1e3)
    eng.check_device()
    ts.sort()
    return ts[len(ts) // 2]


def builtin_choice(eng, n, value_count, torch, mg, src):
    """The (tile, chains) the built-in rules run, from the kernel names."""
    eng.enable_timing(True)
    time_sort(torch, mg, eng, n, value_count, src, 1)
    names = [k for k in eng.read_timing() if k.startswith("mgpu_rank_scatter")]
    eng.enable_timing(False)
    return ("large" if any("_b1024_" in k for k in names) else "small",
            "xcd" if any("_x_" in k for k in names) else "one")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "lib", "mgpusort_sort_tuning.txt"))
    ap.add_argument("--min-log2", type=int, default=16)
    ap.add_argument("--max-log2", type=int, default=28)
    ap.add_argument("--reps", type=int, default=7)
    ap.add_argument("--margin", type=float, default=0.03)
    args = ap.parse_args()

    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    import mgpu_old_amd as mg
    from mgpu_old_amd import workloads

    # The table being written must not steer the measurement.
    if os.path.exists(args.out):
        os.remove(args.out)
    engines = {c: make_engine(mg, torch, env) for c, env in CANDIDATES.items()}
    engines["builtin"] = make_engine(mg, torch, {})
    choices = {cls: [] for cls in CLASSES}
    report = {}
    for lg in range(args.min_log2, args.max_log2 + 1):
        n = 1 << lg
        src = workloads.uniform_keys_torch(n, 300 + lg, "cuda", torch=torch)
        for cls, vc in CLASSES.items():
            times = {c: time_sort(torch, mg, e, n, vc, src, args.reps) for c, e in engines.items()}
            built = builtin_choice(engines["builtin"], n, vc, torch, mg, src)
            pick = choose(times, built, args.margin)
            choices[cls].append((lg, pick))
            report[f"{cls} 2^{lg}"] = {"pick": " ".join(pick), "builtin": " ".join(built),
                                      "ms": {(" ".join(c) if c != "builtin" else c): t * 1e3
                                             for c, t in times.items()}}
            print(f"[tune] {cls:5s} 2^{lg}: " + "  ".join(
                f"{' '.join(c) if c != 'builtin' else c} {t * 1e3:.3f}" for c, t in times.items())
                + f"  -> {' '.join(pick)} (built-in {' '.join(built)})", flush=True)
        del src
        torch.cuda.empty_cache()
    rows = table_rows(choices)
    with open(args.out, "w") as f:
        f.write("# sortArray tuning table: max ceil(log2 n), values (keys|index|kv|multi), "
                "tile (small|large), lookback chains (one|xcd)\n")
        f.write("# written by mgpu_old_amd/tune_sort.py (uniform keys, bits [0,32); "
                f"margin {args.margin})\n")
        for r in rows:
            f.write(" ".join(str(x) for x in r) + "\n")
    print(json.dumps({"rows": rows, "report": report}), flush=True)
    for e in engines.values():
        e.close()


if __name__ == "__main__":
    main()
