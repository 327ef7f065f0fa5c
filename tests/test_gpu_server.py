"""Sort server (continuous batching) parity, needs a gfx950 GPU: every served
request must equal the oracle's stable sort of that request alone, and the
same request sorted alone by sortArray, whatever it was batched with."""
import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import GpuSort, make_engine, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    eng = make_engine()
    yield eng
    eng.close()


def _make_requests(count, seed, lo_log2=10, hi_log2=20):
    sizes = workloads.log_uniform_sizes(count, lo_log2, hi_log2, seed=seed)
    choice = workloads.words(seed + 1, 0, 2 * count)
    widths = np.array([16, 24, 32])[choice[:count] % 3]
    u = choice[count:] % 100
    vcs = np.where(u < 50, 0, np.where(u < 85, 1, 2))
    reqs, host, tensors = [], [], []
    for i in range(count):
        n, w, vc = int(sizes[i]), int(widths[i]), int(vcs[i])
        keys = workloads.uniform_keys(n, seed=seed * 1000 + i, mask_bits=w)
        vals = [workloads.uniform_keys(n, seed=seed * 1000 + 500 + i * 7 + j) for j in range(vc)]
        tk = to_dev(keys)
        tv = [to_dev(v) for v in vals]
        r = mg.SortRequest()
        r.keys = tk.data_ptr()
        for j, t in enumerate(tv):
            r.values[j] = t.data_ptr()
        r.numElements, r.valueCount, r.firstBit, r.endBit = n, vc, 0, w
        reqs.append(r)
        host.append((keys, vals, w))
        tensors.append((tk, tv))
    return reqs, host, tensors


def _poisson(count, mean_gap_ms, seed):
    u = (workloads.words(seed, 0, count).astype(np.float64) + 0.5) / 4294967296.0
    return np.cumsum(-np.log(u) * mean_gap_ms)


def test_served_requests_match_oracle_and_alone(engine, oracle):
    import torch
    reqs, host, tensors = _make_requests(48, seed=7)
    arrivals = _poisson(len(reqs), 0.2, seed=8)
    torch.cuda.synchronize()
    with mg.Server(engine, 1 << 30, max_in_flight=3, max_elements=1 << 20) as srv:
        comps = srv.replay(reqs, arrivals)
    torch.cuda.synchronize()
    batches = set()
    for i, c in enumerate(comps):
        assert c.status == 0, f"request {i} failed: {mg.status_string(c.status)}"
        assert c.doneMs >= c.launchMs >= c.submitMs - 1e-6
        batches.add(c.batch)
        keys, vals, w = host[i]
        tk, tv = tensors[i]
        n = keys.size
        rk, rv = oracle.stable_sort(keys, vals, first_bit=0, end_bit=w)
        assert np.array_equal(to_host(tk, n), rk), f"request {i} keys differ"
        for j, t in enumerate(tv):
            assert np.array_equal(to_host(t, n), rv[j]), f"request {i} values[{j}] differ"
        if i % 6 == 0:  # the same request alone through sortArray
            alone = GpuSort(keys, vals, first_bit=0, end_bit=w).run(engine)
            assert np.array_equal(alone.keys(), rk)
    assert len(batches) > 1  # work really was split into several batches


def test_budget_admission_and_rejection(engine, oracle):
    import torch
    reqs, host, tensors = _make_requests(12, seed=21, lo_log2=14, hi_log2=18)
    big = max(r.numElements * 4 * (1 + max(r.valueCount, 0)) for r in reqs)
    # Room for about one large request at a time: admission must serialize.
    budget = big + (1 << 20)
    too_big = mg.SortRequest()
    extra = to_dev(workloads.uniform_keys(1 << 20, 3))
    too_big.keys = extra.data_ptr()
    too_big.numElements, too_big.valueCount, too_big.firstBit, too_big.endBit = 1 << 20, 0, 0, 32
    torch.cuda.synchronize()
    with mg.Server(engine, budget, max_in_flight=2, max_elements=1 << 20) as srv:
        comps = srv.replay(reqs + [too_big], [0.0] * (len(reqs) + 1))
    assert comps[-1].status == mg.SortStatus.CONFIG_NOT_SUPPORTED
    for i, c in enumerate(comps[:-1]):
        assert c.status == 0
        keys, vals, w = host[i]
        rk, rv = oracle.stable_sort(keys, vals, first_bit=0, end_bit=w)
        assert np.array_equal(to_host(tensors[i][0], keys.size), rk)


def test_submit_and_poll(engine, oracle):
    import torch
    reqs, host, tensors = _make_requests(10, seed=31, lo_log2=10, hi_log2=16)
    torch.cuda.synchronize()
    with mg.Server(engine, 1 << 28, max_in_flight=2, max_elements=1 << 16) as srv:
        for i, r in enumerate(reqs):
            r.tag = 100 + i
            srv.submit(r)
        got = {}
        while len(got) < len(reqs):
            for c in srv.poll(timeout_ms=100):
                got[c.tag] = c
        bad = mg.SortRequest()
        bad.numElements = (1 << 16) + 1
        with pytest.raises(mg.MgpuError):
            srv.submit(bad)
    for i in range(len(reqs)):
        assert got[100 + i].status == 0
        keys, vals, w = host[i]
        rk, _ = oracle.stable_sort(keys, vals, first_bit=0, end_bit=w)
        assert np.array_equal(to_host(tensors[i][0], keys.size), rk)


def test_batch_limits(engine, oracle):
    """A batch holds at most 1024 requests and, unless it is one request,
    at most 2^26 keys; requests past either limit run in later batches with
    the same results."""
    import torch
    # 1100 tiny requests at once on one slot: at least two batches.
    sizes = workloads.log_uniform_sizes(1100, 0, 6, seed=41)
    reqs, host, tensors = [], [], []
    for i, n in enumerate(sizes):
        keys = workloads.uniform_keys(int(n), seed=4100 + i)
        vals = [workloads.uniform_keys(int(n), seed=9100 + i)]
        tk, tv = to_dev(keys), to_dev(vals[0])
        r = mg.SortRequest()
        r.keys, r.values[0] = tk.data_ptr(), tv.data_ptr()
        r.numElements, r.valueCount, r.firstBit, r.endBit = int(n), 1, 0, 32
        reqs.append(r)
        host.append((keys, vals))
        tensors.append((tk, tv))
    torch.cuda.synchronize()
    with mg.Server(engine, 1 << 26, max_in_flight=1, max_elements=1 << 10) as srv:
        comps = srv.replay(reqs, np.zeros(len(reqs)))
    torch.cuda.synchronize()
    per_batch = {}
    for i, c in enumerate(comps):
        assert c.status == 0, f"request {i} failed: {mg.status_string(c.status)}"
        per_batch[c.batch] = per_batch.get(c.batch, 0) + 1
        keys, vals = host[i]
        rk, rv = oracle.stable_sort(keys, vals)
        assert np.array_equal(to_host(tensors[i][0], keys.size), rk), f"request {i} keys"
        assert np.array_equal(to_host(tensors[i][1], keys.size), rv[0]), f"request {i} values"
    assert len(per_batch) >= 2 and max(per_batch.values()) <= 1024

    # Five keys-only requests just under 2^24 keys (80M in all) at once on
    # one slot: no batch of several requests passes 2^26 keys.
    n = (1 << 24) - 1000
    reqs, host, tensors = [], [], []
    for i in range(5):
        keys = workloads.uniform_keys(n, seed=5100 + i)
        tk = to_dev(keys)
        r = mg.SortRequest()
        r.keys = tk.data_ptr()
        r.numElements, r.valueCount, r.firstBit, r.endBit = n, 0, 0, 32
        reqs.append(r)
        host.append(keys)
        tensors.append(tk)
    torch.cuda.synchronize()
    with mg.Server(engine, 1 << 30, max_in_flight=1, max_elements=1 << 24) as srv:
        comps = srv.replay(reqs, np.zeros(len(reqs)))
    torch.cuda.synchronize()
    keys_per_batch, reqs_per_batch = {}, {}
    for i, c in enumerate(comps):
        assert c.status == 0, f"request {i} failed: {mg.status_string(c.status)}"
        keys_per_batch[c.batch] = keys_per_batch.get(c.batch, 0) + n
        reqs_per_batch[c.batch] = reqs_per_batch.get(c.batch, 0) + 1
        # Keys alone: the stable sort is the plain sorted order.
        assert np.array_equal(to_host(tensors[i], n), np.sort(host[i])), f"request {i} keys"
    for b, k in keys_per_batch.items():
        assert reqs_per_batch[b] == 1 or k <= 1 << 26


def test_submit_from_several_threads(engine, oracle):
    """sortServerSubmit from four threads at once (ctypes drops the GIL in the
    call) while the main thread polls: every request completes once, sorted."""
    import threading
    import time

    import torch
    reqs, host, tensors = _make_requests(64, seed=51, lo_log2=8, hi_log2=17)
    torch.cuda.synchronize()
    errors = []
    with mg.Server(engine, 1 << 30, max_in_flight=3, max_elements=1 << 17) as srv:
        def submit(part):
            try:
                for i in range(part, len(reqs), 4):
                    reqs[i].tag = 1000 + i
                    srv.submit(reqs[i])
            except Exception as e:  # reported below
                errors.append(e)

        threads = [threading.Thread(target=submit, args=(p,)) for p in range(4)]
        for t in threads:
            t.start()
        got = {}
        deadline = time.monotonic() + 60
        while len(got) < len(reqs) and not errors and time.monotonic() < deadline:
            for c in srv.poll(timeout_ms=100):
                assert c.tag not in got, f"request {c.tag} completed twice"
                got[c.tag] = c
        for t in threads:
            t.join()
    assert not errors, errors
    assert len(got) == len(reqs), f"{len(reqs) - len(got)} requests never completed"
    torch.cuda.synchronize()
    for i in range(len(reqs)):
        assert got[1000 + i].status == 0
        keys, vals, w = host[i]
        rk, rv = oracle.stable_sort(keys, vals, first_bit=0, end_bit=w)
        assert np.array_equal(to_host(tensors[i][0], keys.size), rk), f"request {i} keys"
        for j, t in enumerate(tensors[i][1]):
            assert np.array_equal(to_host(t, keys.size), rv[j]), f"request {i} values[{j}]"


def test_served_full_size_law(engine, oracle):
    """The served workload's whole size law, [2^10, 2^24]: beside requests
    drawn from it, a dozen in (2^20, 2^24] with one or two value arrays and
    16-, 24- and 32-bit keys, arriving as a Poisson stream; every request
    bit-exact against the oracle's sort of that request alone, and every
    completion SUCCESS."""
    import torch
    reqs, host, tensors = _make_requests(24, seed=41, lo_log2=10, hi_log2=24)
    sizes = workloads.log_uniform_sizes(12, 20, 24, seed=42, lo=(1 << 20) + 1)
    for i, n in enumerate(sizes):
        n, w, vc = int(n), (16, 24, 32)[i % 3], 1 + (i % 2)
        keys = workloads.uniform_keys(n, seed=4300 + i, mask_bits=w)
        vals = [workloads.uniform_keys(n, seed=4400 + 7 * i + j) for j in range(vc)]
        tk, tv = to_dev(keys), [to_dev(v) for v in vals]
        r = mg.SortRequest()
        r.keys = tk.data_ptr()
        for j, t in enumerate(tv):
            r.values[j] = t.data_ptr()
        r.numElements, r.valueCount, r.firstBit, r.endBit = n, vc, 0, w
        reqs.append(r)
        host.append((keys, vals, w))
        tensors.append((tk, tv))
    order = np.argsort(workloads.words(44, 0, len(reqs)), kind="stable")
    reqs = [reqs[i] for i in order]
    host = [host[i] for i in order]
    tensors = [tensors[i] for i in order]
    arrivals = _poisson(len(reqs), 1.0, seed=45)
    torch.cuda.synchronize()
    with mg.Server(engine, 1 << 31, max_in_flight=4, max_elements=1 << 24) as srv:
        comps = srv.replay(reqs, arrivals)
    torch.cuda.synchronize()
    big = 0
    for i, c in enumerate(comps):
        assert c.status == 0, f"request {i} failed: {mg.status_string(c.status)}"
        keys, vals, w = host[i]
        tk, tv = tensors[i]
        n = keys.size
        big += n > (1 << 20)
        rk, rv = oracle.stable_sort(keys, vals, first_bit=0, end_bit=w)
        assert np.array_equal(to_host(tk, n), rk), f"request {i} (n={n}, {w} bits) keys differ"
        for j, t in enumerate(tv):
            assert np.array_equal(to_host(t, n), rv[j]), f"request {i} values[{j}] differ"
    assert big >= 12
