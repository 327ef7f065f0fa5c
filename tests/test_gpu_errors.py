"""Kernel self-check failures reach the C ABI (needs a gfx950 GPU).

The failure is forced from the host: an engine created with
MGPUSORT_SPIN_LIMIT=0 gives up every lookback at once, as a wait that timed
out would (error code 1), and every kernel still runs to its end.  Only
passes with more than one tile in a lookback chain look back at all, so a
one-tile sort and small segments stay correct and
report SUCCESS, while everything with a multi-tile pass reports
SORT_STATUS_KERNEL_ERROR through the call the contract names:
  * sortArray / sortArrayEx / sortSegmentedPlanExecute: at their next entry
    (before queueing anything), or sortCheckDevice;
  * sortHost, sortDevice, sortSegmentedArray, sortSegmentedBatch: on return;
  * the sort server: in the completion of each request of the failed batch.
"""
import ctypes

import numpy as np
import pytest

import mgpu_old_amd as mg
from mgpu_old_amd import workloads
from gpu_helpers import GpuSort, empty_dev, make_engine, to_dev, to_host

pytestmark = pytest.mark.gpu

KERNEL_ERROR = mg.SortStatus.KERNEL_ERROR
SUCCESS = mg.SortStatus.SUCCESS


@pytest.fixture(scope="module")
def failing():
    eng = make_engine({"MGPUSORT_SPIN_LIMIT": "0"})
    yield eng
    eng.close()


def _sync():
    import torch
    torch.cuda.synchronize()


def _raw_sort(engine, data):
    return mg.lib().sortArray(engine.handle, ctypes.byref(data))


def test_sort_array_reports_at_next_entry(failing):
    keys = workloads.uniform_keys(1 << 20, seed=11)
    g = GpuSort(keys)
    _sync()
    assert _raw_sort(failing, g.data) == SUCCESS  # queued; fails on the device
    _sync()
    before = (g.data.keys[0], g.data.keys[1], g.data.parity)
    assert _raw_sort(failing, g.data) == KERNEL_ERROR  # the earlier sort's failure
    assert (g.data.keys[0], g.data.keys[1], g.data.parity) == before, "nothing may run"
    assert mg.lib().sortCheckDevice(failing.handle) == SUCCESS  # reported once
    ex = mg.lib().sortArrayEx
    assert ex(failing.handle, ctypes.byref(g.data), 256, 8, 6, False) == SUCCESS
    _sync()
    assert failing.info()["deviceError"] == 1  # a lookback wait gave up
    assert mg.lib().sortCheckDevice(failing.handle) == KERNEL_ERROR
    assert mg.lib().sortCheckDevice(failing.handle) == SUCCESS


def test_one_tile_sort_is_unaffected(failing, oracle):
    keys = workloads.uniform_keys(5000, seed=12)
    vals = np.arange(5000, dtype=np.uint32)
    g = GpuSort(keys, [vals]).run(failing)
    assert mg.lib().sortCheckDevice(failing.handle) == SUCCESS
    rk, (rv,) = oracle.stable_sort(keys, [vals])
    assert np.array_equal(g.keys(), rk) and np.array_equal(g.values()[0], rv)


def test_sort_host_and_device_report_their_own_failure(failing):
    n = 1 << 20
    keys = workloads.uniform_keys(n, seed=13)
    host = np.ascontiguousarray(keys.copy())
    p = host.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    assert mg.lib().sortHost(failing.handle, p, None, n, 32) == KERNEL_ERROR
    t = to_dev(keys)
    _sync()
    assert mg.lib().sortDevice(failing.handle, t.data_ptr(), 0, n, 32) == KERNEL_ERROR
    assert mg.lib().sortCheckDevice(failing.handle) == SUCCESS


def test_segmented_calls_report(failing, oracle):
    # Only small segments: sorted in LDS, no lookback wait: SUCCESS.
    sizes = [1, 70, 4096, 3000, 0, 900]
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(starts[-1])
    keys = workloads.uniform_keys(n, seed=14)
    vals = [np.arange(n, dtype=np.uint32)]
    g = GpuSort(keys, vals)
    failing.sort_segmented(g.data, starts)
    out_k, (out_v,) = g.keys(), g.values()
    for i in range(len(sizes)):
        a, b = int(starts[i]), int(starts[i + 1])
        rk, (rv,) = oracle.stable_sort(keys[a:b], [vals[0][a:b]])
        assert np.array_equal(out_k[a:b], rk) and np.array_equal(out_v[a:b], rv)
    # A segment with several tiles takes the device-wide passes: KERNEL_ERROR.
    sizes = [100, 200_000]
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    g = GpuSort(workloads.uniform_keys(int(starts[-1]), seed=15))
    arr = (ctypes.c_int * len(starts))(*[int(x) for x in starts])
    rc = mg.lib().sortSegmentedArray(failing.handle, ctypes.byref(g.data), len(sizes), arr)
    assert rc == KERNEL_ERROR
    seg = mg.SortSegment()
    t = [to_dev(workloads.uniform_keys(200_000, seed=16)), empty_dev(200_000)]
    seg.keys[0], seg.keys[1] = t[0].data_ptr(), t[1].data_ptr()
    seg.numElements = 200_000
    _sync()
    rc = mg.lib().sortSegmentedBatch(failing.handle, ctypes.byref(seg), 1, 0, 0, 32)
    assert rc == KERNEL_ERROR
    assert mg.lib().sortCheckDevice(failing.handle) == SUCCESS


def test_plan_execute_reports_at_next_entry(failing):
    sizes = [10, 100_000]
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(starts[-1])
    g = GpuSort(workloads.uniform_keys(n, seed=17))
    plan = failing.segmented_plan(g.data, starts)
    ex = mg.lib().sortSegmentedPlanExecute
    _sync()
    assert ex(plan._h, ctypes.byref(g.data)) == SUCCESS
    _sync()
    assert ex(plan._h, ctypes.byref(g.data)) == KERNEL_ERROR
    assert ex(plan._h, ctypes.byref(g.data)) == SUCCESS
    _sync()
    assert mg.lib().sortCheckDevice(failing.handle) == KERNEL_ERROR
    plan.close()


def _request(keys_t, n, vc=0, vals=()):
    r = mg.SortRequest()
    r.keys = keys_t.data_ptr()
    for j, t in enumerate(vals):
        r.values[j] = t.data_ptr()
    r.numElements, r.valueCount, r.firstBit, r.endBit = n, vc, 0, 32
    return r


def test_server_reports_per_batch(failing, oracle):
    """A batch whose device-wide pass fails completes with KERNEL_ERROR; the
    next batch (small requests only, on a slot whose word was cleared)
    completes with SUCCESS and correct results."""
    big = to_dev(workloads.uniform_keys(1 << 20, seed=18))
    small_keys = [workloads.uniform_keys(n, seed=19 + i) for i, n in enumerate([5, 700, 4096,
                                                                                 2500])]
    small = [to_dev(k) for k in small_keys]
    _sync()
    with mg.Server(failing, 1 << 28, max_in_flight=1, max_elements=1 << 20) as srv:
        r = _request(big, 1 << 20)
        r.tag = 1
        srv.submit(r)
        got = []
        while not got:
            got = srv.poll(timeout_ms=1000)
        assert got[0].tag == 1 and got[0].status == KERNEL_ERROR
        for i, (t, k) in enumerate(zip(small, small_keys)):
            r = _request(t, k.size)
            r.tag = 10 + i
            srv.submit(r)
        done = []
        while len(done) < len(small):
            done += srv.poll(timeout_ms=1000)
        assert all(c.status == SUCCESS for c in done), [c.status for c in done]
    _sync()
    for t, k in zip(small, small_keys):
        rk, _ = oracle.stable_sort(k)
        assert np.array_equal(to_host(t, k.size), rk)


def test_engines_have_their_own_error_word(failing, oracle):
    """A healthy engine beside the failing one reports nothing."""
    eng = make_engine()
    try:
        keys = workloads.uniform_keys(1 << 20, seed=23)
        GpuSort(keys).run(failing)  # sets the failing engine's word
        g = GpuSort(keys).run(eng)
        assert mg.lib().sortCheckDevice(eng.handle) == SUCCESS
        rk, _ = oracle.stable_sort(keys)
        assert np.array_equal(g.keys(), rk)
        assert mg.lib().sortCheckDevice(failing.handle) == KERNEL_ERROR
    finally:
        eng.close()
