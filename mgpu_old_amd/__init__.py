"""MI355X (gfx950) LSD radix sort behind the mgpusort C API.

Python mirror of the reference's host interface (inc/mgpusort.h and the C++
wrapper inc/mgpusort.hpp of seanbaxter/mgpu_old), bound with ctypes to
libmgpusort.so.  The C library and the gfx950 code objects are the product;
this module only marshals arguments.  There is no CPU fallback: if the native
library is missing, importing the binding raises.
"""
import ctypes
import os

__all__ = [
    "SortStatus", "MgpuError", "SortData", "lib", "library_path", "kernel_path",
    "Engine", "MgpuSortData", "plan_passes", "SortSegment", "SegmentedPlan", "SortRequest",
    "SortCompletion", "Server",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")


def library_path():
    """libmgpusort.so beside the code objects, or the build MGPUSORT_LIBRARY
    names (a sanitizer build, say; the code objects still come from LIB_DIR)."""
    return os.environ.get("MGPUSORT_LIBRARY") or os.path.join(LIB_DIR, "libmgpusort.so")


def kernel_path():
    """Prefix from which the engine loads its code objects."""
    return LIB_DIR + os.sep


class SortStatus:
    """sortStatus_t (inc/mgpusort.h:21-36)."""
    SUCCESS = 0
    NOT_INITIALIZED = 1
    DEVICE_ALLOC_FAILED = 2
    HOST_ALLOC_FAILED = 3
    CONFIG_NOT_SUPPORTED = 4
    CONTEXT_MISMATCH = 5
    INVALID_CONTEXT = 6
    KERNEL_NOT_FOUND = 7
    KERNEL_ERROR = 8
    LAUNCH_ERROR = 9
    INVALID_VALUE = 10
    DEVICE_ERROR = 11
    INTERNAL_ERROR = 12
    UNSUPPORTED_DEVICE = 13


SORT_VALUE_INDEX = -1
KEYS_UINT32 = 0     # SORT_KEYS_UINT32: sortTransformKeys / sortArrayTyped key types
KEYS_INT32 = 1      # SORT_KEYS_INT32
KEYS_FLOAT32 = 2    # SORT_KEYS_FLOAT32
SORT_MAX_ELEMENTS = (1 << 30) - 1


class SortData(ctypes.Structure):
    """sortData_d (inc/mgpusort.h:63-147), LP64 layout: 144 bytes."""
    _fields_ = [
        ("maxElements", ctypes.c_int),
        ("numElements", ctypes.c_int),
        ("valueCount", ctypes.c_int),
        ("firstBit", ctypes.c_int),
        ("endBit", ctypes.c_int),
        ("preserveEndKeys", ctypes.c_bool),
        ("earlyExit", ctypes.c_bool),
        ("parity", ctypes.c_int),
        ("keys", ctypes.c_uint64 * 2),
        ("values", (ctypes.c_uint64 * 2) * 6),
    ]


class SortSegment(ctypes.Structure):
    """sortSegment_d: one array of a segmented batch with its own buffers."""
    _fields_ = [
        ("keys", ctypes.c_uint64 * 2),
        ("values", (ctypes.c_uint64 * 2) * 6),
        ("numElements", ctypes.c_int),
    ]


class SortRequest(ctypes.Structure):
    """sortRequest_d: one request to the sort server (sorted in place)."""
    _fields_ = [
        ("keys", ctypes.c_uint64),
        ("values", ctypes.c_uint64 * 6),
        ("numElements", ctypes.c_int),
        ("valueCount", ctypes.c_int),
        ("firstBit", ctypes.c_int),
        ("endBit", ctypes.c_int),
        ("tag", ctypes.c_ulonglong),
    ]


class SortCompletion(ctypes.Structure):
    """sortCompletion_d: outcome and host timestamps of one request."""
    _fields_ = [
        ("tag", ctypes.c_ulonglong),
        ("status", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("submitMs", ctypes.c_double),
        ("launchMs", ctypes.c_double),
        ("doneMs", ctypes.c_double),
    ]


class KernelTiming(ctypes.Structure):
    """sortKernelTiming_t: per-kernel aggregate of event-timed launches."""
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("launches", ctypes.c_int),
        ("totalMs", ctypes.c_double),
        ("minMs", ctypes.c_double),
        ("maxMs", ctypes.c_double),
    ]


class EngineInfo(ctypes.Structure):
    """sortEngineInfo_t: what the engine chose at creation, and the last
    device inconsistency code (not cleared)."""
    _fields_ = [
        ("device", ctypes.c_int),
        ("numCUs", ctypes.c_int),
        ("xcdSegments", ctypes.c_int),
        ("rankMethod", ctypes.c_int),
        ("smallTileKeys", ctypes.c_int),
        ("largeTileKeys", ctypes.c_int),
        ("deviceError", ctypes.c_uint),
        ("reserved", ctypes.c_int),
    ]


_lib = None


def lib():
    """The loaded libmgpusort.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise OSError(f"libmgpusort.so not built at {path}; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        eng = ctypes.c_void_p
        sd = P(SortData)
        sigs = {
            "sortStatusString": (ctypes.c_char_p, [ctypes.c_int]),
            "sortCreateEngine": (ctypes.c_int, [ctypes.c_char_p, P(eng)]),
            "sortLoadKernel": (ctypes.c_int, [eng, ctypes.c_int, ctypes.c_int, ctypes.c_bool,
                                              ctypes.c_int]),
            "sortIncEngine": (ctypes.c_int, [eng]),
            "sortReleaseEngine": (ctypes.c_int, [eng]),
            "sortDestroyEngine": (ctypes.c_int, [eng]),
            "sortCreateData": (ctypes.c_int, [eng, ctypes.c_int, ctypes.c_int, P(sd)]),
            "sortDestroyData": (ctypes.c_int, [sd]),
            "sortAllocData": (ctypes.c_int, [eng, sd]),
            "sortFreeData": (ctypes.c_int, [sd]),
            "sortArray": (ctypes.c_int, [eng, sd]),
            "sortArrayTyped": (ctypes.c_int, [eng, sd, ctypes.c_int]),
            "sortArrayEx": (ctypes.c_int, [eng, sd, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_bool]),
            "sortHost": (ctypes.c_int, [eng, P(ctypes.c_uint32), P(ctypes.c_uint32),
                                        ctypes.c_int, ctypes.c_int]),
            "sortDevice": (ctypes.c_int, [eng, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_int]),
            "sortSetStream": (ctypes.c_int, [eng, ctypes.c_void_p]),
            "sortGetStream": (ctypes.c_int, [eng, P(ctypes.c_void_p)]),
            "sortPlanPasses": (ctypes.c_int, [ctypes.c_int, P(ctypes.c_int)]),
            "sortSegmentedArray": (ctypes.c_int, [eng, sd, ctypes.c_int, P(ctypes.c_int)]),
            "sortSegmentedBatch": (ctypes.c_int, [eng, P(SortSegment), ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "sortSegmentedPlanCreate": (ctypes.c_int, [eng, sd, ctypes.c_int, P(ctypes.c_int),
                                                       P(ctypes.c_void_p)]),
            "sortSegmentedPlanExecute": (ctypes.c_int, [ctypes.c_void_p, sd]),
            "sortSegmentedPlanDestroy": (ctypes.c_int, [ctypes.c_void_p]),
            "sortServerCreate": (ctypes.c_int, [eng, ctypes.c_ulonglong, ctypes.c_int,
                                                ctypes.c_int, P(ctypes.c_void_p)]),
            "sortServerSubmit": (ctypes.c_int, [ctypes.c_void_p, P(SortRequest)]),
            "sortServerPoll": (ctypes.c_int, [ctypes.c_void_p, P(SortCompletion), ctypes.c_int,
                                              ctypes.c_int]),
            "sortServerReplay": (ctypes.c_int, [ctypes.c_void_p, P(SortRequest),
                                                P(ctypes.c_double), ctypes.c_int,
                                                P(SortCompletion)]),
            "sortServerDestroy": (ctypes.c_int, [ctypes.c_void_p]),
            "sortCheckDevice": (ctypes.c_int, [eng]),
            "sortEngineInfo": (ctypes.c_int, [eng, P(EngineInfo)]),
            "sortTransformKeys": (ctypes.c_int, [eng, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_bool]),
            "sortEnableTiming": (ctypes.c_int, [eng, ctypes.c_bool]),
            "sortReadTiming": (ctypes.c_int, [eng, P(KernelTiming), ctypes.c_int]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(status):
    s = lib().sortStatusString(int(status))
    return s.decode() if s else None


class MgpuError(RuntimeError):
    """MgpuException (inc/mgpusort.hpp:12-20): a non-success sortStatus_t."""

    def __init__(self, status, what=""):
        self.status = int(status)
        super().__init__(f"{what}{': ' if what else ''}{status_string(status)}")


def _check(status, what=""):
    if status != SortStatus.SUCCESS:
        raise MgpuError(status, what)


def plan_passes(num_bits):
    """Digit widths sortArray uses for a num_bits-bit range (host only)."""
    widths = (ctypes.c_int * 32)()
    n = lib().sortPlanPasses(int(num_bits), widths)
    return [widths[i] for i in range(n)]


class Engine:
    """MgpuSort (inc/mgpusort.hpp:79-114): a reference-counted engine handle."""

    def __init__(self, kernel_path_prefix=None):
        self._h = ctypes.c_void_p()
        path = kernel_path() if kernel_path_prefix is None else kernel_path_prefix
        _check(lib().sortCreateEngine(path.encode() if path else None, ctypes.byref(self._h)),
               "sortCreateEngine")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().sortReleaseEngine(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_stream(self, stream_handle):
        _check(lib().sortSetStream(self._h, ctypes.c_void_p(stream_handle)), "sortSetStream")

    def check_device(self):
        """Raise if a kernel flagged an inconsistency since the last check."""
        _check(lib().sortCheckDevice(self._h), "sortCheckDevice")

    def transform_keys(self, keys_ptr, n, key_type, to_radix):
        """sortTransformKeys: int32 (KEYS_INT32) or float32 (KEYS_FLOAT32) keys
        at a device pointer to the sort's unsigned order (to_radix) or back."""
        _check(lib().sortTransformKeys(self._h, keys_ptr, n, key_type, bool(to_radix)),
               "sortTransformKeys")

    def info(self):
        """sortEngineInfo as a dict."""
        out = EngineInfo()
        _check(lib().sortEngineInfo(self._h, ctypes.byref(out)), "sortEngineInfo")
        return {name: getattr(out, name) for name, _ in EngineInfo._fields_
                if not name.startswith("reserved")}

    def enable_timing(self, enable=True):
        _check(lib().sortEnableTiming(self._h, bool(enable)), "sortEnableTiming")

    def read_timing(self):
        """{kernel name: (launches, total ms, min ms, max ms)}; clears the record."""
        buf = (KernelTiming * 64)()
        n = lib().sortReadTiming(self._h, buf, 64)
        if n < 0:
            raise MgpuError(SortStatus.DEVICE_ERROR, "sortReadTiming")
        return {buf[i].name.decode(): (buf[i].launches, buf[i].totalMs, buf[i].minMs,
                                       buf[i].maxMs) for i in range(min(n, 64))}

    def sort_array(self, data):
        _check(lib().sortArray(self._h, ctypes.byref(data)), "sortArray")

    def sort_array_typed(self, data, key_type):
        """sortArrayTyped: sortArray over int32 (KEYS_INT32) or float32
        (KEYS_FLOAT32) keys in numeric order, the order mapping done inside
        the passes."""
        _check(lib().sortArrayTyped(self._h, ctypes.byref(data), int(key_type)), "sortArrayTyped")

    def sort_segmented(self, data, starts):
        """sortSegmentedArray: starts = numSegments + 1 offsets into data."""
        arr = (ctypes.c_int * len(starts))(*[int(x) for x in starts])
        _check(lib().sortSegmentedArray(self._h, ctypes.byref(data), len(starts) - 1, arr),
               "sortSegmentedArray")

    def sort_batch(self, segments, value_count, first_bit=0, end_bit=32):
        """sortSegmentedBatch over a list of SortSegment."""
        arr = (SortSegment * max(1, len(segments)))(*segments)
        _check(lib().sortSegmentedBatch(self._h, arr, len(segments), value_count, first_bit,
                                        end_bit), "sortSegmentedBatch")

    def segmented_plan(self, data, starts):
        return SegmentedPlan(self, data, starts)

    def sort_array_ex(self, data, num_threads=256, values_per_thread=8, bit_pass=6,
                      use_trans_list=False):
        _check(lib().sortArrayEx(self._h, ctypes.byref(data), num_threads, values_per_thread,
                                 bit_pass, use_trans_list), "sortArrayEx")


class SegmentedPlan:
    """sortSegmentedPlanCreate / Execute / Destroy: a graph-captured segmented
    sort replayed on the engine's stream."""

    def __init__(self, engine, data, starts):
        self._engine = engine
        self._h = ctypes.c_void_p()
        arr = (ctypes.c_int * len(starts))(*[int(x) for x in starts])
        _check(lib().sortSegmentedPlanCreate(engine.handle, ctypes.byref(data), len(starts) - 1,
                                             arr, ctypes.byref(self._h)),
               "sortSegmentedPlanCreate")

    def execute(self, data):
        _check(lib().sortSegmentedPlanExecute(self._h, ctypes.byref(data)),
               "sortSegmentedPlanExecute")

    def close(self):
        if self._h:
            lib().sortSegmentedPlanDestroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """sortServerCreate / Submit / Poll / Replay / Destroy."""

    def __init__(self, engine, hbm_budget_bytes, max_in_flight=3, max_elements=1 << 24):
        self._engine = engine
        self._h = ctypes.c_void_p()
        _check(lib().sortServerCreate(engine.handle, int(hbm_budget_bytes), int(max_in_flight),
                                      int(max_elements), ctypes.byref(self._h)),
               "sortServerCreate")

    def submit(self, request):
        _check(lib().sortServerSubmit(self._h, ctypes.byref(request)), "sortServerSubmit")

    def poll(self, max_count=256, timeout_ms=0):
        buf = (SortCompletion * max_count)()
        n = lib().sortServerPoll(self._h, buf, max_count, int(timeout_ms))
        if n < 0:
            raise MgpuError(SortStatus.INVALID_VALUE, "sortServerPoll")
        return [buf[i] for i in range(n)]

    def replay(self, requests, arrival_ms):
        """Submit requests[i] at arrival_ms[i]; returns the completions in order."""
        n = len(requests)
        reqs = (SortRequest * max(1, n))(*requests)
        arr = (ctypes.c_double * max(1, n))(*[float(x) for x in arrival_ms])
        out = (SortCompletion * max(1, n))()
        _check(lib().sortServerReplay(self._h, reqs, arr, n, out), "sortServerReplay")
        return [out[i] for i in range(n)]

    def close(self):
        if self._h:
            lib().sortServerDestroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class MgpuSortData(SortData):
    """MgpuSortData (inc/mgpusort.hpp:25-75): tracks caller-attached buffers so
    that reset() frees only what the library allocated."""

    def __init__(self):
        super().__init__()
        self._attached_keys = [False, False]
        self._attached_vals = [[False, False] for _ in range(6)]

    def attach_key(self, ptr, i=0):
        self.keys[i] = ptr
        self._attached_keys[i ^ self.parity] = True

    def attach_val(self, j, ptr, i=0):
        self.values[j][i] = ptr
        self._attached_vals[j][i ^ self.parity] = True

    def alloc(self, engine, max_elements, value_count):
        self.maxElements = max_elements
        self.numElements = max_elements
        self.valueCount = value_count
        _check(lib().sortAllocData(engine.handle, ctypes.byref(self)), "sortAllocData")

    def reset(self):
        for i in range(2):
            if self._attached_keys[i]:
                self.keys[i ^ self.parity] = 0
            for j in range(6):
                if self._attached_vals[j][i]:
                    self.values[j][i ^ self.parity] = 0
        lib().sortFreeData(ctypes.byref(self))
        self._attached_keys = [False, False]
        self._attached_vals = [[False, False] for _ in range(6)]
