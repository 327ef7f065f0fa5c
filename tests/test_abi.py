"""CPU checks of the C ABI boundary: exported symbols, struct layout, status
strings, pass planning and argument validation that needs no device."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT
import mgpu_old_amd as mg

HEADER = os.path.join(ROOT, "include", "mgpusort.h")

# Status names in the order of inc/mgpusort.h:21-36.
REFERENCE_STATUS = [
    "SORT_STATUS_SUCCESS", "SORT_STATUS_NOT_INITIALIZED", "SORT_STATUS_DEVICE_ALLOC_FAILED",
    "SORT_STATUS_HOST_ALLOC_FAILED", "SORT_STATUS_CONFIG_NOT_SUPPORTED",
    "SORT_STATUS_CONTEXT_MISMATCH", "SORT_STATUS_INVALID_CONTEXT", "SORT_STATUS_KERNEL_NOT_FOUND",
    "SORT_STATUS_KERNEL_ERROR", "SORT_STATUS_LAUNCH_ERROR", "SORT_STATUS_INVALID_VALUE",
    "SORT_STATUS_DEVICE_ERROR", "SORT_STATUS_INTERNAL_ERROR", "SORT_STATUS_UNSUPPORTED_DEVICE",
]
# Exports of sort/src/mgpusort/mgpusort.def:1-19.
REFERENCE_EXPORTS = [
    "sortStatusString", "sortCreateEngine", "sortLoadKernel", "sortIncEngine",
    "sortReleaseEngine", "sortCreateData", "sortDestroyData", "sortAllocData", "sortFreeData",
    "sortArray", "sortArrayEx", "sortHost", "sortDevice",
]


def declared_functions(header):
    text = open(header).read()
    return sorted(set(re.findall(r"SORTAPI\s+(\w+)\s*\(", text)))


def exported_symbols(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_function():
    declared = declared_functions(HEADER)
    assert len(declared) >= len(REFERENCE_EXPORTS)
    exported = exported_symbols(mg.library_path())
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    for name in REFERENCE_EXPORTS + ["sortDestroyEngine"]:
        assert name in declared
        assert hasattr(mg.lib(), name)
    # Nothing but the C API leaks out of the library.
    assert exported == set(declared)


def test_sortdata_layout_matches_reference_lp64():
    S = mg.SortData
    assert ctypes.sizeof(S) == 144
    assert S.maxElements.offset == 0 and S.endBit.offset == 16
    assert S.preserveEndKeys.offset == 20 and S.earlyExit.offset == 21
    assert S.parity.offset == 24 and S.keys.offset == 32 and S.values.offset == 48


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles_and_layout_in_c(tmp_path, lang):
    src = tmp_path / ("abi." + ("c" if lang == "c" else "cpp"))
    src.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "mgpusort.h"
int main(void) {
    sortData_d d;
    printf("%zu %zu %zu %zu %zu %zu %zu %d\n", sizeof(sortData_d),
           offsetof(sortData_d, preserveEndKeys), offsetof(sortData_d, parity),
           offsetof(sortData_d, keys), offsetof(sortData_d, values1),
           offsetof(sortData_d, values6), offsetof(sortData_d, pValues), SORT_VALUE_INDEX);
    d.values[2][1] = 7;
    return d.pValues3[1] == (unsigned int*)7 && sizeof(CUdeviceptr) == 8 ? 0 : 1;
}
''')
    exe = tmp_path / "abi"
    compiler = "gcc" if lang == "c" else "g++"
    std = "-std=c11" if lang == "c" else "-std=c++17"
    subprocess.run([compiler, std, "-Wall", "-Werror", "-Wno-pedantic", "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["144", "20", "24", "32", "48", "128", "48", "-1"]


def test_cpp_wrapper_compiles_links_and_throws_without_device(tmp_path):
    """inc/mgpusort.hpp's classes (MgpuSort, MgpuSortData, MgpuException)
    build against the library; with no device the engine constructor throws."""
    src = tmp_path / "wrap.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cstring>
#include "mgpusort.hpp"
int main() {
    MgpuSortData d;
    unsigned int* fake = reinterpret_cast<unsigned int*>(0x1000);
    d.AttachKey(fake);
    d.AttachVal(0, fake + 64);
    bool ok = d.pKeys[0] == fake && d.values1[0] == (CUdeviceptr)(fake + 64);
    d.Reset();  // attached buffers are not freed
    ok = ok && d.keys[0] == 0 && d.values1[0] == 0;
    try {
        MgpuSort engine(static_cast<const char*>(nullptr));
        std::printf("engine %d\n", ok ? 1 : 0);
    } catch (const MgpuException& e) {
        std::printf("%s %d\n", e.what(), ok ? 1 : 0);
    }
    MgpuSort none(static_cast<sortEngine_t>(nullptr));
    MgpuSort copy(none);
    copy = none;
    return 0;
}
''')
    exe = tmp_path / "wrap"
    lib_dir = mg.LIB_DIR
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", lib_dir, f"-Wl,-rpath,{lib_dir}", "-lmgpusort"],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out[-1] == "1"
    if not _gpu_visible():
        assert out[0] == "SORT_STATUS_INVALID_CONTEXT"


def test_status_strings_match_reference():
    for code, name in enumerate(REFERENCE_STATUS):
        assert mg.status_string(code) == name
    assert mg.lib().sortStatusString(len(REFERENCE_STATUS)) is None
    assert mg.lib().sortStatusString(-1) is None


@pytest.mark.parametrize("bits,widths", [
    (32, [8, 8, 8, 8]), (24, [8, 8, 8]), (16, [8, 8]), (8, [8]), (1, [1]), (0, []),
    (18, [6, 6, 6]), (25, [6, 6, 6, 7]), (9, [4, 5]), (33, []),
])
def test_plan_passes(bits, widths):
    assert mg.plan_passes(bits) == widths
    assert sum(widths) == (bits if bits <= 32 else 0)


def test_null_arguments_rejected_without_device():
    L = mg.lib()
    d = mg.SortData()
    assert L.sortArray(None, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE
    assert L.sortArrayEx(None, ctypes.byref(d), 256, 8, 6, False) == mg.SortStatus.INVALID_VALUE
    for key_type in (mg.KEYS_UINT32, mg.KEYS_INT32, mg.KEYS_FLOAT32, 3, -1):
        assert L.sortArrayTyped(None, ctypes.byref(d), key_type) == mg.SortStatus.INVALID_VALUE
    assert L.sortIncEngine(None) == mg.SortStatus.INVALID_VALUE
    assert L.sortReleaseEngine(None) == mg.SortStatus.SUCCESS
    assert L.sortDestroyEngine(None) == mg.SortStatus.SUCCESS
    assert L.sortLoadKernel(None, 256, 8, False, 1) == mg.SortStatus.INVALID_VALUE
    assert L.sortFreeData(None) == mg.SortStatus.INVALID_VALUE
    assert L.sortDestroyData(None) == mg.SortStatus.SUCCESS
    assert L.sortSetStream(None, None) == mg.SortStatus.INVALID_VALUE
    out = ctypes.POINTER(mg.SortData)()
    assert L.sortCreateData(None, 100, 7, ctypes.byref(out)) == mg.SortStatus.INVALID_VALUE
    assert L.sortCreateData(None, 100, -2, ctypes.byref(out)) == mg.SortStatus.INVALID_VALUE
    d.valueCount = 7
    assert L.sortAllocData(None, ctypes.byref(d)) == mg.SortStatus.INVALID_VALUE


def test_free_data_of_empty_descriptor_is_success():
    d = mg.SortData()
    d.maxElements = 5
    assert mg.lib().sortFreeData(ctypes.byref(d)) == mg.SortStatus.SUCCESS
    assert d.maxElements == 0


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-device error path")
def test_create_engine_without_device_reports_invalid_context():
    h = ctypes.c_void_p()
    st = mg.lib().sortCreateEngine(mg.kernel_path().encode(), ctypes.byref(h))
    assert st == mg.SortStatus.INVALID_CONTEXT
    assert not h.value
    with pytest.raises(mg.MgpuError) as e:
        mg.Engine()
    assert e.value.status == mg.SortStatus.INVALID_CONTEXT


def test_code_objects_built_for_gfx950():
    for name in ["mgpusort_radix.hsaco"]:
        path = os.path.join(mg.LIB_DIR, name)
        assert os.path.exists(path), path
        data = open(path, "rb").read()
        assert data[:4] == b"\x7fELF"
        assert b"gfx950" in data
        assert b"mgpu_digit_count" in data and b"mgpu_rank_scatter_b1024_i16_r1_v0" in data


@pytest.mark.parametrize("end_bit", [32, 24])
def test_segmented_index_mode_needs_value_side_zero_without_device(end_bit):
    """Index generation with a null values[0][0] is rejected by argument
    validation, before the engine is used (the handle here is a zeroed block
    standing in for an engine that the validation never reads)."""
    L = mg.lib()
    fake = ctypes.create_string_buffer(4096)
    eng = ctypes.cast(fake, ctypes.c_void_p)
    d = mg.SortData()
    d.maxElements = d.numElements = 4096
    d.valueCount, d.firstBit, d.endBit = -1, 0, end_bit
    d.keys[0], d.keys[1] = 0x10000, 0x20000
    d.values[0][1] = 0x30000
    starts = (ctypes.c_int * 3)(0, 1000, 4096)
    assert L.sortSegmentedArray(eng, ctypes.byref(d), 2, starts) == mg.SortStatus.INVALID_VALUE
    plan = ctypes.c_void_p()
    assert L.sortSegmentedPlanCreate(eng, ctypes.byref(d), 2, starts, ctypes.byref(plan)) == \
        mg.SortStatus.INVALID_VALUE
    seg = mg.SortSegment()
    seg.keys[0], seg.keys[1] = 0x10000, 0x20000
    seg.values[0][1] = 0x30000
    seg.numElements = 4096
    assert L.sortSegmentedBatch(eng, ctypes.byref(seg), 1, -1, 0, end_bit) == \
        mg.SortStatus.INVALID_VALUE
