"""Host-side sanitizer runs (CPU; no GPU needed).

* The CPU oracle and a driver over it (tests/cpp/oracle_check.cpp) built with
  AddressSanitizer and UndefinedBehaviorSanitizer into one executable: the
  checker's own memory and arithmetic errors, on empty and ragged inputs too.
* libmgpusort.so's host code built with UndefinedBehaviorSanitizer
  (mgpu_old_amd.build.build_ubsan; a shared library that needs no preloaded
  runtime) and loaded through MGPUSORT_LIBRARY by a process that runs the
  C-ABI tests: argument validation, status strings, pass plans and the
  device-less paths.  The same build, loaded the same way, runs the GPU tests
  on a GPU.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-pthread",
                    os.path.join(ROOT, "oracle", "oracle.cpp"),
                    os.path.join(ROOT, "tests", "cpp", "oracle_check.cpp"), "-o", exe],
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cases clean" in r.stdout


def test_host_library_under_ubsan(tmp_path):
    from mgpu_old_amd import build
    noalloc = os.path.join(build.LIB, "noalloc.o")
    if not os.path.exists(noalloc):
        pytest.skip("the normal build has not run (no noalloc.o)")
    lib = build.build_ubsan(str(tmp_path))
    env = dict(os.environ, MGPUSORT_LIBRARY=lib,
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # The sanitized build is the one a process loads...
    probe = ("import mgpu_old_amd as mg; mg.lib(); "
             "maps = open('/proc/self/maps').read(); "
             f"assert {lib!r} in maps and 'libubsan' in maps, 'sanitized library not loaded'")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # ...and the C-ABI tests run clean under it.
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p",
                        "no:cacheprovider", os.path.join(ROOT, "tests", "test_abi.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "runtime error" not in r.stdout + r.stderr
