"""The C++ drop-in client (tests/cpp/client.cpp, built by build()) run
against libmgpusort on the GPU: reference-style usage through mgpusort.hpp,
checked against std::stable_sort inside the client."""
import os
import subprocess

import pytest

import mgpu_old_amd as mg
from conftest import ROOT

pytestmark = pytest.mark.gpu

CLIENT = os.path.join(ROOT, "tests", "cpp", "client")


def test_cpp_client_runs_bit_exact():
    assert os.path.exists(CLIENT), "tests/cpp/client not built: run __graft_entry__.build()"
    out = subprocess.run([CLIENT, mg.kernel_path()], capture_output=True, text=True, timeout=110)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "CLIENT OK" in out.stdout


CLIENT_ASAN = os.path.join(ROOT, "tests", "cpp", "client_asan")


def test_cpp_client_under_asan():
    """The same client linked with AddressSanitizer + UBSan against the host
    code built likewise: memory errors, undefined behaviour and leaks of the
    engine's host side during real sorts (the ROCm runtime's own allocations,
    kept until exit, are suppressed; it preloads nothing, and the link-order
    check is off because the process may start with other preloaded code)."""
    assert os.path.exists(CLIENT_ASAN), "tests/cpp/client_asan not built: run __graft_entry__.build()"
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               LSAN_OPTIONS="suppressions=" + os.path.join(ROOT, "tests", "cpp", "lsan_rocm.supp"),
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    out = subprocess.run([CLIENT_ASAN, mg.kernel_path()], capture_output=True, text=True,
                         timeout=110, env=env)
    assert out.returncode == 0, out.stdout + out.stderr[-4000:]
    assert "CLIENT OK" in out.stdout
