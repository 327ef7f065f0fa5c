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
