"""The header-only C++ classes on the GPU: tests/cpp/stencil_api.cpp runs the
reference's stencil test scenarios (known solution, the solver suite through
run_iterative_solver's iteration callbacks, Dirichlet residual, CSR copy,
smoothers, adjoint closed form, effective diffusivity of the compressed
system) through them, and reads, writes and solves the reference's UFZ_CT_02
files from tests/golden."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "stencil_api")


def test_cpp_api_program():
    assert os.path.exists(EXE), "build first: python __graft_entry__.py build (tests/cpp/Makefile)"
    print(f"running {EXE}", flush=True)
    golden = os.path.join(ROOT, "tests", "golden")
    res = subprocess.run([EXE, golden], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("ok  ") == 14
