"""The C ABI library loads and exports every symbol include/lineal_amd.h declares (no GPU needed)."""

import ctypes
import os
import re

import lineal_amd as la

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lineal_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    decl = declared_functions()
    assert len(decl) >= 30
    assert set(decl) == set(la.SIGNATURES), set(decl) ^ set(la.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(la.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_status_strings_without_gpu():
    L = la.lib()
    assert L.lgpu_status_string(la.OK) == b"ok"
    assert L.lgpu_status_string(la.SYNC_TIMEOUT).startswith(b"wavefront")


def test_invalid_arguments_are_reported_not_raised():
    L = la.lib()
    # a NULL context is rejected with a status and a message, no exception crosses
    assert L.lgpu_vec_create(None, 10, la.F64, ctypes.byref(ctypes.c_void_p())) == la.INVALID_ARG
    assert b"invalid" in L.lgpu_last_error()
    p = la._AmgParams()
    L.lgpu_amg_default_params(ctypes.byref(p))
    assert (p.max_level_num, p.min_coarsen_size, p.min_aggregate_size, p.max_aggregate_size) == (16, 32, 6, 9)


def test_cpp_header_compiles_against_the_c_abi(tmp_path):
    """The header-only C++ classes build with g++ alone (host code)."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    assert gxx, "g++ is part of the toolchain"
    src = tmp_path / "use.cpp"
    src.write_text('#include "lineal-amd/lineal-amd.hpp"\n'
                   "int main() { lineal::gpu::HierarchyWrightParams p; return int(p.max_level_num) - 16; }\n")
    res = subprocess.run([gxx, "-std=c++23", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          f"-I{os.path.join(ROOT, 'include')}", str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
