"""The C-ABI library loads on a CPU-only host and exports every entry point
declared in include/qsi_gpu.h; argument validation fails cleanly before any
GPU work."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT
from qsi_subspace_ed_amd import _lib


def _declared():
    txt = open(os.path.join(ROOT, "include", "qsi_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qsi_[a-z0-9_]+)\s*\(", txt)))


def test_all_declared_symbols_exported():
    lib = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_version_and_validation():
    lib = _lib.load()
    assert b"gfx950" in lib.qsi_version()
    h = C.c_void_p()
    bad = np.array([[5, 0], [3, 0]], dtype=np.uint64)  # not increasing
    rc = lib.qsi_basis_create(bad.ctypes.data_as(C.c_void_p), 2, 8, 0, C.byref(h))
    assert rc == _lib.QSI_E_INVALID
    assert b"strictly increasing" in lib.qsi_last_error()
    big = np.array([[1 << 10, 0]], dtype=np.uint64)  # bit 10 with 8 spins
    rc = lib.qsi_basis_create(big.ctypes.data_as(C.c_void_p), 1, 8, 0, C.byref(h))
    assert rc == _lib.QSI_E_INVALID
    rc = lib.qsi_basis_create(None, 1, 0, 0, C.byref(h))
    assert rc == _lib.QSI_E_INVALID
    # ranks are 32-bit and one thread per row must fit a dispatch: more than
    # 2^32 - 512 states are refused before any read
    for dim in ((1 << 32) - 1, (1 << 32) - 511):
        rc = lib.qsi_basis_create(big.ctypes.data_as(C.c_void_p), dim, 8, 0, C.byref(h))
        assert rc == _lib.QSI_E_INVALID
        assert b"dim must be" in lib.qsi_last_error()
    # bucket keys of at most 32 bits (and at most n_spins), on both basis paths
    one = np.array([[1, 0]], dtype=np.uint64)
    for bits, spins in ((33, 64), (9, 8), (-1, 8)):
        rc = lib.qsi_basis_create(one.ctypes.data_as(C.c_void_p), 1, spins, bits, C.byref(h))
        assert rc == _lib.QSI_E_INVALID
        assert b"bucket_bits" in lib.qsi_last_error()
    tet = np.zeros(4, dtype=np.int32)
    order = np.arange(8, dtype=np.int32)
    n = C.c_int64(0)
    rc = lib.qsi_enumerate_basis_device(8, 1, tet.ctypes.data_as(C.c_void_p),
                                        order.ctypes.data_as(C.c_void_p), 0, 33, C.byref(h),
                                        C.byref(n))
    assert rc == _lib.QSI_E_INVALID
    assert b"bucket_bits" in lib.qsi_last_error()


def test_cpp_face_compiles():
    """include/qsi_gpu.hpp (the C++ face over the ABI) and its check program
    compile with the host compiler (no GPU needed)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        import pytest
        pytest.skip("no g++")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(rocm, "include"),
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "wrapper_check.cpp")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
