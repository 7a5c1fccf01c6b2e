"""The allocation-free entry point qsi_lanczos_coeffs_noalloc, called through
the C ABI, against qsi_lanczos_coeffs (the entry point it mirrors) on the same
operator and start vector: the same coefficients, bit for bit, on two calls
in a row, the second on a smaller basis (the static scratch keeps the first
call's contents).  Sizes beyond the scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_basis
import qsi_subspace_ed_amd as q
from qsi_subspace_ed_amd import _lib
from qsi_subspace_ed_amd.device import randn

pytestmark = pytest.mark.gpu


def _noalloc(op, v, u, M):
    a = np.zeros(M)
    b = np.zeros(M)
    n = C.c_int(0)
    P = C.POINTER(C.c_double)
    rc = _lib.load().qsi_lanczos_coeffs_noalloc(op.handle, v.data_ptr(), u.data_ptr(), M,
                                                 a.ctypes.data_as(P), b.ctypes.data_as(P),
                                                 C.byref(n))
    return rc, a[: n.value], b[: n.value]


def test_noalloc_matches_lanczos_coeffs(lattices, cuda):
    torch = cuda
    M = 150
    for W, g in ((48, [-0.8, -0.35, -0.6, -0.5]), (32, [-1.0, -0.7, -0.4, -0.9])):
        b = load_basis(W)
        H = q.build_hamiltonian_g(lattices[W], g)
        q.add_local_field(H, lattices[W], [0.04, -0.03, 0.02])
        op = q.GpuOp(q.GpuBasis(b, W), H)
        v0 = randn(b.shape[0], 100 + W)
        a_ref, b_ref = op.lanczos_coeffs(M, v0=v0)
        v = v0.clone()
        u = torch.empty_like(v)
        torch.cuda.synchronize()
        rc, a, bb = _noalloc(op, v, u, M)
        assert rc == _lib.QSI_OK, _lib.load().qsi_last_error()
        assert a.shape == a_ref.shape
        assert np.array_equal(a, a_ref) and np.array_equal(bb, b_ref), W


def test_noalloc_refuses_beyond_scratch(lattices, cuda):
    torch = cuda
    b = load_basis(32)
    op = q.GpuOp(q.GpuBasis(b, 32), q.build_hamiltonian_g(lattices[32], [-1.0] * 4))
    v = randn(b.shape[0], 3)
    u = torch.empty_like(v)
    rc, _, _ = _noalloc(op, v, u, 4097)
    assert rc == _lib.QSI_E_INVALID
    assert b"4096" in _lib.load().qsi_last_error()
    rc, _, _ = _noalloc(op, v, v, 10)
    assert rc == _lib.QSI_E_INVALID
