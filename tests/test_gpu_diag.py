"""diagonalise_real's algorithm dispatch (include/matrix_diag_bits.hpp:120-204)
on the device: every algorithm against the dense spectrum of the oracle's
triplets, the reference's default choice by size, and its errors."""
import numpy as np
import pytest

import qsi_subspace_ed_amd as q
from conftest import load_basis
from oracle import pyoracle as orc
from qsi_subspace_ed_amd import diag

pytestmark = pytest.mark.gpu


def _dense_w(b, H):
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    D = b.shape[0]
    M = np.zeros((D, D))
    np.add.at(M, (r, c), v)
    return np.linalg.eigvalsh(M), M


def _check(evals, vecs, w, M):
    import torch
    scale = max(1.0, np.max(np.abs(w)))
    for e in evals:
        assert np.min(np.abs(w - e)) <= 1e-9 * scale, e
    assert abs(evals[0] - w[0]) <= 1e-9 * scale
    V = vecs.cpu().numpy() if isinstance(vecs, torch.Tensor) else vecs
    for k in range(V.shape[1]):
        v = V[:, k] / np.linalg.norm(V[:, k])
        assert np.linalg.norm(M @ v - evals[k] * v) <= 1e-6 * scale


@pytest.mark.parametrize("W,algo", [(8, None), (8, "dense"), (32, None), (32, "dense"),
                                    (32, "sparse"), (32, "mfsparse"), (32, "mfeig0")])
def test_algorithms_against_dense(W, algo, lattices, cuda):
    jd = lattices[W]
    b = load_basis(W)
    H = q.build_hamiltonian_g(jd, [-1.0, -0.7, -0.4, -0.9])
    q.add_local_field(H, jd, [0.05, -0.03, 0.08])
    op = q.GpuOp(q.GpuBasis(b, W), H)
    w, M = _dense_w(b, H)
    if algo == "mfeig0":
        opts = diag.Options(algorithm="mfeig0", ncv=30, atol=-12, rtol=-12)
    else:
        opts = diag.Options(n_eigvals=5, n_eigvecs=3, ncv=15, atol=-10, algorithm=algo)
    print(f"[diag] W={W} algo={algo or diag.choose_algorithm(b.shape[0])}", flush=True)
    evals, vecs = diag.diagonalise_real(op, opts)
    assert len(evals) == (1 if algo == "mfeig0" else 5)
    assert vecs.shape[1] == (1 if algo == "mfeig0" else (5 if (algo or diag.choose_algorithm(
        b.shape[0])) == "dense" else 3))
    _check(evals, vecs, w, M)
    if algo == "mfsparse":
        # the matrix-free solve ran on the low-key kernel, left the operator's
        # kernel choice as it was, and the low keys stay on the basis
        assert op.basis.has_lowkeys()
        assert q._lib.load().qsi_op_get_variant(op.handle) == q._lib.QSI_APPLY_AUTO


def test_errors(lattices, cuda):
    jd = lattices[32]
    b = load_basis(32)
    op = q.GpuOp(q.GpuBasis(b, 32), q.build_hamiltonian_g(jd, [-1.0] * 4))
    with pytest.raises(RuntimeError, match="n_eigvals must be 1"):
        diag.diagonalise_real(op, diag.Options(algorithm="mfeig0", n_eigvals=2))
    with pytest.raises(RuntimeError, match="Unrecognised"):
        diag.diagonalise_real(op, diag.Options(algorithm="lapack"))
    with pytest.raises(RuntimeError, match="decomposition failed"):
        diag.diagonalise_real(op, diag.Options(algorithm="mfsparse", n_eigvals=5, atol=-15,
                                               max_iters=1))
