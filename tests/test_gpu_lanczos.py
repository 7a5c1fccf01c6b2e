"""GPU Krylov solvers against the oracle, dense diagonalisation and the
reference's golden mfeig0 result (test/test_correct.sh).

Tolerances: eigenvalues within 1e-9 relative; the mfeig0 eigenvector within
1e-4 per entry up to an overall sign (the golden file's h5diff tolerance; the
Ritz vector's sign is arbitrary in any tridiagonal eigensolver)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_basis
from oracle import pyoracle as orc
import qsi_subspace_ed_amd as q

pytestmark = pytest.mark.gpu

EIG_TOL = 1e-9


def _dense(b, H):
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    D = b.shape[0]
    M = np.zeros((D, D))
    np.add.at(M, (r, c), v)
    return M


def test_eigval0_matches_oracle_32(lattices, cuda):
    jd = lattices[32]
    b = load_basis(32)
    H = q.build_hamiltonian_g(jd, [-1.0, -0.55, -0.3, -0.8])
    q.add_local_field(H, jd, [0.07, -0.02, 0.05])
    op = q.GpuOp(q.GpuBasis(b, 32), H)
    st = q.Settings(krylov_dim=15, abs_tol=1e-10, rel_tol=1e-10, min_iterations=10,
                    max_iterations=300, x0_seed=0, calc_eigenvector=True)
    e_gpu, v_gpu, res = op.lanczos_eigval0(st)
    ost = orc.settings(15, 1e-10, 1e-10, 10, 300, 0, True)
    e_orc, v_orc, ores, _, _ = orc.lanczos_eigval0(orc.Basis(b), H, ost)
    w = np.linalg.eigvalsh(_dense(b, H))
    assert abs(e_gpu - w[0]) <= EIG_TOL * abs(w[0])
    assert abs(e_gpu - e_orc) <= EIG_TOL * abs(w[0])
    assert res.eigvec_converged and ores.eigvec_converged
    assert res.n_iterations == ores.n_iterations
    vg = v_gpu.cpu().numpy()
    s = np.sign(vg @ v_orc)
    assert np.max(np.abs(vg - s * v_orc)) < 1e-6


def test_eigval0_golden_48(lattices, cuda):
    """diag_DOQSI_ham --Jpm 0.02 --B 0.1 0.1 0.1 -n 1 -N 1 --algorithm mfeig0
    -k 20 --rng_seed 11 --atol -9 --rtol -9 (test/test_correct.sh:8)."""
    gold = np.load(os.path.join(GOLDEN, "eigs48_jpm0.02_B0.1.npz"))
    jd = lattices[48]
    b = load_basis(48)
    H = q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    op = q.GpuOp(q.GpuBasis(b, 48), H)
    # parse_lanczos_settings (include/lanczos_cli.hpp:30-39)
    st = q.Settings(krylov_dim=20, abs_tol=1e-9, rel_tol=1e-9, min_iterations=10,
                    max_iterations=5000, x0_seed=11, calc_eigenvector=True,
                    verify_eigenvector=True)
    e0, vec, res = op.lanczos_eigval0(st)
    g0 = gold["eigenvalues"][0]
    assert abs(e0 - g0) <= EIG_TOL * abs(g0), (e0, g0)
    v = vec.cpu().numpy()
    gv = gold["eigenvector"]
    s = np.sign(v @ gv)
    assert np.max(np.abs(v - s * gv)) < 1e-4
    assert res.eigvec_error < 1e-6


def _check_against_dense(evals, w, nev):
    """Each Krylov eigenvalue is an eigenvalue of H, and the list covers the
    lowest eigenvalues (a single start vector sees one copy of each
    degenerate eigenspace, so compare against distinct values)."""
    scale = max(1.0, np.max(np.abs(w)))
    for e in evals:
        assert np.min(np.abs(w - e)) <= EIG_TOL * scale, e
    distinct = []
    for x in w:
        if not distinct or x - distinct[-1] > 1e-7 * scale:
            distinct.append(x)
    got = sorted(set(np.round(evals / scale, 8)))
    want = np.round(np.array(distinct[:len(got)]) / scale, 8)
    assert np.allclose(got, want, atol=2e-8), (evals, distinct[:nev])
    assert abs(evals[0] - w[0]) <= EIG_TOL * scale


@pytest.mark.parametrize("g", [[1.0, 1.0, 1.0, 1.0], [0.9, 0.35, 0.6, 0.45]])
def test_eigsh_32_vs_dense(g, lattices, cuda):
    """Config 1 shape: 32 spins, H = -sum g_s (O + O^dagger), nev=5, ncv=15,
    tol 1e-10."""
    jd = lattices[32]
    b = load_basis(32)
    H = q.build_hamiltonian_g(jd, [-x for x in g])
    op = q.GpuOp(q.GpuBasis(b, 32), H)
    evals, vecs, nconv, napp = op.eigsh(nev=5, ncv=15, tol=1e-10, seed=0, n_evecs=2)
    assert nconv == 5
    w = np.linalg.eigvalsh(_dense(b, H))
    _check_against_dense(evals, w, 5)
    # Ritz vectors are unit eigenvectors
    M = _dense(b, H)
    for k in range(2):
        v = vecs[:, k].cpu().numpy()
        assert abs(np.linalg.norm(v) - 1) < 1e-10
        assert np.linalg.norm(M @ v - evals[k] * v) < 1e-8


def test_eigsh_48_vs_sparse(lattices, cuda):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    jd = lattices[48]
    b = load_basis(48)
    H = q.build_hamiltonian_g(jd, [-0.9, -0.25, -0.6, -0.45])
    q.add_local_field(H, jd, [0.05, 0.02, -0.03])
    op = q.GpuOp(q.GpuBasis(b, 48), H)
    evals, _, nconv, _ = op.eigsh(nev=5, ncv=15, tol=1e-10, seed=1)
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    M = sp.csr_matrix((v, (r, c)), shape=(b.shape[0],) * 2)
    w = np.sort(sla.eigsh(M, k=12, which="SA", tol=1e-13)[0])
    _check_against_dense(evals, w, 5)
