"""The oracle's Lanczos restatement (lanczos.cpp) and its small eigensolvers,
pinned by numpy/LAPACK and by the reference's own mock-Hamiltonian scenario
(test/test_lanczos.cpp: 100x100 symmetric matrix from std::mt19937(0) normal
draws, eigenvalue and eigenvector error <= 1e-4)."""
import numpy as np
import pytest

from oracle import pyoracle as orc


def _mock_matrix(dim, seed):
    """generate_random_Herm (test/test_lanczos.cpp:44-60): upper triangle
    row by row from one normal stream, mirrored."""
    draws = orc.normal_stream(seed, dim * (dim + 1) // 2)
    M = np.zeros((dim, dim))
    k = 0
    for i in range(dim):
        for j in range(i, dim):
            M[i, j] = M[j, i] = draws[k]
            k += 1
    return M


def test_mock_lanczos_scenario():
    M = _mock_matrix(100, 0)
    # settings of test_lanczos.cpp:104-130 (defaults: krylov 30, tol 1e-8,
    # min 30, max 5000); the reference seeds the start vector from
    # Settings::x0_seed (default 10)
    st = orc.settings(krylov_dim=30, abs_tol=1e-8, rel_tol=1e-8, min_iterations=30,
                      max_iterations=5000, x0_seed=10, calc_eigenvector=True)
    e, v, res = orc.lanczos_dense(M, st)
    w, V = np.linalg.eigh(M)
    assert res.eigval_converged
    assert abs(e - w[0]) < 1e-10
    err_vec = 1 - abs(V[:, 0] @ v / np.linalg.norm(v))
    assert err_vec < 1e-4


def test_random_unit_is_normalised_gaussian():
    v = orc.random_unit(11, 1000)
    raw = orc.normal_stream(11, 1000)
    assert abs(np.linalg.norm(v) - 1) < 1e-15
    assert np.allclose(v, raw / np.linalg.norm(raw), rtol=1e-15, atol=0)


@pytest.mark.parametrize("n", [1, 2, 7, 40])
def test_jacobi_vs_numpy(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, V = orc.jacobi_eig(A)
    wn = np.linalg.eigvalsh(A)
    assert np.allclose(w, wn, atol=1e-12)
    assert np.allclose(A @ V, V * w, atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 5, 60, 400])
def test_tridiag_lowest_pair_vs_numpy(n):
    rng = np.random.default_rng(100 + n)
    d, e = rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, V = np.linalg.eigh(T)
    e0, v = orc.tridiag_one(d, e)
    assert abs(e0 - w[0]) < 1e-13 * max(1, abs(w[0]))
    assert 1 - abs(v @ V[:, 0]) < 1e-12
