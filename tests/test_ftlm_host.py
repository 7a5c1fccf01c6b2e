"""FTLM quadrature on the host (ftlm.quadrature / thermodynamics) for samples
whose recurrence ends early (M > D): the zeroed tail must not add nodes, so
the shift e_ref and the Boltzmann sums stay those of the exact spectrum."""
import numpy as np

from qsi_subspace_ed_amd import ftlm


def _coeffs(A, v0, M):
    """The 3-term recurrence on a small dense A, stopped (later coefficients
    0) when beta vanishes, as the device recurrence reports it."""
    v = v0 / np.linalg.norm(v0)
    u = np.zeros_like(v)
    a = np.zeros(M)
    b = np.zeros(M)
    for j in range(M):
        w = A @ v - (b[j - 1] * u if j else 0.0)
        a[j] = v @ w
        w -= a[j] * v
        w -= (v @ w) * v  # keep the tiny case exact enough to end cleanly
        nb = np.linalg.norm(w)
        if nb < 1e-12:
            break
        b[j] = nb
        u, v = v, w / nb
    return a, b


def test_truncated_samples_positive_spectrum():
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.standard_normal((4, 4)))
    w_exact = np.array([0.5, 1.0, 2.0, 3.5])  # positive: exp(-e/T) underflows at low T
    A = Q @ np.diag(w_exact) @ Q.T
    M = 10
    al = np.zeros((3, M))
    be = np.zeros((3, M))
    for r in range(3):
        al[r], be[r] = _coeffs(A, rng.standard_normal(4), M)
        assert be[r][3] == 0.0  # the recurrence ended at step 4 < M
    nodes = ftlm.quadrature(al, be)
    for e, w in nodes:
        assert e.size == 4
        np.testing.assert_allclose(np.sort(e), w_exact, atol=1e-10)
        assert abs(w.sum() - 1.0) < 1e-12
    th = ftlm.thermodynamics(nodes, 4, [0.001, 0.01, 0.1, 1.0])
    assert abs(th["E_ref"] - 0.5) < 1e-10
    assert np.all(np.isfinite(th["E"])) and np.all(np.isfinite(th["C"]))
    assert abs(th["E"][0] - 0.5) < 1e-9
