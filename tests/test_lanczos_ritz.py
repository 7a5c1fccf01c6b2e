"""Cullum-Willoughby selection of Ritz values (qsi_lanczos_ritz), host only.

Pinned by what the mathematics fixes: a long run of the 3-term recurrence
without reorthogonalisation on a matrix with a known spectrum produces ghost
copies of its converged extreme eigenvalues and spurious values in between;
the selection must report every converged eigenvalue once (to ~1e-12), never
report a spurious value as converged (small residual), and give exact values
with zero residual when the recurrence ends on an invariant subspace."""
import numpy as np
import pytest

from qsi_subspace_ed_amd import lanczos_ritz


def three_term(diag, v0, m):
    """The reference's recurrence (lanczos.cpp:111-146) on diag(d): alphas
    and betas, beta[j] = norm after step j, no reorthogonalisation."""
    v = v0 / np.linalg.norm(v0)
    u = np.zeros_like(v)
    a = np.zeros(m)
    b = np.zeros(m)
    for j in range(m):
        w = diag * v - (b[j - 1] * u if j else 0.0)
        a[j] = v @ w
        w -= a[j] * v
        b[j] = np.linalg.norm(w)
        u, v = v, w / b[j]
    return a, b


@pytest.fixture(scope="module")
def run():
    rng = np.random.default_rng(5)
    # a few well separated low eigenvalues below a dense bulk
    low = np.array([-13.0, -12.2, -11.9, -11.5, -11.0])
    bulk = rng.uniform(-10.0, 10.0, 3000)
    d = np.concatenate([low, bulk])
    a, b = three_term(d, rng.standard_normal(d.size), 400)
    return d, a, b


def test_plain_ritz_values_carry_ghosts(run):
    # the premise: the plain tridiagonal spectrum holds repeated copies
    d, a, b = run
    th = np.linalg.eigvalsh(np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1))
    assert np.sum(np.abs(th - (-13.0)) < 1e-9) >= 2


def test_converged_values_reported_once(run):
    d, a, b = run
    vals, res, mult = lanczos_ritz(a, b)
    assert np.all(np.diff(vals) > 0)
    low = np.sort(d)[:5]
    np.testing.assert_allclose(vals[:5], low, rtol=0, atol=1e-11)
    assert np.all(res[:5] < 1e-8)
    assert mult[0] >= 2  # the ghost copies of the lowest are merged


def test_no_spurious_value_looks_converged(run):
    d, a, b = run
    vals, res, _ = lanczos_ritz(a, b)
    spec = np.sort(d)
    conv = vals[res < 1e-8]
    dist = np.min(np.abs(conv[:, None] - spec[None, :]), axis=1)
    assert np.all(dist < 1e-9), dist.max()
    # and the filter removes values the plain spectrum holds
    th = np.linalg.eigvalsh(np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1))
    assert vals.size < th.size


@pytest.mark.parametrize("seed", [5, 6, 7, 8, 10, 11])
def test_ghosts_on_their_way_are_dropped(seed):
    # mid-run, T holds copies still converging towards an eigenvalue (such
    # as -12.19998 beside -12.2): plain Ritz values below the bulk that are
    # no eigenvalue exist, the selected ones below the bulk are all exact
    rng = np.random.default_rng(seed)
    low = np.array([-13.0, -12.2, -11.9, -11.5, -11.0])
    d = np.concatenate([low, rng.uniform(-10.0, 10.0, 3000)])
    a, b = three_term(d, rng.standard_normal(d.size), 250)
    th = np.linalg.eigvalsh(np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1))
    stray = [x for x in th if x < -10.0 and np.min(np.abs(low - x)) > 1e-6]
    assert stray
    vals, res, _ = lanczos_ritz(a, b)
    below = vals[vals < -10.0]
    np.testing.assert_allclose(below, low, atol=1e-9)


def test_invariant_subspace_exact():
    d = np.array([-2.0, -1.0, 0.5, 3.0])
    a, b = three_term(d, np.ones(4), 4)
    b[-1] = 0.0  # the recurrence ended: T_4 is exact
    vals, res, mult = lanczos_ritz(a, b)
    np.testing.assert_allclose(vals, d, atol=1e-13)
    assert np.all(res == 0.0) and np.all(mult == 1)


def test_max_out_and_validation():
    d = np.linspace(-1, 1, 50)
    a, b = three_term(d, np.ones(50), 20)
    vals, res, mult = lanczos_ritz(a, b, max_out=3)
    assert vals.size == 3
    with pytest.raises(ValueError):
        lanczos_ritz(a, b[:-1])
