"""The oracle's basis search, push apply and sparse triplets, pinned by the
reference's compiled search (oracle/_ref), a brute-force dense H built with
a hash map instead of a search, the fixture sparsity counts, Hermiticity and
the row-sum identity."""
import numpy as np
import pytest

from conftest import load_basis
from oracle import pyoracle as orc
import qsi_subspace_ed_amd as q


def _as_int(lohi):
    return [int(a) | (int(b) << 64) for a, b in lohi]


def _lohi(states):
    return np.array([[s & ((1 << 64) - 1), s >> 64] for s in states], dtype=np.uint64)


@pytest.mark.parametrize("W", [8, 32, 48])
def test_search_oracle_vs_reference(W):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    b = load_basis(W)
    ob, rb = orc.Basis(b), orc.Basis(b, ref=True)
    rng = np.random.default_rng(W)
    queries = np.concatenate([b, b ^ rng.integers(0, 2, size=b.shape).astype(np.uint64)])
    queries[:, 1] = 0
    for bst in (0, 1):
        a = ob.search(queries, bst=bst)
        r = rb.search(queries, bst=bst)
        assert np.array_equal(a, r)
    assert np.array_equal(ob.search(b), np.arange(b.shape[0]))


def _dense_push(states, H):
    """H as a dense matrix by the push rule with a dict lookup."""
    idx = {s: i for i, s in enumerate(states)}
    D = len(states)
    M = np.zeros((D, D))
    for c, op, _ in H.off_diag_terms:
        for j, s in enumerate(states):
            sg, s2 = op.apply_state(s)
            if sg and s2 in idx:
                M[idx[s2], j] += c * sg
    for c, op, _ in H.diagonal_terms:
        for j, s in enumerate(states):
            sg, _ = op.apply_state(s)
            M[j, j] += c * sg
    return M


def _ham32(jd, field=True):
    H = q.build_hamiltonian_g(jd, [-1.0, -0.55, -0.3, -0.8])
    if field:
        q.add_local_field(H, jd, [0.07, -0.02, 0.05])
    return H


@pytest.mark.parametrize("W", [8, 32])
def test_apply_vs_dense(W, lattices):
    jd = lattices[W]
    b = load_basis(W)
    H = _ham32(jd)
    M = _dense_push(_as_int(b), H)
    assert np.allclose(M, M.T, atol=0)  # ring terms come in O, O^dagger pairs
    rng = np.random.default_rng(7)
    x = rng.standard_normal(b.shape[0])
    y0 = rng.standard_normal(b.shape[0])
    ob = orc.Basis(b)
    y = orc.evaluate_add(ob, H, x, y0.copy())
    ref = y0 + M @ x
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-14


def test_apply_tol_skip(lattices):
    """|x_j| <= 1e-15 contributes nothing off the diagonal (APPLY_TOL)."""
    jd = lattices[32]
    b = load_basis(32)
    H = q.build_hamiltonian_g(jd, [1.0, 1.0, 1.0, 1.0])
    x = np.full(b.shape[0], 1e-16)
    y = orc.evaluate_add(orc.Basis(b), H, x, np.zeros(b.shape[0]))
    assert np.all(y == 0.0)


def test_jpm_hamiltonian_shape(lattices):
    jd = lattices[48]
    H = q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    assert len(H.off_diag_terms) == 96
    assert len(H.diagonal_terms) == len(jd["bonds"]) == 144


@pytest.mark.parametrize("W,nnz,D", [(32, 15360, 2970), (48, 703680, 87546)])
def test_triplet_counts(W, nnz, D, lattices):
    jd = lattices[W]
    b = load_basis(W)
    assert b.shape[0] == D
    H = q.build_hamiltonian_g(jd, [1.0, 1.0, 1.0, 1.0])
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    off = r != c
    assert off.sum() == nnz
    assert (r == c).sum() == D
    if W == 32:
        hist = np.bincount(np.bincount(c[off], minlength=D), minlength=9)
        assert hist.tolist() == [474, 0, 0, 0, 672, 0, 768, 384, 672]


def test_row_sum_identity(lattices):
    """(H 1)_i = sum over flippable hexagons of g_sl + diagonal(i)."""
    jd = lattices[32]
    b = load_basis(32)
    g = [-1.0, -0.55, -0.3, -0.8]
    H = q.build_hamiltonian_g(jd, g)
    _, hB = q.add_local_field(H, jd, [0.07, -0.02, 0.05])
    y = orc.evaluate_add(orc.Basis(b), H, np.ones(b.shape[0]), np.zeros(b.shape[0]))
    sl = q.lattice.sublattice_of(jd)
    states = _as_int(b)
    ops, _, sls = q.get_ring_ops(jd)
    for i in range(0, len(states), 37):
        s = states[i]
        exp = 0.0
        for o, slr in zip(ops, sls):
            m = o.X
            pat = s & m
            if pat == o.up or pat == o.down:  # one of the two alternating patterns
                exp += g[slr]
        for site in range(32):
            exp += -0.5 * hB[sl[site]] * (1 if (s >> site) & 1 else -1)
        assert abs(y[i] - exp) < 1e-13


def test_hermitian_jpm48(lattices):
    jd = lattices[48]
    b = load_basis(48)
    H = q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    ob = orc.Basis(b)
    rng = np.random.default_rng(11)
    u, v = rng.standard_normal((2, b.shape[0]))
    Hu = orc.evaluate_add(ob, H, u, np.zeros(b.shape[0]))
    Hv = orc.evaluate_add(ob, H, v, np.zeros(b.shape[0]))
    assert abs(u @ Hv - v @ Hu) < 1e-12 * np.linalg.norm(Hu) * np.linalg.norm(v)


def test_golden_e0_sparse(lattices):
    """The 48-spin Jpm=0.02, B=(0.1,0.1,0.1) Hamiltonian assembled from the
    oracle's triplets has the golden ground energy of test/test_correct.sh."""
    import os
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from conftest import GOLDEN
    gold = np.load(os.path.join(GOLDEN, "eigs48_jpm0.02_B0.1.npz"))
    jd = lattices[48]
    b = load_basis(48)
    H = q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    M = sp.csr_matrix((v, (r, c)), shape=(b.shape[0],) * 2)
    w = sla.eigsh(M, k=1, which="SA", tol=1e-14, v0=np.ones(b.shape[0]))[0]
    e0 = gold["eigenvalues"][0]
    assert abs(w[0] - e0) < 1e-9 * abs(e0)
    # the golden eigenvector is an eigenvector of the same H
    vec = gold["eigenvector"]
    res = M @ vec - e0 * vec
    assert np.linalg.norm(res) < 1e-6 * abs(e0)


def test_row_range_apply_sums_to_full(lattices):
    """Pushes of row ranges add up to the full apply (the sampled CPU
    baseline of very large bases uses a range)."""
    from conftest import load_basis
    b = load_basis(32)
    D = b.shape[0]
    H = q.build_hamiltonian_g(lattices[32], [-0.9, -0.35, -0.6, -0.45])
    q.add_local_field(H, lattices[32], [0.05, -0.02, 0.04])
    ob = orc.Basis(b)
    x = np.random.default_rng(1).standard_normal(D)
    full = orc.evaluate_add(ob, H, x, np.zeros(D))
    parts = np.zeros(D)
    for r0 in range(0, D, 1000):
        orc.evaluate_add_rows(ob, H, x, parts, r0, r0 + 1000)
    np.testing.assert_allclose(parts, full, rtol=1e-13, atol=1e-13)
