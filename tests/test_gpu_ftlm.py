"""Batched block-Lanczos / FTLM on the GPU.

Pins: every column's (alpha, beta) against the oracle's single-vector
recurrence from the same start vector (first 25 steps, 1e-10 relative —
later steps drift apart as orthogonality is lost in both); multi-sector and
graph launches against plain ones; and the quadrature pipeline against an
exact trace: with the complete set of basis vectors e_i as start vectors,
sum_i sum_k w_ik exp(-e_ik/T) = Tr exp(-H/T), checked against the dense
spectrum to 1e-10 relative."""
import numpy as np
import pytest

from conftest import load_basis
from oracle import pyoracle as orc
import qsi_subspace_ed_amd as q
from qsi_subspace_ed_amd import ftlm

pytestmark = pytest.mark.gpu


def _ham(jd, g, B=None):
    H = q.build_hamiltonian_g(jd, [-x for x in g])
    if B is not None:
        q.add_local_field(H, jd, B)
    return H


@pytest.mark.parametrize("R", [1, 5, 64])
def test_block_coefficients_match_oracle(R, lattices, cuda):
    torch = cuda
    jd = lattices[32]
    b = load_basis(32)
    H = _ham(jd, [0.9, 0.35, 0.6, 0.45], [0.05, -0.02, 0.04])
    op = q.GpuOp(q.GpuBasis(b, 32), H)
    V0 = ftlm.random_sign_block(b.shape[0], R, seed=R)
    M = 40
    al, be, _ = ftlm.lanczos_block([op], [torch.from_numpy(V0).cuda()], M)
    ob = orc.Basis(b)
    for c in range(min(R, 6)):
        oa, obb = orc.lanczos_coeffs(ob, H, V0[:, c], M)
        k = 25
        scale = np.max(np.abs(oa[:k]))
        assert np.max(np.abs(al[0, c, :k] - oa[:k])) <= 1e-10 * scale
        assert np.max(np.abs(be[0, c, :k] - obb[:k])) <= 1e-10 * scale


def test_multi_sector_and_graph(lattices, cuda):
    torch = cuda
    jd = lattices[48]
    H = _ham(jd, [0.7, 0.3, 0.9, 0.5])
    bases = [q.enumerate_basis(jd, n, n_threads=8) for n in (0, 1)]
    ops = [q.GpuOp(q.GpuBasis(bb, 48), H) for bb in bases]
    R, M = 16, 30
    V0s = [torch.from_numpy(ftlm.random_sign_block(bb.shape[0], R, seed=7 + k)).cuda()
           for k, bb in enumerate(bases)]
    a2, b2, _ = ftlm.lanczos_block(ops, V0s, M, use_graph=True)
    a2n, b2n, _ = ftlm.lanczos_block(ops, V0s, M, use_graph=False)
    assert np.array_equal(a2, a2n) and np.array_equal(b2, b2n)
    for s in range(2):
        a1, b1, _ = ftlm.lanczos_block([ops[s]], [V0s[s]], M)
        sc = np.max(np.abs(a1))
        assert np.max(np.abs(a1[0] - a2[s])) <= 1e-11 * sc
        assert np.max(np.abs(b1[0] - b2[s])) <= 1e-11 * sc


def test_exact_trace_with_complete_start_set(lattices, cuda):
    """FTLM quadrature with all D basis vectors as start vectors is exact."""
    torch = cuda
    jd = lattices[32]
    b = load_basis(32)
    D = b.shape[0]
    H = _ham(jd, [1.0, 0.8, 0.6, 0.9], [0.1, 0.0, -0.05])
    op = q.GpuOp(q.GpuBasis(b, 32), H)
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    Md = np.zeros((D, D))
    np.add.at(Md, (r, c), v)
    w = np.linalg.eigvalsh(Md)
    temps = np.array([0.5, 1.0, 2.0, 5.0])
    e_ref = w[0]
    Z_exact = np.array([np.sum(np.exp(-(w - e_ref) / T)) for T in temps])
    E_exact = np.array([np.sum(w * np.exp(-(w - e_ref) / T)) for T in temps]) / Z_exact
    M, R = 80, 64
    nodes = []
    for c0 in range(0, D, R):
        cols = np.arange(c0, min(c0 + R, D))
        V0 = np.zeros((D, R))
        V0[cols, np.arange(len(cols))] = 1.0
        if len(cols) < R:  # pad with duplicates, dropped below
            V0[:, len(cols):] = V0[:, :1]
        al, be, _ = ftlm.lanczos_block([op], [torch.from_numpy(V0).cuda()], M)
        for k in range(len(cols)):
            e, wt = ftlm.tridiag_eig(al[0, k], be[0, k])
            assert abs(wt.sum() - 1) < 1e-12
            nodes.append((e, wt))
    th = ftlm.thermodynamics(nodes, D, temps, e_ref=e_ref, normalise=False)
    assert np.allclose(th["Z_shifted"], Z_exact, rtol=1e-10)
    assert np.allclose(th["E"], E_exact, rtol=1e-10, atol=1e-12)


def test_cached_connectivity_same_coefficients(lattices, cuda):
    """Block Lanczos through cached connectivity: the same recurrence (the
    per-wave reduction grid differs, so 1e-12 relative, not bits)."""
    torch = cuda
    jd = lattices[48]
    H = _ham(jd, [0.7, 0.3, 0.9, 0.5], [0.02, 0.01, -0.03])
    bases = [q.enumerate_basis(jd, n, n_threads=8) for n in (0, 1)]
    ops = [q.GpuOp(q.GpuBasis(bb, 48), H) for bb in bases]
    R, M = 8, 30
    V0s = [torch.from_numpy(ftlm.random_sign_block(bb.shape[0], R, seed=3 + k)).cuda()
           for k, bb in enumerate(bases)]
    a0, b0, _ = ftlm.lanczos_block(ops, V0s, M)
    for op in ops:
        assert op.cache_connectivity(True) > 0
    a1, b1, _ = ftlm.lanczos_block(ops, V0s, M)
    sc = np.max(np.abs(a0))
    assert np.max(np.abs(a0 - a1)) <= 1e-12 * sc
    assert np.max(np.abs(b0 - b1)) <= 1e-12 * sc


@pytest.mark.parametrize("R", [6, 16])
def test_cached_column_pairs_same_coefficients(R, lattices, cuda):
    """Block Lanczos with the cached kernel's two columns per lane: the same
    recurrence as one column per lane (the alpha partials cross the row
    groups in another order, so 1e-12 relative, not bits)."""
    torch = cuda
    jd = lattices[48]
    H = _ham(jd, [0.7, 0.3, 0.9, 0.5], [0.02, 0.01, -0.03])
    bases = [q.enumerate_basis(jd, n, n_threads=8) for n in (0, 1)]
    ops = [q.GpuOp(q.GpuBasis(bb, 48), H) for bb in bases]
    M = 30
    V0s = [torch.from_numpy(ftlm.random_sign_block(bb.shape[0], R, seed=9 + k)).cuda()
           for k, bb in enumerate(bases)]
    for op in ops:
        op.cache_connectivity(True)
        op.set_block_cfg(1)
    a0, b0, _ = ftlm.lanczos_block(ops, V0s, M)
    for op in ops:
        op.set_block_cfg(9)
    a1, b1, _ = ftlm.lanczos_block(ops, V0s, M)
    sc = np.max(np.abs(a0))
    assert np.max(np.abs(a0 - a1)) <= 1e-12 * sc
    assert np.max(np.abs(b0 - b1)) <= 1e-12 * sc


def test_block_autotune_multi_sector(lattices, cuda):
    """ftlm.autotune times the configurations on the two-sector launch and
    sets one that lanczos_block then uses; the recurrence stays the same
    (1e-12 relative: the persistent grid's reduction layout may change)."""
    torch = cuda
    jd = lattices[48]
    H = _ham(jd, [0.7, 0.3, 0.9, 0.5], [0.02, 0.01, -0.03])
    bases = [q.enumerate_basis(jd, n, n_threads=8) for n in (0, 1)]
    ops = [q.GpuOp(q.GpuBasis(bb, 48), H) for bb in bases]
    R, M = 16, 20
    V0s = [torch.from_numpy(ftlm.random_sign_block(bb.shape[0], R, seed=5 + k)).cuda()
           for k, bb in enumerate(bases)]
    for op in ops:
        op.cache_connectivity(True)
    a0, b0, _ = ftlm.lanczos_block(ops, V0s, M)
    cfg, ms = ftlm.autotune(ops, R, reps=2)
    assert cfg in (0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14) and ms > 0
    a1, b1, _ = ftlm.lanczos_block(ops, V0s, M)
    sc = np.max(np.abs(a0))
    assert np.max(np.abs(a0 - a1)) <= 1e-12 * sc
    assert np.max(np.abs(b0 - b1)) <= 1e-12 * sc


def test_c3_two_sectors_against_oracle(lattices, cuda):
    """The batched FTLM workload itself: the 48-spin sectors n = 0 and n = 1
    (uneven sizes) in one launch, R = 64 random-sign start vectors each,
    M = 100 steps, g_s ~ U[0.2, 1] (seeded), B = 0, with the autotuned block
    configuration and graph capture.
      * 4 columns per sector: (alpha, beta) of the first 25 steps against
        the oracle's single-vector recurrence from the same start vector
        (lanczos.cpp:111-146), 1e-10 relative;
      * every column at full M: the Gauss quadrature of T_M reproduces the
        sample's moments <r|H^p|r> for p = 0..3 (exact for p < 2M),
        computed with the oracle's H (the reference's push apply), 1e-10
        relative."""
    torch = cuda
    jd = lattices[48]
    rng = np.random.default_rng(77)
    g = rng.uniform(0.2, 1.0, size=4)
    H = _ham(jd, g)
    bases = [q.enumerate_basis(jd, n, n_threads=8) for n in (0, 1)]
    assert bases[0].shape[0] == 87546
    ops = [q.GpuOp(q.GpuBasis(bb, 48), H) for bb in bases]
    R, M = 64, 100
    V0 = [ftlm.random_sign_block(bb.shape[0], R, seed=77 + k) for k, bb in enumerate(bases)]
    ftlm.autotune(ops, R, reps=1)
    al, be, _ = ftlm.lanczos_block(ops, [torch.from_numpy(v).cuda() for v in V0], M,
                                   use_graph=True)
    for s in range(2):
        ob = orc.Basis(bases[s])
        D = bases[s].shape[0]
        for c in (0, 21, 42, 63):
            oa, obb = orc.lanczos_coeffs(ob, H, V0[s][:, c], 25)
            scale = np.max(np.abs(oa))
            assert np.max(np.abs(al[s, c, :25] - oa)) <= 1e-10 * scale, (s, c)
            assert np.max(np.abs(be[s, c, :25] - obb)) <= 1e-10 * scale, (s, c)
        nodes = ftlm.quadrature(al[s], be[s])
        for c in range(R):
            r = V0[s][:, c]
            h1 = orc.evaluate_add(ob, H, r, np.zeros(D))
            h2 = orc.evaluate_add(ob, H, h1, np.zeros(D))
            mom = [r @ r, r @ h1, h1 @ h1, h1 @ h2]
            e, w = nodes[c]
            got = [np.sum(w * e ** p) for p in range(4)]
            for p in range(4):
                assert abs(got[p] - mom[p]) <= 1e-10 * max(1.0, abs(mom[p])), (s, c, p)


def test_sample_operators_against_dense(lattices, cuda):
    """qsi_ftlm_sample (the operator-expectation sample) against the dense
    propagator: psi_b = exp(-beta_b (H - s) / 2) r from the dense spectrum of
    the oracle's H; <psi|psi> and <psi|O|psi> for a ring term O, 1e-9
    relative; the coefficients equal qsi_lanczos_coeffs' (same recurrence)."""
    torch = cuda
    jd = lattices[32]
    b = load_basis(32)
    D = b.shape[0]
    H = _ham(jd, [0.9, 0.35, 0.6, 0.45], [0.05, -0.02, 0.04])
    basis = q.GpuBasis(b, 32)
    op = q.GpuOp(basis, H)
    ring, _, _ = q.get_ring_ops(jd)
    O = q.SymbolicOpSum()
    O.add_term(1.0, ring[3])
    oop = q.GpuOp(basis, O)
    r = np.random.default_rng(4).standard_normal(D)
    rd = torch.from_numpy(r).cuda()
    betas = np.array([0.2, 0.7, 1.5, 3.0])
    res = ftlm.sample_operators(op, [oop], rd, 90, betas)
    a2, b2 = op.lanczos_coeffs(90, v0=rd)
    assert np.array_equal(res["alpha"], a2) and np.array_equal(res["beta"], b2)
    ob = orc.Basis(b)
    rr, cc, vv = orc.sparse_triplets(ob, H)
    Md = np.zeros((D, D))
    np.add.at(Md, (rr, cc), vv)
    w, U = np.linalg.eigh(Md)
    s = res["e_shift"]
    e, _ = ftlm.tridiag_eig(res["alpha"], res["beta"])
    assert s == e.min()
    rn = r / np.linalg.norm(r)
    for j, bt in enumerate(betas):
        psi = U @ (np.exp(-0.5 * bt * (w - s)) * (U.T @ rn))
        n_ref = psi @ psi
        o_ref = psi @ orc.evaluate_add(ob, O, psi, np.zeros(D))
        assert abs(res["norm"][j] - n_ref) <= 1e-9 * n_ref
        assert abs(res["expect"][0, j] - o_ref) <= 1e-9 * max(abs(o_ref), 1e-3 * n_ref)
