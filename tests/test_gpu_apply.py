"""GPU parity of the device basis lookup, the matrix-free apply (single vector
and block) and the connectivity materialiser against the oracle (the
reference's push algorithm on the CPU) on the fixture bases.

Tolerances: ranks and (row, col) connectivity bit-exact; matrix elements
exact (products of a coefficient and +-1); H·v within 1e-13 relative 2-norm.
"""
import numpy as np
import pytest

from conftest import load_basis
from oracle import pyoracle as orc
import qsi_subspace_ed_amd as q
from qsi_subspace_ed_amd._lib import QsiError
from qsi_subspace_ed_amd.lattice import add_ring_terms

pytestmark = pytest.mark.gpu

HV_TOL = 1e-13


def _ham(jd, W, kind):
    if kind == "ring":
        return q.build_hamiltonian_g(jd, [-1.0, -0.55, -0.3, -0.8])
    if kind == "ring+field":
        H = q.build_hamiltonian_g(jd, [-0.9, -0.25, -0.6, -0.45])
        q.add_local_field(H, jd, [0.07, -0.02, 0.05])
        return H
    if kind == "jpm":
        return q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    raise ValueError(kind)


def _rand_x(D, seed, tiny=True):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(D)
    if tiny:  # exercise the |x| <= 1e-15 skip
        x[rng.integers(0, D, size=max(1, D // 50))] = 3e-16
    return x


@pytest.mark.parametrize("W", [8, 32, 48])
@pytest.mark.parametrize("kbits", [0, 1, 5])
def test_search_ranks(W, kbits, cuda):
    b = load_basis(W)
    basis = q.GpuBasis(b, W, bucket_bits=min(kbits, W))
    assert np.array_equal(basis.search(b), np.arange(b.shape[0]))
    rng = np.random.default_rng(W)
    members = set(int(x) for x in b[:, 0])
    cand = rng.integers(0, 1 << W if W < 64 else (1 << 63), size=4000, dtype=np.uint64)
    cand = np.array([c for c in cand if int(c) not in members][:2000], dtype=np.uint64)
    qs = np.zeros((cand.shape[0], 2), dtype=np.uint64)
    qs[:, 0] = cand
    assert np.all(basis.search(qs) == -1)
    qs[:, 1] = 1  # hi word set: outside a <= 64-spin basis
    assert np.all(basis.search(qs[:10]) == -1)
    # the same through the slot index
    basis.build_slots()
    assert basis.has_slots()
    assert np.array_equal(basis.search(b), np.arange(b.shape[0]))
    assert np.all(basis.search(qs) == -1)
    qs[:, 1] = 0
    assert np.all(basis.search(qs) == -1)
    basis.build_slots(False)
    assert not basis.has_slots()


@pytest.mark.parametrize("variant", ["auto", "window", "lane", "lane8", "hash", "slot4", "low4",
                                     "cached"])
@pytest.mark.parametrize("W,kind", [(8, "ring"), (32, "ring"), (32, "ring+field"),
                                    (48, "ring+field"), (48, "jpm")])
def test_apply_parity(W, kind, variant, lattices, cuda):
    torch = cuda
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    H = _ham(jd, W, kind)
    gb = q.GpuBasis(b, W, hash_index=True)
    if variant == "slot4":
        gb.build_slots()
    if variant == "low4":
        gb.build_lowkeys()
    op = q.GpuOp(gb, H)
    if variant == "cached":
        op.cache_connectivity(True)
    op.set_variant(variant)
    ob = orc.Basis(b)
    x = _rand_x(D, W)
    y0 = np.random.default_rng(W + 1).standard_normal(D)
    xd = torch.from_numpy(x).cuda()
    # evaluate: y = H x
    yd = torch.full((D,), 7.0, dtype=torch.float64, device="cuda")
    op.evaluate(xd, yd)
    ref = orc.evaluate_add(ob, H, x, np.zeros(D))
    got = yd.cpu().numpy()
    assert np.linalg.norm(got - ref) <= HV_TOL * np.linalg.norm(ref)
    # evaluate_add: y += H x
    yd = torch.from_numpy(y0.copy()).cuda()
    op.evaluate_add(xd, yd)
    ref = orc.evaluate_add(ob, H, x, y0.copy())
    assert np.linalg.norm(yd.cpu().numpy() - ref) <= HV_TOL * np.linalg.norm(ref)
    # host-vector path
    assert np.linalg.norm(op.evaluate_host(x) - orc.evaluate_add(ob, H, x, np.zeros(D))) \
        <= HV_TOL * np.linalg.norm(ref)


@pytest.mark.parametrize("W,kind", [(32, "ring+field"), (48, "jpm")])
def test_connectivity_bit_exact(W, kind, lattices, cuda):
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    H = _ham(jd, W, kind)
    op = q.GpuOp(q.GpuBasis(b, W), H)
    r, c, v, d = op.connectivity()
    orr, occ, ovv = orc.sparse_triplets(orc.Basis(b), H)
    off = orr != occ
    # reference triplets are (J, j, H[J, j]); the GPU lists H[row, col]
    o_key = np.lexsort((occ[off], orr[off]))
    g_key = np.lexsort((c, r))
    assert np.array_equal(r[g_key], orr[off][o_key])
    assert np.array_equal(c[g_key], occ[off][o_key])
    assert np.array_equal(v[g_key], ovv[off][o_key])
    diag = np.zeros(D)
    diag[orr[~off]] = ovv[~off]
    # summation order differs (merged field groups), so compare to the scale
    assert np.max(np.abs(d - diag)) <= 1e-14 * np.max(np.abs(diag))
    if kind == "ring+field" and W == 32:
        assert r.shape[0] == 15360


@pytest.mark.parametrize("W,kbits", [(32, 0), (48, 0), (48, 3)])
def test_variants_bit_identical(W, kbits, lattices, cuda):
    """Window and lane kernels add the same terms in the same order."""
    torch = cuda
    jd = lattices[W]
    b = load_basis(W)
    H = _ham(jd, W, "jpm" if W == 48 else "ring+field")
    basis = q.GpuBasis(b, W, bucket_bits=kbits, hash_index=True)
    basis.build_slots()
    basis.build_lowkeys()
    op = q.GpuOp(basis, H)
    x = torch.from_numpy(_rand_x(b.shape[0], 9)).cuda()
    ys = []
    for v in ("window", "lane", "lane2", "lane4", "lane8", "hash", "hash4", "slot2", "slot4",
              "slot3", "low2", "low4", "low8", "cached"):
        if v == "cached":
            op.cache_connectivity(True)
        op.set_variant(v)
        y = torch.zeros_like(x)
        op.evaluate(x, y)
        ys.append(y)
    # every gather depth of the cached kernel
    for pipe in (4, 8, 16, 32):
        op.set_cached_pipe(pipe)
        assert op.cached_pipe() == pipe
        y = torch.zeros_like(x)
        op.evaluate(x, y)
        ys.append(y)
    for y in ys[1:]:
        assert torch.equal(ys[0], y)
    # the low-key kernels after the verification pass (ring flips on a full
    # sector reach only basis states, so no compare is needed there) give the
    # same bits
    exact = op.verify_lowkeys()
    if W == 32:
        assert exact == 2
    for v in ("low2", "low4", "low8"):
        op.set_variant(v)
        y = torch.zeros_like(x)
        op.evaluate(x, y)
        assert torch.equal(ys[0], y), (v, exact)
    op.set_variant("cached")
    # y += H x and y = 0.5 y + H x through the cached kernel match too
    y0 = torch.from_numpy(_rand_x(b.shape[0], 10)).cuda()
    y1 = y0.clone()
    op.evaluate_add(x, y1)
    op.set_variant("lane8")
    y2 = y0.clone()
    op.evaluate_add(x, y2)
    assert torch.equal(y1, y2)
    assert op.cache_connectivity(False) == -1


def test_apply_many_terms(lattices, cuda):
    """More off-diagonal terms than one LDS staging chunk holds (256): the
    32-spin ring terms five times over, with two coefficient sets in turn (320
    flips, five coefficient classes, so the connectivity can be cached), plus
    the local field.  Every lookup kernel stages the terms in chunks and gives
    the oracle's y, with the same bits across variants."""
    torch = cuda
    jd = lattices[32]
    b = load_basis(32)
    D = b.shape[0]
    H = q.SymbolicOpSum()
    for rep in range(5):
        add_ring_terms(H, jd, [-1.0 + 0.4 * (rep % 2), -0.55, -0.3, -0.8])
    q.add_local_field(H, jd, [0.07, -0.02, 0.05])
    assert len(H.off_diag_terms) > 256
    basis = q.GpuBasis(b, 32)
    basis.build_lowkeys()
    op = q.GpuOp(basis, H)
    x = _rand_x(D, 17)
    ref = orc.evaluate_add(orc.Basis(b), H, x, np.zeros(D))
    xd = torch.from_numpy(x).cuda()
    ys = []
    for v in ("lane", "lane4", "lane8", "low2", "low4", "low8", "window", "slot4", "cached"):
        if v == "slot4":
            basis.build_slots()
        if v == "cached":
            op.cache_connectivity(True)
        op.set_variant(v)
        y = torch.zeros_like(xd)
        op.evaluate(xd, y)
        assert np.linalg.norm(y.cpu().numpy() - ref) <= HV_TOL * np.linalg.norm(ref), v
        ys.append((v, y))
    assert op.verify_lowkeys() == 2
    for v in ("low2", "low4", "low8"):
        op.set_variant(v)
        y = torch.zeros_like(xd)
        op.evaluate(xd, y)
        ys.append((v + " verified", y))
    for v, y in ys[1:]:
        assert torch.equal(ys[0][1], y), v
    # AUTO without the cache takes the low-key kernel (LOW8 after a mode-2 check)
    op.cache_connectivity(False)
    op.set_variant("auto")
    y = torch.zeros_like(xd)
    op.evaluate(xd, y)
    assert torch.equal(ys[0][1], y)
    # the dot-product instances (the Lanczos step) through the chunked kernels
    for v in ("lane4", "low4", "low8"):
        op.set_variant(v)
        a, bt = op.lanczos_coeffs(6, seed=3)
        if v == "lane4":
            a0, b0 = a, bt
        else:
            assert np.array_equal(a, a0) and np.array_equal(bt, b0)


def test_apply_partial_basis(lattices, cuda):
    """A basis missing some connected states (every third state dropped):
    targets absent from the basis are skipped, as the reference's search
    returning 0 makes the pushed contribution vanish."""
    torch = cuda
    jd = lattices[32]
    b = load_basis(32)[::3].copy()
    H = _ham(jd, 32, "ring+field")
    for v in ("window", "lane", "lane2", "lane4", "lane8", "hash", "hash4", "slot2", "slot4",
              "slot3", "low2", "low4", "low8"):
        basis = q.GpuBasis(b, 32, hash_index=True)
        basis.build_slots()
        basis.build_lowkeys()
        op = q.GpuOp(basis, H)
        if v.startswith("low"):
            op.verify_lowkeys()  # whatever it finds, the result must stay exact
        op.set_variant(v)
        x = _rand_x(b.shape[0], 4, tiny=False)
        y = torch.zeros(b.shape[0], dtype=torch.float64, device="cuda")
        op.evaluate(torch.from_numpy(x).cuda(), y)
        ref = orc.evaluate_add(orc.Basis(b), H, x, np.zeros(b.shape[0]))
        assert np.linalg.norm(y.cpu().numpy() - ref) <= HV_TOL * np.linalg.norm(ref)


def _relabel(jd, f, W2):
    """Copy of a lattice with spin i renamed f(i) (for >64-spin code paths)."""
    import copy
    j2 = copy.deepcopy(jd)
    atoms = [None] * W2
    for i, a in enumerate(jd["atoms"]):
        atoms[f(i)] = a
    filler = {"xyz": [0, 0, 0], "sl": "0"}
    j2["atoms"] = [a if a is not None else filler for a in atoms]
    for t in j2["tetrahedra"]:
        t["member_spin_idx"] = [f(s) for s in t["member_spin_idx"]]
    for r in j2["rings"]:
        r["member_spin_idx"] = [f(s) for s in r["member_spin_idx"]]
    for bd in j2["bonds"]:
        bd["from_idx"], bd["to_idx"] = f(bd["from_idx"]), f(bd["to_idx"])
    return j2


def test_apply_parity_two_words(lattices, cuda):
    """48 spins spread over bits 20..114: exercises the 128-bit state path."""
    torch = cuda
    W2 = 115
    f = lambda i: 2 * i + 20  # noqa: E731
    jd = _relabel(lattices[48], f, W2)
    b48 = load_basis(48)
    states = []
    for lo in b48[:, 0]:
        s = 0
        lo = int(lo)
        for i in range(48):
            if (lo >> i) & 1:
                s |= 1 << f(i)
        states.append(s)
    states.sort()
    lohi = q.as_lohi(states)
    H = q.build_hamiltonian_g(jd, [-0.9, -0.25, -0.6, -0.45])
    for i in range(48):  # a field on the real spins only
        H.add_term(0.013 * ((i % 4) - 1.5), q.SymbolicPMROperator.from_ops("z", [f(i)]))
    H.add_term(0.021, q.SymbolicPMROperator.from_ops("zz", [f(3), f(40)]))
    gb = q.GpuBasis(lohi, W2, hash_index=True)
    gb.build_slots()
    gb.build_lowkeys()
    op = q.GpuOp(gb, H)
    x = _rand_x(len(states), 5)
    ref = orc.evaluate_add(orc.Basis(lohi), H, x, np.zeros(len(states)))
    for v in ("window", "lane", "lane2", "lane4", "lane8", "hash", "hash4", "slot2", "slot4",
              "slot3", "low2", "low4", "low8"):
        op.set_variant(v)
        yd = torch.zeros(len(states), dtype=torch.float64, device="cuda")
        op.evaluate(torch.from_numpy(x).cuda(), yd)
        assert np.linalg.norm(yd.cpu().numpy() - ref) <= HV_TOL * np.linalg.norm(ref)
    basis = q.GpuBasis(lohi, W2)
    assert np.array_equal(basis.search(lohi), np.arange(len(states)))
    basis.build_slots()
    assert np.array_equal(basis.search(lohi), np.arange(len(states)))
    # a slot variant falls back to the bisection once the index is dropped
    op.set_variant("slot4")
    gb.build_slots(False)
    yd = torch.zeros(len(states), dtype=torch.float64, device="cuda")
    op.evaluate(torch.from_numpy(x).cuda(), yd)
    assert np.linalg.norm(yd.cpu().numpy() - ref) <= HV_TOL * np.linalg.norm(ref)
    # and a low-key variant once the low keys are dropped
    op.set_variant("low8")
    gb.build_lowkeys(False)
    assert not gb.has_lowkeys()
    yd = torch.zeros(len(states), dtype=torch.float64, device="cuda")
    op.evaluate(torch.from_numpy(x).cuda(), yd)
    assert np.linalg.norm(yd.cpu().numpy() - ref) <= HV_TOL * np.linalg.norm(ref)
    with pytest.raises(QsiError, match="no low keys"):
        op.set_variant("low4")


@pytest.mark.parametrize("R", [1, 3, 16, 64])
def test_block_apply_matches_columns(R, lattices, cuda):
    torch = cuda
    W = 32
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    H = _ham(jd, W, "ring+field")
    op = q.GpuOp(q.GpuBasis(b, W), H)
    X = torch.from_numpy(np.random.default_rng(R).standard_normal((D, R))).cuda()
    Y = torch.zeros((D, R), dtype=torch.float64, device="cuda")
    op.apply_block(X, Y)
    for r in range(R):
        y = torch.zeros(D, dtype=torch.float64, device="cuda")
        op.evaluate(X[:, r].contiguous(), y)
        assert torch.equal(Y[:, r], y)


def test_block_apply_per_column_coefficients(lattices, cuda):
    torch = cuda
    W = 48
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    R = 5
    rng = np.random.default_rng(3)
    gs = rng.uniform(0.0, 1.0, size=(R, 4))
    Bs = rng.normal(size=(R, 3)) * 0.1
    # slot s < 4: ring sublattice s; slot 4 + s: field on sublattice s
    Hs = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, [0.0] * 4, slots=True)
    q.add_local_field(Hs, jd, [0.0, 0.0, 0.0], slots=True)
    bas = q.GpuBasis(b, W)
    op = q.GpuOp(bas, Hs, n_slots=8)
    coeffs = np.zeros((8, R))
    loc = q.get_loc_z()
    for r in range(R):
        coeffs[:4, r] = -gs[r]
        coeffs[4:, r] = [-0.5 * float(Bs[r] @ loc[s]) for s in range(4)]
    X = torch.from_numpy(rng.standard_normal((D, R))).cuda()
    Y = torch.zeros((D, R), dtype=torch.float64, device="cuda")
    op.apply_block(X, Y, coeffs=torch.from_numpy(coeffs).cuda())
    ob = orc.Basis(b)
    for r in range(R):
        H = q.build_hamiltonian_g(jd, list(-gs[r]))
        q.add_local_field(H, jd, Bs[r])
        ref = orc.evaluate_add(ob, H, X[:, r].cpu().numpy(), np.zeros(D))
        got = Y[:, r].cpu().numpy()
        assert np.linalg.norm(got - ref) <= HV_TOL * np.linalg.norm(ref)


@pytest.mark.parametrize("R", [1, 5, 16, 64])
def test_cached_connectivity_block_apply_identical(R, lattices, cuda):
    """Block applies through the cached connectivity give the same bits as
    the searching kernel, with and without per-column coefficients."""
    torch = cuda
    W = 48
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    Hs = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, [0.0] * 4, slots=True)
    q.add_local_field(Hs, jd, [0.0, 0.0, 0.0], slots=True)
    op = q.GpuOp(q.GpuBasis(b, W), Hs, n_slots=8)
    rng = np.random.default_rng(R)
    X = torch.from_numpy(rng.standard_normal((D, R))).cuda()
    co = torch.from_numpy(rng.standard_normal((8, R))).cuda()
    Y0 = torch.zeros((D, R), dtype=torch.float64, device="cuda")
    Y1 = torch.zeros_like(Y0)
    op.apply_block(X, Y0, coeffs=co)
    nnz = op.cache_connectivity(True)
    assert nnz == 703680  # off-diagonal nonzeros of the 48-spin n=0 fixture
    op.apply_block(X, Y1, coeffs=co)
    assert torch.equal(Y0, Y1)
    op.apply_block(X, Y0)
    op.apply_block(X, Y1, accumulate=False)
    assert torch.equal(Y0, Y1)
    assert op.cache_connectivity(False) == -1


@pytest.mark.parametrize("R,ldx", [(2, 2), (3, 3), (3, 4), (5, 6), (6, 6), (6, 8), (8, 8),
                                   (16, 16), (64, 64)])
def test_cached_block_apply_column_pairs(R, ldx, lattices, cuda):
    """The cached block apply with two columns per lane (configuration bit 3)
    gives the bits of the one-column kernel at every gather depth, for X rows
    padded or not (an odd unpadded R runs the one-column kernel)."""
    torch = cuda
    W = 48
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    Hs = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, [0.0] * 4, slots=True)
    q.add_local_field(Hs, jd, [0.0, 0.0, 0.0], slots=True)
    op = q.GpuOp(q.GpuBasis(b, W), Hs, n_slots=8)
    op.cache_connectivity(True)
    rng = np.random.default_rng(100 + R)
    Xp = torch.from_numpy(rng.standard_normal((D, ldx))).cuda()
    X = Xp[:, :R]
    co = torch.from_numpy(rng.standard_normal((8, R))).cuda()
    for k in range(3):
        ys = []
        for cfg in (k, k | 8, k | 8 | 4):
            op.set_block_cfg(cfg)
            assert op.block_cfg() == cfg
            Y = torch.full((D, R), 7.0, dtype=torch.float64, device="cuda")
            op.apply_block(X, Y, coeffs=co)
            Ya = torch.ones((D, R), dtype=torch.float64, device="cuda")
            op.apply_block(X, Ya, coeffs=co, accumulate=True)
            ys.append((Y, Ya))
        for Y, Ya in ys[1:]:
            assert torch.equal(Y, ys[0][0]), k
            assert torch.equal(Ya, ys[0][1]), k


def test_cached_layouts_identical(lattices, cuda):
    """The cached connectivity padded to groups of four (the default when
    memory allows) and unpadded (first and last group of a row masked) give
    the searching kernel's bits, single-vector at every gather depth and in
    block applies of one and two columns per lane."""
    torch = cuda
    W = 48
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    Hs = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, [0.0] * 4, slots=True)
    q.add_local_field(Hs, jd, [0.0, 0.0, 0.0], slots=True)
    H = q.build_hamiltonian_g(jd, [-0.9, -0.35, -0.6, -0.45])
    q.add_local_field(H, jd, [0.05, -0.02, 0.04])
    op = q.GpuOp(q.GpuBasis(b, W), H)
    ops = q.GpuOp(q.GpuBasis(b, W), Hs, n_slots=8)
    rng = np.random.default_rng(31)
    x = torch.from_numpy(rng.standard_normal(D)).cuda()
    R = 6
    X = torch.from_numpy(rng.standard_normal((D, R))).cuda()
    co = torch.from_numpy(rng.standard_normal((8, R))).cuda()
    op.set_variant("lane4")
    y_ref = torch.zeros(D, dtype=torch.float64, device="cuda")
    op.evaluate(x, y_ref)
    Y_ref = torch.zeros((D, R), dtype=torch.float64, device="cuda")
    ops.apply_block(X, Y_ref, coeffs=co)
    for padded, pad in ((False, 1), (None, 4)):
        assert op.cache_connectivity(True, padded=padded) == 703680
        assert op.connectivity_padding() == pad
        op.set_variant("cached")
        for pipe in (4, 8, 16, 32):
            op.set_cached_pipe(pipe)
            y = torch.zeros(D, dtype=torch.float64, device="cuda")
            op.evaluate(x, y)
            assert torch.equal(y, y_ref), (pad, pipe)
        op.cache_connectivity(False)
        assert op.connectivity_padding() == -1
        ops.cache_connectivity(True, padded=padded)
        assert ops.connectivity_padding() == pad
        for cfg in (0, 1, 2, 8, 9, 10):
            ops.set_block_cfg(cfg)
            Y = torch.zeros((D, R), dtype=torch.float64, device="cuda")
            ops.apply_block(X, Y, coeffs=co)
            assert torch.equal(Y, Y_ref), (pad, cfg)
        ops.cache_connectivity(False)


def test_autotune_keeps_results(lattices, cuda):
    """The autotuner picks a kernel by timing; every pick gives the same bits."""
    torch = cuda
    W = 48
    jd = lattices[W]
    b = load_basis(W)
    D = b.shape[0]
    Hs = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, [0.0] * 4, slots=True)
    q.add_local_field(Hs, jd, [0.0, 0.0, 0.0], slots=True)
    op = q.GpuOp(q.GpuBasis(b, W), Hs, n_slots=8)
    x = torch.from_numpy(_rand_x(D, 3)).cuda()
    y_ref = torch.zeros_like(x)
    op.set_variant("lane8")
    op.evaluate(x, y_ref)
    name, ms = op.autotune(0, reps=2)
    assert name in q._lib.VARIANTS and ms > 0
    y = torch.zeros_like(x)
    op.evaluate(x, y)
    assert torch.equal(y, y_ref)
    R = 8
    X = torch.from_numpy(np.random.default_rng(4).standard_normal((D, R))).cuda()
    co = torch.from_numpy(np.random.default_rng(5).standard_normal((8, R))).cuda()
    Y0 = torch.zeros_like(X)
    op.apply_block(X, Y0, coeffs=co)
    for cache in (False, True):
        op.cache_connectivity(cache)
        cfg, ms = op.autotune(R, reps=2)
        assert 0 <= cfg < 16 and ms > 0
        Y1 = torch.zeros_like(X)
        op.apply_block(X, Y1, coeffs=co)
        assert torch.equal(Y0, Y1)


def _expected_key_fields(ints, n_spins, K):
    """The bucket key fields the library documents, from the sorted states as
    Python integers: informative positions (where a neighbouring pair first
    differs) from the top, runs as fields, at most 4 fields (the last one then
    grows down over gaps), K bits in all."""
    info = set()
    for a, c in zip(ints[:-1], ints[1:]):
        info.add((a ^ c).bit_length() - 1)
    fields, width = [], 0
    for p in range(n_spins - 1, -1, -1):
        if width >= K:
            break
        if p not in info:
            continue
        if fields and fields[-1][0] == p + 1:
            fields[-1] = (p, fields[-1][1] + 1)
            width += 1
        elif len(fields) < 4:
            fields.append((p, 1))
            width += 1
        else:
            grow = fields[-1][0] - p
            if width + grow > K:
                break
            fields[-1] = (p, fields[-1][1] + grow)
            width += grow
    return fields


@pytest.mark.parametrize("W,kbits", [(32, 0), (48, 0), (48, 7), (48, 30)])
def test_bucket_key_fields(W, kbits, cuda):
    """The device's choice of bucket key fields matches the rule recomputed on
    the host, the key is non-decreasing along the basis, and every state is
    found at its rank."""
    b = load_basis(W)
    basis = q.GpuBasis(b, W, bucket_bits=kbits)
    ints = [int(v) for v in b[:, 0]]
    K = basis.bucket_bits()
    fields = basis.key_fields()
    assert fields == _expected_key_fields(ints, W, K if kbits == 0 else kbits)
    assert sum(w for _, w in fields) == K
    keys = []
    for s in ints:
        k = 0
        for lo, w in fields:
            k = (k << w) | ((s >> lo) & ((1 << w) - 1))
        keys.append(k)
    assert all(a <= c for a, c in zip(keys[:-1], keys[1:]))
    assert np.array_equal(basis.search(b), np.arange(b.shape[0]))


def test_bucket_key_fields_two_words(lattices, cuda):
    """Scattered informative positions over both words (48 spins on every
    other bit of 20..114): four fields, the last grown over the gaps."""
    f = lambda i: 2 * i + 20  # noqa: E731
    b48 = load_basis(48)
    ints = sorted(sum(1 << f(i) for i in range(48) if (int(lo) >> i) & 1) for lo in b48[:, 0])
    basis = q.GpuBasis(q.as_lohi(ints), 115)
    fields = basis.key_fields()
    K = basis.bucket_bits()
    assert len(fields) == 4 and fields == _expected_key_fields(ints, 115, K)
    assert np.array_equal(basis.search(q.as_lohi(ints)), np.arange(len(ints)))
