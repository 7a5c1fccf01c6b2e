"""Continuously batched solves (qsi_engine_*) and the serving loop.

Pins: each request's lowest eigenvalue against dense diagonalisation of the
same (g, h) point built without slots (oracle connectivity, 1e-9 relative)
and against the single-solve qsi_eigsh, its further eigenvalues against
levels of the dense spectrum (one copy of a degenerate level is what a
single-vector Krylov solve finds); a request's result is bit-identical
whether it runs alone or in a mix that puts it in another column; the server
gives every request the same bits as a lone engine run even when the HBM
budget forces spaces to be evicted and re-uploaded between requests."""
import numpy as np
import pytest

from conftest import load_basis
from oracle import pyoracle as orc
import qsi_subspace_ed_amd as q
from qsi_subspace_ed_amd import serve

pytestmark = pytest.mark.gpu

EIG_TOL = 1e-9

POINTS = [
    ([-0.9, -0.35, -0.6, -0.45], [0.05, -0.02, 0.04]),
    ([-0.5, -0.5, -0.5, -0.5], [0.0, 0.0, 0.0]),
    ([-0.2, -0.8, -0.4, -0.7], [0.1, 0.1, 0.1]),
    ([-1.0, -0.3, -0.3, -0.3], [0.0, 0.0, 0.3]),
    ([-0.6, -0.6, -0.2, -0.9], [-0.2, 0.05, 0.0]),
]


def _h_of(B):
    loc = q.get_loc_z()
    return [float(np.asarray(B) @ loc[s]) for s in range(4)]


def _plain_h(jd, g, B):
    H = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, g)
    q.add_local_field(H, jd, B)
    return H


def _dense(b, H):
    r, c, v = orc.sparse_triplets(orc.Basis(b), H)
    D = b.shape[0]
    M = np.zeros((D, D))
    np.add.at(M, (r, c), v)
    return M


def _close(a, b):
    return abs(a - b) <= EIG_TOL * max(1.0, abs(b))


@pytest.mark.parametrize("ncv", [12, 24, 30, 40])
def test_engine_vs_dense_and_eigsh(ncv, lattices, cuda):
    """Every Krylov-basis width class of the engine's kernels (up to 16, 24,
    32 and 64 vectors)."""
    jd = lattices[32]
    b = load_basis(32)
    basis = q.GpuBasis(b, 32)
    op = q.GpuOp(basis, serve.slotted_hamiltonian(jd))
    assert op.n_slots == serve.N_SLOTS
    nev = 3
    eng = serve.Engine(op, max_cols=8, nev=nev, ncv=ncv, tol=1e-11)
    ids = {}
    for k, (g, B) in enumerate(POINTS):
        ids[eng.submit(serve.point_coeffs(g, _h_of(B)), seed=100 + k)] = k
    done = eng.run_all()
    assert len(done) == len(POINTS)
    for f in done:
        g, B = POINTS[ids[f.id]]
        assert f.nconv == nev
        H = _plain_h(jd, g, B)
        w = np.linalg.eigvalsh(_dense(b, H))
        # a single-vector Krylov method sees one copy of a degenerate level:
        # the lowest must be w[0], each further one some level of the spectrum
        assert _close(f.evals[0], w[0]), (ids[f.id], f.evals[0], w[0])
        for i in range(1, nev):
            k = int(np.argmin(np.abs(w - f.evals[i])))
            assert _close(f.evals[i], w[k]), (ids[f.id], i, f.evals[i], w[k])
        assert np.all(np.diff(f.evals) >= 0)
        ev, _, _, _ = q.GpuOp(basis, H).eigsh(nev=nev, ncv=ncv, tol=1e-11)
        assert _close(f.evals[0], ev[0])


def test_same_alone_and_in_mix(lattices, cuda):
    jd = lattices[48]
    b = q.enumerate_basis(jd, 1, n_threads=8)
    op = q.GpuOp(q.GpuBasis(b, 48), serve.slotted_hamiltonian(jd))
    target = serve.point_coeffs(POINTS[0][0], _h_of(POINTS[0][1]))
    eng = serve.Engine(op, max_cols=8, nev=2, ncv=20, tol=1e-10)
    eng.submit(target, seed=42)
    alone = eng.run_all()
    assert len(alone) == 1
    # now three others first (columns 0-2), the target in column 3, and two
    # more queued behind so columns are reused while it runs
    for k in (1, 2, 3):
        eng.submit(serve.point_coeffs(POINTS[k][0], _h_of(POINTS[k][1])), seed=k)
    want = eng.submit(target, seed=42)
    for k in (4, 2):
        eng.submit(serve.point_coeffs(POINTS[k][0], _h_of(POINTS[k][1])), seed=50 + k)
    mixed = {f.id: f for f in eng.run_all()}
    assert len(mixed) == 6
    got = mixed[want]
    assert np.array_equal(got.evals, alone[0].evals)
    assert got.steps == alone[0].steps
    # the repeated point with another seed converges to the same eigenvalues
    other = [f for f in mixed.values() if f.id != want]
    assert all(np.all(np.isfinite(f.evals)) for f in other)


def test_server_budget_eviction(lattices, cuda):
    jd32, jd48 = lattices[32], lattices[48]
    b32 = load_basis(32)
    b48 = load_basis(48)
    kw = dict(max_cols=4, nev=1, ncv=16, tol=1e-10)
    probe = serve.SolveServer(1 << 40, **kw)
    s32 = probe.add_space("32", jd32, b32, 32)
    s48 = probe.add_space("48", jd48, b48, 48)
    f32 = s32.footprint(kw["max_cols"], kw["ncv"])
    f48 = s48.footprint(kw["max_cols"], kw["ncv"])
    # room for the larger space but not for both
    srv = serve.SolveServer(max(f32, f48) + min(f32, f48) // 2, **kw)
    srv.add_space("32", jd32, b32, 32)
    srv.add_space("48", jd48, b48, 48)
    reqs = []
    for k in range(8):
        g, B = POINTS[k % len(POINTS)]
        reqs.append(serve.Request("32" if k % 2 == 0 else "48",
                                  serve.point_coeffs(g, _h_of(B)), seed=k, tag=k))
    arrivals = np.linspace(0.0, 0.05, len(reqs))
    out = srv.run(reqs, arrivals)
    assert len(out) == len(reqs)
    assert srv.stats.evictions >= 1
    st = srv.stats.summary()
    assert st["solves"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()
    # each request alone on a fresh engine gives the same bits
    ref = {}
    for key, jd, b, W in (("32", jd32, b32, 32), ("48", jd48, b48, 48)):
        op = q.GpuOp(q.GpuBasis(b, W), serve.slotted_hamiltonian(jd))
        eng = serve.Engine(op, **kw)
        for r in reqs:
            if r.space == key:
                eng.submit(r.coeffs, r.seed)
                ref[r.tag] = eng.run_all()[0].evals
        eng.close()
    for r in out:
        assert np.array_equal(r.result.evals, ref[r.tag]), r.tag


def test_randn_deterministic(cuda):
    """qsi_vec_randn: entries fixed by (seed, i); standard normal moments."""
    from qsi_subspace_ed_amd.device import randn
    a = randn(200001, 7).cpu().numpy()
    b = randn(200001, 7).cpu().numpy()
    c = randn(1000, 7).cpu().numpy()
    assert np.array_equal(a, b) and np.array_equal(a[:1000], c)
    assert not np.array_equal(a[:1000], randn(1000, 8).cpu().numpy())
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1.0) < 0.01


def test_engine_matches_eigsh_same_start(lattices, cuda):
    """An engine request and qsi_eigsh started from the same device Gaussian
    vector converge to the same nev = 5 eigenvalues (1e-9 relative)."""
    from qsi_subspace_ed_amd.device import randn
    jd = lattices[48]
    b = load_basis(48)
    basis = q.GpuBasis(b, 48)
    op = q.GpuOp(basis, serve.slotted_hamiltonian(jd))
    eng = serve.Engine(op, max_cols=4, nev=5, ncv=15, tol=1e-10)
    pts = POINTS[:3]
    ids = {eng.submit(serve.point_coeffs(g, _h_of(B)), seed=500 + k): k
           for k, (g, B) in enumerate(pts)}
    for f in eng.run_all():
        k = ids[f.id]
        g, B = pts[k]
        ev, _, nconv, _ = q.GpuOp(basis, _plain_h(jd, g, B)).eigsh(
            nev=5, ncv=15, tol=1e-10, v0=randn(b.shape[0], 500 + k))
        assert f.nconv == 5 and nconv == 5
        for i in range(5):
            assert _close(f.evals[i], ev[i]), (k, i, f.evals[i], ev[i])


def test_engine_cached_connectivity_bit_identical(lattices, cuda):
    """The engine gives the same bits with and without cached connectivity,
    and with the cached kernel's one or two columns per lane (4, 3, 2 and 1
    live columns as requests finish)."""
    jd = lattices[32]
    b = load_basis(32)
    basis = q.GpuBasis(b, 32)
    out = []
    for cache, cfg in ((False, None), (True, 1), (True, 9), (True, 8)):
        op = q.GpuOp(basis, serve.slotted_hamiltonian(jd))
        if cfg is not None:
            op.set_block_cfg(cfg)
        eng = serve.Engine(op, max_cols=4, nev=3, ncv=16, tol=1e-10, cache_connectivity=cache,
                           autotune=cfg is None)
        for k, (g, B) in enumerate(POINTS[:4]):
            eng.submit(serve.point_coeffs(g, _h_of(B)), seed=70 + k)
        out.append({f.id: f.evals for f in eng.run_all()})
        eng.close()
    for o in out[1:]:
        assert out[0].keys() == o.keys()
        for k in out[0]:
            assert np.array_equal(out[0][k], o[k])


def test_server_concurrent_engines_same_bits(lattices, cuda):
    """Several resident spaces step side by side on their own streams; every
    request gets the same bits as on a lone engine."""
    jd32, jd48 = lattices[32], lattices[48]
    spaces = {"32/0": (jd32, load_basis(32), 32),
              "32/1": (jd32, q.enumerate_basis(jd32, 1, n_threads=8), 32),
              "48/0": (jd48, load_basis(48), 48)}
    kw = dict(max_cols=4, nev=3, ncv=12, tol=1e-10)
    srv = serve.SolveServer(1 << 40, **kw)
    for key, (jd, b, W) in spaces.items():
        srv.add_space(key, jd, b, W)
    keys = list(spaces)
    reqs = [serve.Request(keys[k % 3], serve.point_coeffs(POINTS[k % 5][0], _h_of(POINTS[k % 5][1])),
                          seed=900 + k, tag=k) for k in range(10)]
    out = srv.run(reqs)
    srv.close()
    assert len(out) == len(reqs)
    for key, (jd, b, W) in spaces.items():
        op = q.GpuOp(q.GpuBasis(b, W), serve.slotted_hamiltonian(jd))
        eng = serve.Engine(op, **kw)
        for r in reqs:
            if r.space == key:
                eng.submit(r.coeffs, r.seed)
                ref = eng.run_all()[0]
                got = r.result
                assert got.steps == ref.steps, (r.tag, got.steps, ref.steps)
                assert np.array_equal(got.evals, ref.evals), (r.tag, got.evals, ref.evals)
        eng.close()


def test_engine_first_step_after_create(lattices, cuda):
    """Requests admitted in the very first step of a freshly created engine
    on a large basis (its creation-time zeroing must be finished by then)
    converge like the same requests admitted later."""
    jd = lattices[48]
    b = q.enumerate_basis(jd, 1, n_threads=8)
    basis = q.GpuBasis(b, 48)
    pts = [serve.point_coeffs(g, _h_of(B)) for g, B in POINTS[:4]]
    res = []
    for warm in (False, True):
        op = q.GpuOp(basis, serve.slotted_hamiltonian(jd))
        eng = serve.Engine(op, max_cols=16, nev=3, ncv=15, tol=1e-10)
        if warm:
            eng.submit(pts[0], seed=1)
            eng.run_all()
        ids = [eng.submit(c, seed=40 + k) for k, c in enumerate(pts)]
        got = {f.id: f for f in eng.run_all()}
        res.append([got[i] for i in ids])
        eng.close()
    for a, b2 in zip(*res):
        assert a.nconv == 3 and b2.nconv == 3
        assert a.steps > 1
        np.testing.assert_allclose(a.evals, b2.evals, rtol=1e-9, atol=0)


def test_lookahead_same_as_one_step_at_a_time(lattices, cuda):
    """Two steps in flight (qsi_engine_step keeps the next step enqueued while
    the host finishes the previous; a request whose basis fills sits out until
    its restart) give every request the same bits and step counts as strictly
    one step at a time."""
    jd = lattices[48]
    b = q.enumerate_basis(jd, 0, n_threads=8)
    op = q.GpuOp(q.GpuBasis(b, 48), serve.slotted_hamiltonian(jd))

    def run(lookahead):
        eng = serve.Engine(op, max_cols=4, nev=3, ncv=12, tol=1e-10)
        for k, (g, B) in enumerate(POINTS):
            eng.submit(serve.point_coeffs(g, _h_of(B)), seed=70 + k)
        done = []
        while eng.busy():
            if lookahead:
                eng.step()
            else:
                eng.step_begin()
                assert eng.inflight <= 1
                eng.step_end()
            done += eng.poll()
        eng.close()
        return {f.id: f for f in done}

    one, two = run(False), run(True)
    assert one.keys() == two.keys() and len(one) == len(POINTS)
    for i in one:
        assert one[i].steps == two[i].steps, i
        assert np.array_equal(one[i].evals, two[i].evals), i
