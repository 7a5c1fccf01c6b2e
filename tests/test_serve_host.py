"""Host side of the serving layer (no GPU): the slotted Hamiltonian has the
same terms as the plain one, slot coefficients reproduce the plain
coefficients, the engine footprint agrees with its buffer list, and the
server's admission keeps resident spaces within the HBM budget."""
import ctypes as C

import numpy as np
import pytest

import qsi_subspace_ed_amd as q
from qsi_subspace_ed_amd import _lib, serve


def _lat(W):
    from conftest import LATTICE_OF
    return q.load_lattice(q.lattice_path(LATTICE_OF[W]))


def test_slotted_terms_match_plain():
    jd = _lat(32)
    g = [-0.9, -0.35, -0.6, -0.45]
    B = [0.05, -0.02, 0.04]
    h = [float(np.asarray(B) @ q.get_loc_z()[s]) for s in range(4)]
    Hs = serve.slotted_hamiltonian(jd)
    assert Hs.n_slots() == serve.N_SLOTS
    coeffs = serve.point_coeffs(g, h)
    plain = q.lattice.add_ring_terms(q.SymbolicOpSum(), jd, g)
    q.add_local_field(plain, jd, B)
    got = [(coeffs[s], op) for (_, op, s) in list(Hs.off_diag_terms) + list(Hs.diagonal_terms)]
    want = [(c, op) for (c, op, _) in list(plain.off_diag_terms) + list(plain.diagonal_terms)]
    assert len(got) == len(want)
    for (c1, o1), (c2, o2) in zip(got, want):
        assert (o1.X, o1.Z, o1.up, o1.down, o1.sign) == (o2.X, o2.Z, o2.up, o2.down, o2.sign)
        assert c1 == c2


def test_jpm_point_matches_g_vals():
    g, h = serve.jpm_point(0.02, [0.1, 0.0, 0.0])
    assert g == q.g_vals(0.02, [0.1, 0.0, 0.0])
    assert np.allclose(h, [0.1 / np.sqrt(3)] * 2 + [-0.1 / np.sqrt(3)] * 2)


def test_point_coeffs_validation():
    try:
        serve.point_coeffs([1, 2, 3], [0, 0, 0, 0])
    except ValueError:
        pass
    else:
        raise AssertionError("expected ValueError")


def test_engine_footprint_formula():
    lib = _lib.load()
    for D, R, ncv in ((2970, 8, 24), (2756394, 16, 20), (1, 1, 2)):
        ld = (D + 63) // 64 * 64
        # Krylov bases R x (ncv + 1) x ld, X (D x R), w (R x ld); restarts are
        # done in place
        want = 8 * (R * (ncv + 1) * ld + D * R + R * ld)
        assert lib.qsi_engine_footprint(D, R, ncv) == want


def test_engine_create_validation():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.qsi_engine_create(None, 4, 1, 10, 1e-10, 10, C.byref(h)) == _lib.QSI_E_INVALID
    assert lib.qsi_engine_poll(None, None, None, None, None) == 0
    assert lib.qsi_engine_step(None, None) == _lib.QSI_E_INVALID


def test_default_algorithm_by_size():
    """diagonalise_real's default (include/matrix_diag_bits.hpp:126-136)."""
    from qsi_subspace_ed_amd import diag
    assert diag.choose_algorithm(18) == "dense"
    assert diag.choose_algorithm(99) == "dense"
    assert diag.choose_algorithm(100) == "sparse"
    assert diag.choose_algorithm(9_999_999) == "sparse"
    assert diag.choose_algorithm(10_000_000) == "mfsparse"


class _Closable:
    def close(self):
        pass


class _FakeEngine:
    """Engine stand-in: every request finishes after three finished steps."""

    def __init__(self):
        self.jobs, self.next, self.inflight = {}, 0, 0

    def submit(self, coeffs, seed):
        self.jobs[self.next] = 3
        self.next += 1
        return self.next - 1

    active = property(lambda self: len(self.jobs))
    queued = property(lambda self: 0)

    def busy(self):
        return self.inflight > 0 or bool(self.jobs)

    def ready(self):
        return True

    def step_begin(self):
        if not self.jobs:
            return False
        self.inflight += 1
        return True

    def step_end(self):
        self.inflight -= 1
        for k in self.jobs:
            self.jobs[k] -= 1

    def poll(self):
        done = [k for k, v in self.jobs.items() if v <= 0]
        for k in done:
            del self.jobs[k]
        return [serve.Finished(k, np.zeros(1), 1, 3) for k in done]

    def close(self):
        assert not self.busy(), "a busy engine was evicted"


def test_server_admission_under_budget(monkeypatch):
    """Three spaces of 100 bytes under a 250-byte budget: at most two are
    resident, a space is uploaded only once an idle one is evicted, and every
    request is answered."""
    monkeypatch.setattr(serve.Space, "footprint", lambda self, m, n: self.dim)
    srv = serve.SolveServer(250)

    def upload(sp):
        sp.engine, sp.op, sp.basis = _FakeEngine(), _Closable(), _Closable()
        srv.stats.uploads += 1
        assert srv.resident_bytes() <= srv.budget

    monkeypatch.setattr(srv, "_upload", upload)
    for key in "abc":
        srv.add_space(key, {}, np.zeros((100, 2), np.uint64), 16)
    with pytest.raises(ValueError):
        srv.add_space("big", {}, np.zeros((300, 2), np.uint64), 16)
    reqs = [serve.Request(space=k, coeffs=np.zeros(8), seed=i) for i, k in enumerate("abcab")]
    out = srv.run(reqs)
    assert sorted(r.id for r in out) == list(range(5))
    assert all(r.result is not None and r.result.nconv == 1 for r in out)
    assert srv.stats.uploads >= 3 and srv.stats.evictions >= 1
    assert sum(s.engine is not None for s in srv.spaces.values()) <= 2
    assert not srv.busy()


@pytest.mark.parametrize("W", [32, 48])
def test_padded_nnz_counts_oracle_rows(W):
    # the footprint's connectivity count: per row, the oracle's off-diagonal
    # entries (the reference's toSparseMatrix rows), padded to groups of 4
    from conftest import load_basis
    from oracle import pyoracle as orc
    jd = _lat(W)
    b = load_basis(W)
    sp = serve.Space(f"{W}", jd, q.as_lohi(b), W)
    H = q.build_hamiltonian_g(jd, [-1.0, -1.0, -1.0, -1.0])
    r, c, _ = orc.sparse_triplets(orc.Basis(b), H)
    off = r != c
    f = np.bincount(r[off], minlength=b.shape[0])
    assert sp.padded_nnz() == int(((f + 3) // 4 * 4).sum())
    assert f.max() > W // 4 or W == 32  # rows above n_spins / 4 exist at 48 spins
