"""Edge cases at the boundary: empty and one-state bases, inputs below the
apply tolerance (APPLY_TOL, components/operator/include/operator_matrix.hpp:
113), and the reference's errors for an empty vector
(components/linalg_routines/src/lanczos.cpp: "v called with size 0")."""
import numpy as np
import pytest

import qsi_subspace_ed_amd as q
from conftest import load_basis
from oracle import pyoracle as orc
from qsi_subspace_ed_amd._lib import QsiError

pytestmark = pytest.mark.gpu


def _ham(jd):
    H = q.build_hamiltonian_g(jd, [-1.0, -0.7, -0.4, -0.9])
    q.add_local_field(H, jd, [0.05, -0.03, 0.08])
    return H


def test_empty_basis(lattices, cuda):
    torch = cuda
    basis = q.GpuBasis(np.zeros((0, 2), dtype=np.uint64), 8)
    assert basis.dim() == 0
    op = q.GpuOp(basis, _ham(lattices[8]))
    x = torch.zeros(0, dtype=torch.float64, device="cuda")
    y = torch.zeros(0, dtype=torch.float64, device="cuda")
    op.evaluate(x, y)
    op.evaluate_add(x, y)
    with pytest.raises(QsiError, match="empty basis"):
        op.eigsh(nev=1, ncv=2)
    with pytest.raises(QsiError, match="size 0"):
        op.lanczos_eigval0(q.Settings())


def test_one_state_basis(lattices, cuda):
    """Every flip leaves the basis, so H is the state's diagonal entry."""
    torch = cuda
    jd = lattices[8]
    one = load_basis(8)[5:6]
    H = _ham(jd)
    op = q.GpuOp(q.GpuBasis(one, 8), H)
    x = torch.tensor([0.7], dtype=torch.float64, device="cuda")
    y = torch.zeros(1, dtype=torch.float64, device="cuda")
    op.evaluate(x, y)
    ref = orc.evaluate_add(orc.Basis(one), H, np.array([0.7]), np.zeros(1))
    assert abs(y.item() - ref[0]) <= 1e-13 * max(1.0, abs(ref[0]))
    evals, _, nconv, _ = op.eigsh(nev=1, ncv=15)
    assert nconv == 1
    assert abs(evals[0] - ref[0] / 0.7) <= 1e-12 * max(1.0, abs(ref[0]))


@pytest.mark.parametrize("variant", ["lane4", "window", "cached"])
def test_inputs_below_apply_tolerance(variant, lattices, cuda):
    """x at 1e-16 except 50 entries: contributions from |x_j| <= 1e-15 are
    skipped, so most rows are their diagonal term alone; entry by entry the
    GPU matches the oracle's push apply (which skips the same sources) to
    1e-12 relative plus 1e-15 |x_i| + 1e-28 (the rounding left when a row's
    diagonal terms cancel) — a kept contribution would add at least
    1e-16 x 0.4."""
    torch = cuda
    jd = lattices[32]
    b = load_basis(32)
    D = b.shape[0]
    H = _ham(jd)
    op = q.GpuOp(q.GpuBasis(b, 32), H)
    if variant == "cached":
        op.cache_connectivity(True)
    op.set_variant(variant)
    rng = np.random.default_rng(17)
    x = 1e-16 * np.where(rng.random(D) < 0.5, -1.0, 1.0)
    x[rng.choice(D, 50, replace=False)] = rng.standard_normal(50)
    y = torch.zeros(D, dtype=torch.float64, device="cuda")
    op.evaluate(torch.from_numpy(x).cuda(), y)
    y = y.cpu().numpy()
    ref = orc.evaluate_add(orc.Basis(b), H, x, np.zeros(D))
    assert np.all(np.abs(y - ref) <= 1e-12 * np.abs(ref) + 1e-15 * np.abs(x) + 1e-28)


def test_lowkeys_limits(cuda):
    """Low keys hold at most 32 informative positions below the bucket key.
    States 2^k: neighbours first differ at bit k + 1, so every bit from 1 up
    is informative.  With a 1-bit key, 2^0..2^33 leave 32 positions below it
    (accepted: the LOW kernels give the lane kernel's bits for S+_i S-_j
    flips, each target 2^j -> 2^i a basis state); 2^0..2^63 leave 62
    (refused with an error, the basis unchanged)."""
    torch = cuda
    H = q.SymbolicOpSum()
    for i, j in ((33, 0), (5, 17), (20, 21), (0, 33), (31, 2)):
        H.add_term(0.5 + 0.01 * i, q.SymbolicPMROperator.from_ops("+-", [i, j]))
    st = np.array([[1 << k, 0] for k in range(34)], dtype=np.uint64)
    basis = q.GpuBasis(st, 34, bucket_bits=1)
    basis.build_lowkeys()
    assert basis.has_lowkeys()
    op = q.GpuOp(basis, H)
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(34)).cuda()
    ys = []
    for v in ("lane4", "low2", "low4", "low8"):
        op.set_variant(v)
        y = torch.zeros(34, dtype=torch.float64, device="cuda")
        op.evaluate(x, y)
        ys.append(y)
    ref = orc.evaluate_add(orc.Basis(st), H, x.cpu().numpy(), np.zeros(34))
    assert np.linalg.norm(ys[0].cpu().numpy() - ref) <= 1e-13 * np.linalg.norm(ref)
    assert np.count_nonzero(ref) == 5
    for y in ys[1:]:
        assert torch.equal(ys[0], y)
    wide = np.array([[1 << k, 0] for k in range(64)], dtype=np.uint64)
    b64 = q.GpuBasis(wide, 64, bucket_bits=1)
    with pytest.raises(QsiError, match="more than 32 informative"):
        b64.build_lowkeys()
    assert not b64.has_lowkeys()


def test_lowkeys_verification_finds_a_clash(cuda):
    """Basis {00, 11}: only bit 1 is informative, so with a 1-bit key the
    state 01 (the pre-image of row 11 under S+ on spin 1) has the same key and
    low key as 00.  The verification pass must report the clash, and the
    low-key kernels must then still compare states: y = 0, as the oracle's
    push (which drops the target 10, absent from the basis) gives."""
    torch = cuda
    st = np.array([[0, 0], [3, 0]], dtype=np.uint64)
    H = q.SymbolicOpSum()
    H.add_term(1.0, q.SymbolicPMROperator.from_ops("+", [1]))
    basis = q.GpuBasis(st, 2, bucket_bits=1)
    basis.build_lowkeys()
    op = q.GpuOp(basis, H)
    assert op.verify_lowkeys() == 0
    x = np.array([0.5, 0.25])
    ref = orc.evaluate_add(orc.Basis(st), H, x, np.zeros(2))
    assert np.all(ref == 0.0)
    for v in ("low2", "low4", "low8", "lane4"):
        op.set_variant(v)
        y = torch.full((2,), 9.0, dtype=torch.float64, device="cuda")
        op.evaluate(torch.from_numpy(x).cuda(), y)
        assert np.array_equal(y.cpu().numpy(), ref), v


def test_lowkeys_verification_modes(lattices, cuda):
    """The 32-spin basis with every third state dropped: ring flips now reach
    states outside the basis, so the pass cannot return 2; whatever mode it
    returns, the low-key kernels agree with the oracle and with LANE4 bit for
    bit, and the full basis under the same terms gives 2."""
    torch = cuda
    jd = lattices[32]
    H = _ham(jd)
    for b, full in ((load_basis(32), True), (load_basis(32)[::3].copy(), False)):
        basis = q.GpuBasis(b, 32)
        basis.build_lowkeys()
        op = q.GpuOp(basis, H)
        mode = op.verify_lowkeys()
        assert mode == 2 if full else mode in (0, 1)
        x = np.random.default_rng(8).standard_normal(b.shape[0])
        ref = orc.evaluate_add(orc.Basis(b), H, x, np.zeros(b.shape[0]))
        ys = []
        for v in ("lane4", "low2", "low4", "low8"):
            op.set_variant(v)
            y = torch.zeros(b.shape[0], dtype=torch.float64, device="cuda")
            op.evaluate(torch.from_numpy(x).cuda(), y)
            ys.append(y)
        assert np.linalg.norm(ys[0].cpu().numpy() - ref) <= 1e-13 * np.linalg.norm(ref)
        for y in ys[1:]:
            assert torch.equal(ys[0], y)


def test_widest_bucket_key(cuda):
    """A 32-bit bucket key (the widest; its table has 2^32 + 1 entries, more
    than one dispatch of one thread per entry can hold, so the table is built
    grid-stride) on the 64-spin Q=0 basis, whose states differ at 33
    informative positions: every state is found at its rank, and the apply
    gives the same bits as with the automatic key."""
    torch = cuda
    free, _ = torch.cuda.mem_get_info()
    if free < 40e9:
        pytest.skip("a 32-bit bucket table needs 16 GiB of device memory")
    jd = q.load_lattice(q.lattice_path("pyro_2,0,0_0,2,0_0,0,4.json"))
    H = q.build_hamiltonian_g(jd, [-1.0, -0.55, -0.3, -0.8])
    q.add_local_field(H, jd, [0.07, -0.02, 0.05])
    ref = q.enumerate_basis_device(jd, 0)
    states = ref.states()
    D = states.shape[0]
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(D)).cuda()
    ref.build_lowkeys()
    op_ref = q.GpuOp(ref, H)
    op_ref.set_variant("lane4")
    y_ref = torch.zeros_like(x)
    op_ref.evaluate(x, y_ref)
    wide = q.GpuBasis(states, 64, bucket_bits=32)
    assert wide.bucket_bits() == 32
    assert sum(w for _, w in wide.key_fields()) == 32
    assert np.array_equal(wide.search(states), np.arange(D))
    wide.build_lowkeys()
    op = q.GpuOp(wide, H)
    assert op.verify_lowkeys() == 2
    for v in ("lane4", "low4", "low8"):
        op.set_variant(v)
        y = torch.zeros_like(x)
        op.evaluate(x, y)
        assert torch.equal(y, y_ref), v
    op.close()
    wide.close()
    op_ref.close()
    ref.close()
    torch.cuda.empty_cache()
    with pytest.raises(QsiError):
        q.GpuBasis(states[:10], 64, bucket_bits=33)
