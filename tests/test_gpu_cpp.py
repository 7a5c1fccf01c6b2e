"""The C++ face (include/qsi_gpu.hpp) on the GPU: tests/cpp/wrapper_check
(built by build()) runs qsi::GpuOp::perform_op / evaluate_device (also on the low-key kernels),
qsi::apply_fn, qsi::eigsh and qsi::FtlmComputer on the 32-spin fixture; its outputs must equal the Python
face's (same kernels, same bits) and the dense spectrum."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import qsi_subspace_ed_amd as q
from conftest import ROOT, load_basis
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "wrapper_check")


def _read(d, name):
    return np.fromfile(os.path.join(d, name), dtype=np.float64)


def test_cpp_face(lattices, cuda, tmp_path):
    torch = cuda
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/wrapper_check is not built (run __graft_entry__.build())")
    jd = lattices[32]
    b = load_basis(32)
    D = b.shape[0]
    H = q.build_hamiltonian_g(jd, [-1.0, -0.8, -0.6, -0.9])
    q.add_local_field(H, jd, [0.1, 0.0, -0.05])
    arr, n = H.terms()
    lohi = np.ascontiguousarray(q.device.as_lohi(b), dtype=np.uint64)
    x = np.random.default_rng(3).standard_normal(D)
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{D} 32 {n}\n")
    lohi.tofile(os.path.join(d, "basis.bin"))
    with open(os.path.join(d, "terms.bin"), "wb") as f:
        f.write(bytes(arr))
    x.tofile(os.path.join(d, "x.bin"))
    print(f"[cpp] wrapper_check D={D} terms={n}", flush=True)
    r = subprocess.run([EXE, d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    assert "error ok" in r.stdout, r.stdout
    assert "one-state ok" in r.stdout, r.stdout
    # ring flips and a field on the full 32-spin sector: no compare needed
    assert "lowkeys ok 2" in r.stdout, r.stdout

    op = q.GpuOp(q.GpuBasis(b, 32), H)
    xd = torch.from_numpy(x).cuda()
    y = torch.zeros(D, dtype=torch.float64, device="cuda")
    op.evaluate(xd, y)
    y = y.cpu().numpy()
    assert np.array_equal(_read(d, "y_perform.bin"), y)
    assert np.array_equal(_read(d, "y_device.bin"), y)
    y1 = torch.ones(D, dtype=torch.float64, device="cuda")
    op.evaluate_add(xd, y1)
    assert np.array_equal(_read(d, "y_applyfn.bin"), y1.cpu().numpy())
    ref = orc.evaluate_add(orc.Basis(b), H, x, np.zeros(D))
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)

    ev = op.eigsh(nev=5, ncv=15, tol=1e-10, seed=0)[0]
    assert np.array_equal(_read(d, "evals.bin"), np.asarray(ev))

    rr, cc, vv = orc.sparse_triplets(orc.Basis(b), H)
    Md = np.zeros((D, D))
    np.add.at(Md, (rr, cc), vv)
    w = np.linalg.eigvalsh(Md)
    assert np.allclose(_read(d, "evals.bin"), w[:5], rtol=1e-9, atol=1e-12)

    out = _read(d, "ftlm_exact.bin")
    nT = 6
    T, Z, E, Cv, e_ref = out[:nT], out[nT:2 * nT], out[2 * nT:3 * nT], out[3 * nT:4 * nT], out[-1]
    assert np.allclose(1.0 / T, np.linspace(1 / 5.0, 1 / 0.5, nT), rtol=1e-14)
    assert abs(e_ref - w[0]) <= 1e-9 * abs(w[0])
    bz = np.exp(-np.outer(1.0 / T, w - e_ref))
    Zx = bz.sum(axis=1)
    Ex = bz @ w / Zx
    Cx = (bz @ (w * w) / Zx - Ex ** 2) / T ** 2
    assert np.allclose(Z, Zx, rtol=1e-10)
    assert np.allclose(E, Ex, rtol=1e-10, atol=1e-12)
    assert np.allclose(Cv, Cx, rtol=1e-8, atol=1e-10)

    g = _read(d, "ftlm_rng.bin").reshape(2, 2, nT)  # (one at a time, block) x (Z, E)
    assert np.allclose(g[0], g[1], rtol=1e-12)


FTLM_EXE = os.path.join(ROOT, "tests", "cpp", "ftlm_check")


def test_cpp_ftlm_operator_sums(lattices, cuda, tmp_path):
    """FtlmComputer(ops, H, T_min, T_max, n_Ts) with the complete set of unit
    start vectors: write_to_h5's "<name>_sum" equals
    Tr(exp(-beta (H - e_ref)) O) and "<name>_sum_2" equals
    sum_i ((exp(-beta H / 2) O exp(-beta H / 2))_ii exp(beta e_ref))^2, both from
    the dense spectrum of the oracle's H; the norm sum is the shifted Z.
    Operators: one ring term O_h, and O_h + O_h^dagger of another ring."""
    from qsi_subspace_ed_amd import h5io
    if not os.path.exists(FTLM_EXE):
        pytest.skip("tests/cpp/ftlm_check is not built (no HDF5 headers at build time)")
    jd = lattices[32]
    b = load_basis(32)
    D = b.shape[0]
    H = q.build_hamiltonian_g(jd, [-1.0, -0.8, -0.6, -0.9])
    q.add_local_field(H, jd, [0.1, 0.0, -0.05])
    ring, ring_h, _ = q.get_ring_ops(jd)
    O0 = q.SymbolicOpSum()
    O0.add_term(1.0, ring[0])
    O1 = q.SymbolicOpSum()
    O1.add_term(1.0, ring[5])
    O1.add_term(1.0, ring_h[5])
    arr, n = H.terms()
    oarr = [o.terms() for o in (O0, O1)]
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{D} 32 {n} 2 {oarr[0][1]} {oarr[1][1]}\n")
    np.ascontiguousarray(q.device.as_lohi(b), dtype=np.uint64).tofile(os.path.join(d, "basis.bin"))
    with open(os.path.join(d, "terms.bin"), "wb") as f:
        f.write(bytes(arr))
    with open(os.path.join(d, "ops.bin"), "wb") as f:
        for a, k in oarr:
            f.write(bytes(a)[: k * C.sizeof(a._type_)])
    print(f"[cpp] ftlm_check D={D}", flush=True)
    r = subprocess.run([FTLM_EXE, d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout
    out = np.fromfile(os.path.join(d, "ftlm.bin"))
    nT = 6
    T, norm, e_ref = out[:nT], out[nT:2 * nT], out[-1]
    ob = orc.Basis(b)

    def dense(Hs):
        rr, cc, vv = orc.sparse_triplets(ob, Hs)
        M = np.zeros((D, D))
        np.add.at(M, (rr, cc), vv)
        return M
    w, U = np.linalg.eigh(dense(H))
    assert abs(e_ref - w[0]) <= 1e-9 * abs(w[0])
    for l, Od in enumerate((dense(O0), dense(O1))):
        s1 = h5io.read_dataset(os.path.join(d, "ftlm.h5"), f"op{l}_sum")
        s2 = h5io.read_dataset(os.path.join(d, "ftlm.h5"), f"op{l}_sum_2")
        OU = U.T @ Od @ U
        for j, t in enumerate(T):
            half = U @ np.diag(np.exp(-0.5 * (w - e_ref) / t)) @ U.T
            A = half @ Od @ half
            assert abs(s1[j] - np.trace(A)) <= 1e-9 * max(1.0, abs(np.trace(A))), (l, j)
            assert abs(s1[j] - np.sum(np.exp(-(w - e_ref) / t) * np.diag(OU))) <= \
                1e-9 * max(1.0, abs(s1[j]))
            want2 = np.sum(np.diag(A) ** 2)
            assert abs(s2[j] - want2) <= 1e-9 * max(1.0, want2), (l, j)
    Zx = np.array([np.sum(np.exp(-(w - e_ref) / t)) for t in T])
    assert np.allclose(norm, Zx, rtol=1e-10)
