"""A GPU solve written to the reference's .eigs.h5 layout and read back
(src/diag_DOQSI_ham.cpp:149-173): the golden 48-spin case (Jpm = 0.02,
B = (0.1, 0.1, 0.1), test/test_correct.sh) solved on the device, written by
h5io.write_eigs, read back, and checked against the golden arrays: the
datasets and their dims, the eigenvalue to 1e-9 and the eigenvector to the
h5diff tolerance of test/test_correct.sh (1e-4 per entry), up to sign."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_basis
import qsi_subspace_ed_amd as q
from qsi_subspace_ed_amd import h5io

pytestmark = pytest.mark.gpu

try:
    h5io._lib()
    HAVE_H5 = True
except RuntimeError:
    HAVE_H5 = False


@pytest.mark.skipif(not HAVE_H5, reason="no HDF5 C library in this image")
def test_gpu_solve_eigs_h5_round_trip(lattices, cuda, tmp_path):
    gold = np.load(os.path.join(GOLDEN, "eigs48_jpm0.02_B0.1.npz"))
    jd = lattices[48]
    b = load_basis(48)
    D = b.shape[0]
    H = q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    op = q.GpuOp(q.GpuBasis(b, 48), H)
    evals, vecs, nconv, _ = op.eigsh(nev=2, ncv=20, tol=1e-12, seed=11, n_evecs=2)
    assert nconv == 2
    V = vecs.cpu().numpy()          # (D, 2)
    p = tmp_path / "pyro_2,0,0_0,2,0_0,0,3.0.eigs.h5"
    h5io.write_eigs(p, evals, V, q.lattice_path("pyro_2,0,0_0,2,0_0,0,3.json"))
    ev, V2, lat, dn = h5io.read_eigs(p)
    assert np.array_equal(ev, np.asarray(evals)) and np.array_equal(V2, V)
    assert lat == "pyro_2,0,0_0,2,0_0,0,3.json" and dn == "basis"
    raw = h5io.read_dataset(p, "eigenvectors")
    assert raw.shape == (D, 2)                       # dims (D, k) ...
    assert np.array_equal(raw.reshape(-1)[:D], V[:, 0])  # ... over the column-major buffer
    g0 = gold["eigenvalues"][0]
    assert abs(ev[0] - g0) <= 1e-9 * abs(g0)
    gv = gold["eigenvector"] / np.linalg.norm(gold["eigenvector"])
    v = V2[:, 0]
    assert np.max(np.abs(v - np.sign(v @ gv) * gv)) < 1e-4
