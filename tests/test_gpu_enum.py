"""Device enumeration against the host enumerator (which is pinned to the
reference's basis fixtures in test_enumerate.py): identical sorted states
for 8/32/48/64 spins and spinon sectors n = 0, 1, 2, and a device basis that
searches and applies like one created from host states."""
import numpy as np
import pytest

from conftest import load_basis
import qsi_subspace_ed_amd as q

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,n", [(8, 0), (32, 0), (32, 1), (32, 2), (48, 0), (48, 1), (64, 0)])
def test_device_enumeration_matches_host(W, n, lattices, cuda):
    jd = lattices[W]
    host = q.enumerate_basis(jd, n, n_threads=8)
    dev = q.enumerate_basis_device(jd, n)
    assert dev.dim() == host.shape[0]
    assert np.array_equal(dev.states(), host)


def test_device_basis_fixture_and_apply(lattices, cuda):
    torch = cuda
    jd = lattices[48]
    b = load_basis(48)
    dev = q.enumerate_basis_device(jd, 0)
    assert np.array_equal(dev.states(), b)
    H = q.build_hamiltonian_g(jd, [-0.7, -0.3, -0.9, -0.5])
    ref = q.GpuBasis(b, 48)
    x = torch.from_numpy(np.random.default_rng(5).standard_normal(b.shape[0])).cuda()
    y0 = torch.zeros_like(x)
    y1 = torch.zeros_like(x)
    q.GpuOp(ref, H).evaluate(x, y0)
    q.GpuOp(dev, H).evaluate(x, y1)
    assert torch.equal(y0, y1)
    assert np.array_equal(dev.search(b[::7]), np.arange(0, b.shape[0], 7))
