import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def load_basis(W):
    """Reference fixture basis for W in {8, 32, 48} as (D, 2) uint64."""
    if W == 48:
        return np.load(os.path.join(GOLDEN, "basis_48.npz"))["basis"]
    return np.load(os.path.join(GOLDEN, f"basis_{W}.npy"))


LATTICE_OF = {
    8: "pyro_1,0,0_0,1,0_0,0,2.json",
    32: "pyro_2,0,0_0,2,0_0,0,2.json",
    48: "pyro_2,0,0_0,2,0_0,0,3.json",
    64: "pyro_2,0,0_0,2,0_0,0,4.json",
}


@pytest.fixture(scope="session")
def lattices():
    import qsi_subspace_ed_amd as q
    return {W: q.load_lattice(q.lattice_path(n)) for W, n in LATTICE_OF.items()}


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch
