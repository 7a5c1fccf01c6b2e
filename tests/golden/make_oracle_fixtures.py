"""Fixtures computed by the oracle (the CPU restatement of the reference) for
cases too slow to rerun inside a GPU test.  Output: tests/golden/oracle48.json.

mfeig0 on the golden 48-spin case (test/test_correct.sh:8 settings): pass 1 of
the reference's lanczos_iterate always runs to max_iterations (5000), pass 2
stops at the first converged check.  Recorded: E0, pass-2 stop iteration, and
the pass-2 eigenvector's norm, overlap with and max deviation from the
reference's golden eigenvector (both normalised) — enough to pin the GPU
replica to the oracle without shipping the vector.

Usage: python tests/golden/make_oracle_fixtures.py   (about 3 minutes)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import load_basis  # noqa: E402
from oracle import pyoracle as orc  # noqa: E402
import qsi_subspace_ed_amd as q  # noqa: E402


def main():
    jd = q.load_lattice(q.lattice_path("pyro_2,0,0_0,2,0_0,0,3.json"))
    b = load_basis(48)
    H = q.build_hamiltonian_jpm(jd, 0.02, [0.1, 0.1, 0.1])
    st = orc.settings(20, 1e-9, 1e-9, 10, 5000, 11, True, False)
    e0, vec, res, a, bb = orc.lanczos_eigval0(orc.Basis(b), H, st, cap=5000)
    gold = np.load(os.path.join(HERE, "eigs48_jpm0.02_B0.1.npz"))["eigenvector"]
    gv = gold / np.linalg.norm(gold)
    v = vec / np.linalg.norm(vec)
    s = np.sign(v @ gv)
    out = {
        "case": "48-spin Jpm=0.02 B=(0.1,0.1,0.1), mfeig0 -k 20 --rng_seed 11 --atol -9 --rtol -9",
        "e0": e0,
        "pass1_steps": int(len(a)),
        "pass2_n_iterations": int(res.n_iterations),
        "eigvec_norm": float(np.linalg.norm(vec)),
        "overlap_with_golden": float(abs(v @ gv)),
        "maxdiff_vs_golden": float(np.max(np.abs(v - s * gv))),
        "alpha_0": float(a[0]), "beta_0": float(bb[0]),
        "alpha_4999": float(a[-1]), "beta_4999": float(bb[-1]),
    }
    with open(os.path.join(HERE, "oracle48.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
