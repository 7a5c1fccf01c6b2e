"""CPU eigenvalues of the 64-spin bench workload (diag(2,2,4), Q=0, seeded
g_s ~ U[0.2, 1] and |B| = 0.1; bench.build_workload(1234)), the check the
single-vector GPU solve is held to: the oracle (the CPU restatement of the
reference) materialises H as triplets (toSparseMatrix,
components/operator/include/operator_matrix.hpp:79-105) and scipy's ARPACK
implicitly restarted Lanczos (the method family of the reference's Spectra
SymEigsSolver, include/matrix_diag_bits.hpp:51-95) finds the 12 lowest
eigenvalues to tol 1e-13.  Output: tests/golden/eigs64_bench.json.

Usage: python tests/golden/make_eigs64.py   (about a minute on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    import bench
    from oracle import pyoracle as orc

    t0 = time.perf_counter()
    jd, states, H, g, B, _ = bench.build_workload(1234, n_threads=os.cpu_count())
    D = states.shape[0]
    r, c, v = orc.sparse_triplets(orc.Basis(states), H)
    M = sp.csr_matrix((v, (r, c)), shape=(D, D))
    del r, c, v
    t1 = time.perf_counter()
    w = np.sort(sla.eigsh(M, k=12, which="SA", tol=1e-13, ncv=40,
                          v0=np.random.default_rng(0).standard_normal(D))[0])
    t2 = time.perf_counter()
    out = {
        "workload": "pyro diag(2,2,4) 64 spins, Q=0; bench.build_workload(1234)",
        "g": [float(x) for x in g],
        "B": [float(x) for x in B],
        "dim": int(D),
        "nnz": int(M.nnz),
        "method": "oracle sparse triplets + scipy.sparse.linalg.eigsh(k=12, which='SA', "
                  "tol=1e-13, ncv=40)",
        "evals": [float(x) for x in w],
        "seconds": {"build": round(t1 - t0, 1), "eigsh": round(t2 - t1, 1)},
    }
    path = os.path.join(HERE, "eigs64_bench.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
