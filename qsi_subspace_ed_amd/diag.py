"""The reference's algorithm dispatch for real Hamiltonians,
diagonalise_real (include/matrix_diag_bits.hpp:120-204), over the device
library.  Options mirror diag_DOQSI_ham's (src/diag_DOQSI_ham.cpp:48-72) and
provide_lanczos_options (include/lanczos_cli.hpp:5-39), defaults included.

  dense     (auto when D < 100) H materialised from the toSparseMatrix
            triplets (qsi_op_connectivity; operator_matrix.hpp:79-105) as a
            dense FP64 matrix on the device and fully diagonalised
            (compute_eigenspectrum_dense, matrix_diag_bits.hpp:98-117) with
            torch.linalg.eigh.
  sparse    (auto when D < 10^7) the connectivity kept on the device (ranks and
            coefficient classes: the role of the materialised sparse matrix)
            under the thick-restart Lanczos solver
            (compute_spectrum_iterative<SparseSymMatProd>, :51-95).
  mfsparse  (auto otherwise) the same solver over the matrix-free search apply
            (compute_spectrum_iterative<LazyOpSumProd>).
  mfeig0    lanczos::eigval0 with parse_lanczos_settings' settings
            (lanczos_cli.hpp:28-39): krylov_dim = ncv, min_iterations 10,
            eigenvector on, x0_seed = rng_seed; n_eigvals = n_eigvecs = 1.

Sizing as the reference: ncv = min(ncv, D), n_eigvals = min(ncv - 1,
n_eigvals), n_eigvecs = min(n_eigvecs, n_eigvals); tol = 10^atol.  Errors
are the reference's: an unknown algorithm, mfeig0 with more than one
eigenvalue or eigenvector, or an iterative solve that does not converge
raise RuntimeError.
"""
import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check
from .device import GpuOp, Settings

ALGORITHMS = ("dense", "sparse", "mfsparse", "mfeig0")


@dataclass
class Options:
    n_eigvals: int = 1
    n_eigvecs: int = 1
    ncv: int = 15
    atol: int = -8        # exponent: tolerance 10^atol
    rtol: int = -8        # exponent (mfeig0)
    rng_seed: int = 0
    max_iters: int = 1000
    verify: bool = False  # mfeig0: verify the eigenvector
    algorithm: Optional[str] = None


def choose_algorithm(dim):
    """The reference's default by basis size (matrix_diag_bits.hpp:126-136)."""
    if dim < 100:
        return "dense"
    if dim < 10_000_000:
        return "sparse"
    return "mfsparse"


def dense_matrix(op: GpuOp):
    """H as a dense (D, D) CUDA float64 tensor from the device triplets."""
    import torch
    D = op.rows()
    lib = _lib.load()
    torch.cuda.synchronize()
    nnz = C.c_int64(0)
    check(lib.qsi_op_connectivity(op.handle, C.byref(nnz), None, None, None, None),
          "qsi_op_connectivity")
    n = max(1, nnz.value)
    r = torch.empty(n, dtype=torch.int64, device="cuda")
    c = torch.empty(n, dtype=torch.int64, device="cuda")
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    d = torch.empty(max(1, D), dtype=torch.float64, device="cuda")
    check(lib.qsi_op_connectivity(op.handle, C.byref(nnz), r.data_ptr(), c.data_ptr(),
                                  v.data_ptr(), d.data_ptr()), "qsi_op_connectivity")
    k = nnz.value
    # duplicate (row, col) pairs (two terms reaching the same state) are
    # summed by coalescing (sorted, a fixed order), not by atomic adds
    off = torch.sparse_coo_tensor(torch.stack((r[:k], c[:k])), v[:k], (D, D)).coalesce()
    return off.to_dense() + torch.diag(d[:D])


def diagonalise_real(op: GpuOp, opts: Options = Options()):
    """(eigenvalues ascending as numpy, eigenvectors as a (D, n) CUDA float64
    tensor) by the algorithm opts.algorithm, or the reference's default for
    the basis size.  The matrix-free solve (mfsparse) builds the basis's low
    keys if it has none (4 bytes per state, kept on the basis) and checks them
    against the operator once (about one apply), so its applies run on the
    low-key kernel; a basis whose low keys cannot be built keeps the lane
    kernel.  Every kernel gives the same bits, so the same eigenpairs."""
    import torch
    D = op.rows()
    algo = opts.algorithm or choose_algorithm(D)
    if algo not in ALGORITHMS:
        raise RuntimeError("Unrecognised algorithm name (this is a bug)")
    if algo == "mfeig0":
        if opts.n_eigvals != 1:
            raise RuntimeError("n_eigvals must be 1 for mf-large")
        if opts.n_eigvecs != 1:
            raise RuntimeError("n_eigvecs must be 1 for mf-large")
        st = Settings(krylov_dim=opts.ncv, abs_tol=10.0 ** opts.atol,
                      rel_tol=10.0 ** opts.rtol, min_iterations=10,
                      x0_seed=opts.rng_seed, calc_eigenvector=True,
                      verify_eigenvector=opts.verify)
        e0, vec, _ = op.lanczos_eigval0(st)
        return np.array([e0]), vec.reshape(D, 1)
    if algo == "dense":
        n_eigvals = min(D, opts.n_eigvals)
        w, V = torch.linalg.eigh(dense_matrix(op))
        return w[:n_eigvals].cpu().numpy(), V[:, :n_eigvals]
    ncv = min(opts.ncv, D)
    n_eigvals = min(ncv - 1, opts.n_eigvals)
    n_eigvecs = min(opts.n_eigvecs, n_eigvals)
    lib = _lib.load()
    saved = lib.qsi_op_get_variant(op.handle)
    if algo == "sparse":
        op.cache_connectivity(True)
        op.set_variant("cached")
    elif saved in (_lib.QSI_APPLY_AUTO, _lib.QSI_APPLY_CACHED):
        # matrix-free even when the operator holds a connectivity cache: the
        # low-key kernel, eight lookups per lane when candidates need no
        # compare (check mode 2), else four; the lane kernel without low keys
        variant = "lane4"
        try:
            if not op.basis.has_lowkeys():
                op.basis.build_lowkeys()
            variant = "low8" if op.verify_lowkeys() == 2 else "low4"
        except _lib.QsiError:
            pass
        op.set_variant(variant)
    try:
        evals, vecs, nconv, _ = op.eigsh(nev=n_eigvals, ncv=ncv, tol=10.0 ** opts.atol,
                                         max_restarts=opts.max_iters, seed=opts.rng_seed,
                                         n_evecs=n_eigvecs)
    except _lib.QsiError as e:
        raise RuntimeError("Eigenvalue decomposition failed") from e
    finally:
        check(lib.qsi_op_set_variant(op.handle, saved), "qsi_op_set_variant")
    return evals[:nconv], vecs
