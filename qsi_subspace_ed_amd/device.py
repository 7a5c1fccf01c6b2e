"""Python face of the GPU library: device basis, the LazyOpSum-shaped operator
and the Krylov solvers, all calling the C ABI (include/qsi_gpu.h).

Device vectors are torch tensors on `cuda:0` (float64); torch is only used for
device memory and the current stream.  The reference's interfaces mirrored:
  LazyOpSum  evaluate / evaluate_add / rows / cols / toSparseMatrix
             (components/operator/include/operator_matrix.hpp:10-118)
  LazyOpSumProd::perform_op      (operator_matrix.hpp:120-139)
  lanczos::eigval0 + Settings    (components/linalg_routines/include/lanczos.hpp)
  compute_spectrum_iterative     (include/matrix_diag_bits.hpp:51-95)
"""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _require_cuda_f64(t, n=None, name="vector"):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
            and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous float64 CUDA tensor")
    if n is not None and t.numel() != n:
        raise ValueError(f"{name} has {t.numel()} entries, expected {n}")


def as_lohi(states):
    """(D, 2) uint64 [lo, hi] array from an array of states (uint64 (D,) or
    (D, 2), or a list of Python ints)."""
    a = np.asarray(states)
    if a.dtype == object or (a.ndim == 1 and a.dtype != np.uint64):
        v = [int(x) for x in states]
        out = np.empty((len(v), 2), dtype=np.uint64)
        out[:, 0] = [x & ((1 << 64) - 1) for x in v]
        out[:, 1] = [x >> 64 for x in v]
        return out
    if a.ndim == 1:
        out = np.zeros((a.shape[0], 2), dtype=np.uint64)
        out[:, 0] = a
        return out
    return np.ascontiguousarray(a, dtype=np.uint64)


class GpuBasis:
    """Sorted Z-basis resident in GPU memory with its bucket table (replaces
    ZBasisInterp, components/operator/include/operator.hpp:197-204)."""

    @classmethod
    def _adopt(cls, handle, n_spins):
        """Wrap a basis handle built on the device (states fetched lazily)."""
        self = cls.__new__(cls)
        self._h = handle
        self.n_spins = int(n_spins)
        self._states_host = None
        return self

    def __init__(self, states, n_spins, bucket_bits=0, hash_index=False):
        lib = _lib.load()
        lohi = as_lohi(states)
        self._states_host = lohi
        h = C.c_void_p()
        check(lib.qsi_basis_create_ex(lohi.ctypes.data_as(C.c_void_p), lohi.shape[0],
                                      int(n_spins), int(bucket_bits), int(bool(hash_index)),
                                      C.byref(h)),
              "qsi_basis_create")
        self._h = h
        self.n_spins = int(n_spins)

    @property
    def handle(self):
        return self._h

    def dim(self):
        return int(_lib.load().qsi_basis_dim(self._h))

    def bucket_bits(self):
        return int(_lib.load().qsi_basis_bucket_bits(self._h))

    def build_slots(self, enable=True):
        """Build (or drop) the slot index under the bucket table
        (qsi_basis_build_slots): lookups by hashed slots per bucket instead of
        bisection; the QSI_APPLY_SLOT* kernels and search() use it."""
        check(_lib.load().qsi_basis_build_slots(self._h, int(bool(enable))),
              "qsi_basis_build_slots")

    def has_slots(self):
        return bool(_lib.load().qsi_basis_has_slots(self._h))

    def build_lowkeys(self, enable=True):
        """Build (or drop) the low keys (qsi_basis_build_lowkeys): 4 bytes
        per state that the QSI_APPLY_LOW* kernels bisect a bucket on."""
        check(_lib.load().qsi_basis_build_lowkeys(self._h, int(bool(enable))),
              "qsi_basis_build_lowkeys")

    def has_lowkeys(self):
        return bool(_lib.load().qsi_basis_has_lowkeys(self._h))

    def key_fields(self):
        """Bucket key fields [(lo, width), ...], highest first."""
        n = C.c_int(0)
        lo = (C.c_int * 4)()
        w = (C.c_int * 4)()
        check(_lib.load().qsi_basis_key_fields(self._h, C.byref(n), lo, w), "qsi_basis_key_fields")
        return [(lo[i], w[i]) for i in range(n.value)]

    def states(self):
        """(D, 2) uint64 [lo, hi] host copy of the states."""
        if self._states_host is None:
            out = np.empty((self.dim(), 2), dtype=np.uint64)
            check(_lib.load().qsi_basis_download(self._h, out.ctypes.data_as(C.c_void_p)),
                  "qsi_basis_download")
            self._states_host = out
        return self._states_host

    def search(self, queries_lohi):
        """Device lookup: int64 ranks (-1 where absent) of (n, 2) uint64 states
        given as a CUDA int64 tensor view or numpy array."""
        torch = _torch()
        q = as_lohi(queries_lohi)
        dq = torch.from_numpy(q.view(np.int64)).cuda()
        out = torch.empty(q.shape[0], dtype=torch.int64, device="cuda")
        check(_lib.load().qsi_basis_search(self._h, _ptr(dq), q.shape[0], _ptr(out), _stream()),
              "qsi_basis_search")
        return out.cpu().numpy()

    def close(self):
        if self._h:
            _lib.load().qsi_basis_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Settings:
    """lanczos::Settings (components/linalg_routines/include/lanczos.hpp:9-20)."""
    krylov_dim: int = 10
    abs_tol: float = 1e-8
    rel_tol: float = 1e-8
    min_iterations: int = 30
    max_iterations: int = 5000
    convergence_check_interval: int = 5
    verbosity: int = 0
    x0_seed: int = 10
    calc_eigenvector: bool = False
    verify_eigenvector: bool = False

    def to_c(self):
        s = _lib.LanczosSettings()
        s.krylov_dim = self.krylov_dim
        s.abs_tol = self.abs_tol
        s.rel_tol = self.rel_tol
        s.min_iterations = self.min_iterations
        s.max_iterations = self.max_iterations
        s.x0_seed = self.x0_seed & 0xFFFFFFFF
        s.calc_eigenvector = int(self.calc_eigenvector)
        s.verify_eigenvector = int(self.verify_eigenvector)
        return s


@dataclass
class Result:
    eigval_converged: bool
    eigvec_converged: bool
    n_iterations: int
    eigval_error: float
    eigvec_error: float


class GpuOp:
    """y = H x over a GpuBasis (replaces LazyOpSum<double, ZBasisInterp> and
    LazyOpSumProd).  Borrows the basis, copies the terms."""

    def __init__(self, basis: GpuBasis, H, n_slots=None):
        lib = _lib.load()
        arr, n = H.terms()
        ns = H.n_slots() if n_slots is None else int(n_slots)
        h = C.c_void_p()
        check(lib.qsi_op_create(basis.handle, arr, n, ns, C.byref(h)), "qsi_op_create")
        self._h = h
        self.basis = basis
        self.n_slots = ns
        self._n = basis.dim()

    @property
    def handle(self):
        return self._h

    def set_variant(self, variant):
        """Choose the apply kernel: 'auto' ('cached' when the connectivity is
        cached, else with low keys 'low8' after verify_lowkeys() gave 2 and
        'low4' otherwise, else 'lane4'),
        'cached', 'hash', 'hash4' (hash-index lookups; the basis needs
        hash_index=True), 'lane', 'lane2', 'lane4',
        'lane8' (bucket-table lookups, 1-8 in flight per lane), 'window'
        (LDS-staged bucket windows per wavefront), 'slot2/3/4' (slot index)
        or 'low2/4/8' (low keys)."""
        v = _lib.VARIANTS[variant]
        check(_lib.load().qsi_op_set_variant(self._h, v), "qsi_op_set_variant")

    def verify_lowkeys(self):
        """qsi_op_verify_lowkeys (one pass over the rows; the basis needs low
        keys): 0 when the low-key kernels must compare candidate states, 1
        when they may compare low keys instead, 2 when every connected state
        is in the basis and they need no compare."""
        mode = C.c_int(0)
        check(_lib.load().qsi_op_verify_lowkeys(self._h, C.byref(mode)), "qsi_op_verify_lowkeys")
        return int(mode.value)

    def set_cached_pipe(self, pipe):
        """Pin the cached apply's gathers in flight per lane (4, 8, 16, 32)."""
        check(_lib.load().qsi_op_set_cached_pipe(self._h, int(pipe)), "qsi_op_set_cached_pipe")

    def cached_pipe(self):
        """Gathers in flight per lane of the cached apply."""
        return _lib.load().qsi_op_get_cached_pipe(self._h)

    def set_block_cfg(self, cfg):
        """Pin the block-apply configuration (kernel 0-2, + 4 for a half grid,
        + 8 for two columns per lane in the cached kernel)."""
        check(_lib.load().qsi_op_set_block_cfg(self._h, int(cfg)), "qsi_op_set_block_cfg")

    def block_cfg(self):
        return _lib.load().qsi_op_get_block_cfg(self._h)

    # -- Eigen / Spectra glue
    def rows(self):
        return self._n

    def cols(self):
        return self._n

    def evaluate(self, x, y):
        """y = H x (LazyOpSum::evaluate)."""
        self._apply(x, y, 0)

    def evaluate_add(self, x, y):
        """y += H x (LazyOpSum::evaluate_add; the RealApplyFn contract)."""
        self._apply(x, y, 1)

    def perform_op(self, x_in, y_out):
        """Spectra MatOp: y_out = H x_in (LazyOpSumProd::perform_op)."""
        self._apply(x_in, y_out, 0)

    def _apply(self, x, y, acc):
        _require_cuda_f64(x, self._n, "x")
        _require_cuda_f64(y, self._n, "y")
        check(_lib.load().qsi_op_apply(self._h, _ptr(x), _ptr(y), acc, _stream()), "qsi_op_apply")

    def evaluate_host(self, x, y=None, accumulate=False):
        """Host-vector version (copies through PCIe)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if y is None:
            y = np.zeros(self._n)
        check(_lib.load().qsi_op_apply_host(self._h, x.ctypes.data_as(C.c_void_p),
                                            y.ctypes.data_as(C.c_void_p), int(accumulate)),
              "qsi_op_apply_host")
        return y

    def apply_block(self, X, Y, coeffs=None, accumulate=False):
        """Y[:, r] (+)= H_r X[:, r]; X, Y CUDA float64 (D, R) row-major.
        coeffs: CUDA float64 (n_slots, R) per-column slot coefficients."""
        torch = _torch()
        if X.dim() != 2 or Y.shape != X.shape:
            raise ValueError("X and Y must be (D, R) with the same shape")
        D, R = X.shape
        if D != self._n:
            raise ValueError("row count mismatch")
        for t, nm in ((X, "X"), (Y, "Y")):
            if not (t.is_cuda and t.dtype == torch.float64 and t.stride(1) == 1):
                raise TypeError(f"{nm} must be row-major float64 CUDA")
        cp = None
        if coeffs is not None:
            if tuple(coeffs.shape) != (self.n_slots, R):
                raise ValueError("coeffs must be (n_slots, R)")
            coeffs = coeffs.contiguous()
            cp = _ptr(coeffs)
        check(_lib.load().qsi_op_apply_block(self._h, _ptr(X), X.stride(0), _ptr(Y), Y.stride(0),
                                             R, cp, int(accumulate), _stream()),
              "qsi_op_apply_block")

    def autotune(self, R=0, reps=3, x=None, y=None):
        """Time the apply kernels and keep the fastest (qsi_op_autotune):
        R = 0 the single-vector variant, R >= 1 the block apply for R
        columns.  x, y: optional CUDA float64 scratch (D or D x R).  Returns
        (choice, ms per apply); the choice is a variant name for R = 0 and a
        block configuration number otherwise."""
        best = C.c_int(-1)
        ms = C.c_double(0.0)
        check(_lib.load().qsi_op_autotune(self._h, int(R), int(reps),
                                          _ptr(x) if x is not None else None,
                                          _ptr(y) if y is not None else None,
                                          C.byref(best), C.byref(ms)), "qsi_op_autotune")
        if R == 0:
            names = {v: k for k, v in _lib.VARIANTS.items() if k != "auto"}
            return names.get(best.value, best.value), ms.value
        return best.value, ms.value

    def cache_connectivity(self, enable=True, padded=None):
        """Keep the connected ranks on the device so applies gather instead
        of searching (qsi_op_cache_connectivity).  padded=None lets the
        library pad rows to groups of four when memory allows; False keeps
        them unpadded.  Returns the cached nonzero count (-1 when dropped)."""
        nnz = C.c_int64(-1)
        mode = 0 if not enable else (2 if padded is False else 1)
        check(_lib.load().qsi_op_cache_connectivity(self._h, mode, C.byref(nnz)),
              "qsi_op_cache_connectivity")
        return nnz.value

    def connectivity_padding(self):
        """Rows of the cached connectivity padded to multiples of this many
        entries (4 or 1), or -1 when nothing is cached."""
        return _lib.load().qsi_op_connectivity_padding(self._h)

    def connectivity(self):
        """(rows, cols, vals, diag) numpy arrays: off-diagonal H[row, col] in
        row-major order and the diagonal."""
        torch = _torch()
        lib = _lib.load()
        torch.cuda.synchronize()
        nnz = C.c_int64(0)
        check(lib.qsi_op_connectivity(self._h, C.byref(nnz), None, None, None, None),
              "qsi_op_connectivity")
        n = max(1, nnz.value)
        r = torch.empty(n, dtype=torch.int64, device="cuda")
        c = torch.empty(n, dtype=torch.int64, device="cuda")
        v = torch.empty(n, dtype=torch.float64, device="cuda")
        d = torch.empty(max(1, self._n), dtype=torch.float64, device="cuda")
        check(lib.qsi_op_connectivity(self._h, C.byref(nnz), _ptr(r), _ptr(c), _ptr(v), _ptr(d)),
              "qsi_op_connectivity")
        k = nnz.value
        return (r[:k].cpu().numpy(), c[:k].cpu().numpy(), v[:k].cpu().numpy(),
                d[:self._n].cpu().numpy())

    # -- solvers
    def lanczos_eigval0(self, settings: Settings):
        """lanczos::eigval0 on the GPU: (E0, eigvec tensor or None, Result)."""
        torch = _torch()
        torch.cuda.synchronize()
        ev = C.c_double(0.0)
        vec = torch.zeros(self._n, dtype=torch.float64, device="cuda") \
            if settings.calc_eigenvector else None
        res = _lib.LanczosResult()
        st = settings.to_c()
        check(_lib.load().qsi_lanczos_eigval0(self._h, C.byref(st), C.byref(ev),
                                              _ptr(vec) if vec is not None else None,
                                              C.byref(res)), "qsi_lanczos_eigval0")
        return ev.value, vec, Result(bool(res.eigval_converged), bool(res.eigvec_converged),
                                     int(res.n_iterations), float(res.eigval_error),
                                     float(res.eigvec_error))

    def lanczos_coeffs(self, M, v0=None, seed=0):
        """(alphas, betas) of M steps of the 3-term recurrence from v0 (CUDA
        float64, normalised here) or from randn(D, seed)."""
        torch = _torch()
        torch.cuda.synchronize()
        if v0 is not None:
            _require_cuda_f64(v0, self._n, "v0")
        a = np.zeros(int(M))
        b = np.zeros(int(M))
        n = C.c_int(0)
        check(_lib.load().qsi_lanczos_coeffs(self._h, _ptr(v0) if v0 is not None else None,
                                             int(seed), int(M),
                                             a.ctypes.data_as(C.POINTER(C.c_double)),
                                             b.ctypes.data_as(C.POINTER(C.c_double)),
                                             C.byref(n)), "qsi_lanczos_coeffs")
        return a[: n.value], b[: n.value]

    def eigsh(self, nev=5, ncv=15, tol=1e-10, max_restarts=1000, v0=None, seed=0,
              n_evecs=0, raise_on_noconv=True):
        """Lowest nev eigenpairs (SmallestAlge) by thick-restart Lanczos.
        Returns (evals numpy, evecs CUDA (D, n_evecs) column-major view or
        None, nconv, n_applies)."""
        torch = _torch()
        torch.cuda.synchronize()
        ev = (C.c_double * nev)()
        vecs = None
        if n_evecs:
            vecs = torch.zeros((n_evecs, self._n), dtype=torch.float64, device="cuda")
        if v0 is not None:
            _require_cuda_f64(v0, self._n, "v0")
        nconv = C.c_int(0)
        napp = C.c_int(0)
        rc = _lib.load().qsi_eigsh(self._h, nev, ncv, tol, max_restarts,
                                   _ptr(v0) if v0 is not None else None, seed, ev,
                                   _ptr(vecs) if vecs is not None else None, n_evecs,
                                   C.byref(nconv), C.byref(napp))
        if rc != _lib.QSI_E_NOCONV or raise_on_noconv:
            check(rc, "qsi_eigsh")
        out = np.array(list(ev))
        return out, (vecs.T if vecs is not None else None), nconv.value, napp.value

    def close(self):
        if self._h:
            _lib.load().qsi_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def randn(n, seed):
    """CUDA float64 vector of n standard Gaussian draws fixed by (seed, i)
    (qsi_vec_randn; the start vectors of engine requests)."""
    torch = _torch()
    v = torch.empty(int(n), dtype=torch.float64, device="cuda")
    check(_lib.load().qsi_vec_randn(_ptr(v), int(n), int(seed), _stream()), "qsi_vec_randn")
    return v


def lanczos_ritz(alpha, beta, tol=0.0, max_out=None):
    """Ritz values of a run of the 3-term recurrence without
    reorthogonalisation (the alphas and betas of GpuOp.lanczos_coeffs) after
    the Cullum-Willoughby test (qsi_lanczos_ritz): ghost copies of converged
    eigenvalues merged, spurious eigenvalues dropped.  Returns ascending
    (values, residual estimates |beta_M z_M|, multiplicities in T_M)."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    if a.ndim != 1 or a.shape != b.shape or a.size == 0:
        raise ValueError("lanczos_ritz: alpha and beta must be non-empty 1-d arrays of one length")
    m = a.size
    k = m if max_out is None else int(max_out)
    vals = np.zeros(max(k, 1))
    res = np.zeros(max(k, 1))
    mult = np.zeros(max(k, 1), dtype=np.int32)
    n = C.c_int(0)
    P = C.POINTER(C.c_double)
    check(_lib.load().qsi_lanczos_ritz(a.ctypes.data_as(P), b.ctypes.data_as(P), m, float(tol), k,
                                       vals.ctypes.data_as(P), res.ctypes.data_as(P),
                                       mult.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n)),
          "qsi_lanczos_ritz")
    return vals[: n.value], res[: n.value], mult[: n.value]


def _lattice_arrays(jdata, order):
    from .lattice import greedy_spin_ordering
    W = len(jdata["atoms"])
    tet = np.ascontiguousarray([t["member_spin_idx"] for t in jdata["tetrahedra"]],
                               dtype=np.int32)
    if order == "greedy":
        perm = greedy_spin_ordering(jdata, 0)
    elif order == "none":
        perm = list(range(W))
    else:
        perm = list(order)
    return W, tet, np.ascontiguousarray(perm, dtype=np.int32)


def enumerate_basis_device(jdata, n_spinon_pairs=0, order="greedy", bucket_bits=0):
    """The same basis as enumerate_basis, built and sorted on the device and
    kept there (qsi_enumerate_basis_device): a GpuBasis whose states are never
    copied through the host unless asked for."""
    W, tet, perm = _lattice_arrays(jdata, order)
    h = C.c_void_p()
    dim = C.c_int64(0)
    check(_lib.load().qsi_enumerate_basis_device(
        W, tet.shape[0], tet.ctypes.data_as(C.c_void_p), perm.ctypes.data_as(C.c_void_p),
        int(n_spinon_pairs), int(bucket_bits), C.byref(h), C.byref(dim)),
        "qsi_enumerate_basis_device")
    return GpuBasis._adopt(h, W)


def enumerate_basis(jdata, n_spinon_pairs=0, order="greedy", n_threads=None):
    """Sorted ice-rule basis of a lattice as a (D, 2) uint64 [lo, hi] array
    (gen_spinon_basis, components/build_basis/gen_spinon_basis.cpp:16-156)."""
    W, tet, perm = _lattice_arrays(jdata, order)
    if n_threads is None:
        n_threads = min(16, os.cpu_count() or 1)
    out = C.POINTER(C.c_uint64)()
    dim = C.c_int64(0)
    lib = _lib.load()
    check(lib.qsi_enumerate_basis(W, tet.shape[0], tet.ctypes.data_as(C.c_void_p),
                                  perm.ctypes.data_as(C.c_void_p), int(n_spinon_pairs),
                                  int(n_threads), C.byref(out), C.byref(dim)),
          "qsi_enumerate_basis")
    try:
        n = dim.value
        arr = np.ctypeslib.as_array(out, shape=(max(n, 1) * 2,))[: n * 2].copy().reshape(n, 2)
    finally:
        lib.qsi_free(C.cast(out, C.c_void_p))
    return arr
