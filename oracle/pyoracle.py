"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py).  ctypes wrappers over oracle/liboracle.so (the
CPU restatement of the reference) and oracle/_ref/libref.so (the reference's
own operator core, when built)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libref.so")


class Term(C.Structure):
    _fields_ = [("coeff", C.c_double), ("X", C.c_uint64 * 2), ("Z", C.c_uint64 * 2),
                ("up", C.c_uint64 * 2), ("down", C.c_uint64 * 2),
                ("sign", C.c_int32), ("slot", C.c_int32)]


class LzSettings(C.Structure):
    _fields_ = [("krylov_dim", C.c_int64), ("abs_tol", C.c_double), ("rel_tol", C.c_double),
                ("min_iterations", C.c_int64), ("max_iterations", C.c_int64),
                ("x0_seed", C.c_uint32), ("calc_eigenvector", C.c_int),
                ("verify_eigenvector", C.c_int)]


class LzResult(C.Structure):
    _fields_ = [("eigval_converged", C.c_int), ("eigvec_converged", C.c_int),
                ("n_iterations", C.c_int64), ("eigval_error", C.c_double),
                ("eigvec_error", C.c_double)]


_P = C.c_void_p
_lib = None
_ref = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        lib = C.CDLL(LIB)
        sig = {
            "oracle_num_threads": (C.c_int, []),
            "oracle_set_threads": (None, [C.c_int]),
            "oracle_pmr_construct": (C.c_int, [C.c_char_p, _P, C.c_int, _P, C.POINTER(C.c_int)]),
            "oracle_pmr_from_spec": (C.c_int, [C.c_char_p, _P, C.POINTER(C.c_int)]),
            "oracle_apply_state": (C.c_int, [_P, C.c_int, _P]),
            "oracle_pmr_mul": (None, [_P, C.POINTER(C.c_int), _P, C.c_int]),
            "oracle_basis_new": (_P, [_P, C.c_int64]),
            "oracle_basis_free": (None, [_P]),
            "oracle_basis_search": (None, [_P, _P, C.c_int64, _P, C.c_int]),
            "oracle_evaluate_add": (C.c_int, [_P, _P, C.c_int, _P, _P]),
            "oracle_evaluate_add_rows": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int64,
                                                   C.c_int64]),
            "oracle_sparse_triplets": (C.c_int64, [_P, _P, C.c_int, C.c_double, C.c_int64,
                                                   _P, _P, _P]),
            "oracle_random_unit": (None, [C.c_uint32, _P, C.c_int64]),
            "oracle_normal_stream": (None, [C.c_uint32, _P, C.c_int64]),
            "oracle_lanczos_eigval0": (C.c_int, [_P, _P, C.c_int, C.POINTER(LzSettings),
                                                 C.POINTER(C.c_double), _P, C.POINTER(LzResult),
                                                 _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
            "oracle_lanczos_coeffs": (C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, _P]),
            "oracle_lanczos_dense": (C.c_int, [_P, C.c_int64, C.POINTER(LzSettings),
                                               C.POINTER(C.c_double), _P, C.POINTER(LzResult)]),
            "oracle_jacobi_eig": (None, [_P, C.c_int, _P, _P]),
            "oracle_tridiag_lowest": (C.c_double, [_P, _P, C.c_int64]),
            "oracle_tridiag_one": (None, [_P, _P, C.c_int64, C.POINTER(C.c_double), _P]),
            "oracle_brute_basis": (C.c_int64, [C.c_int, C.c_int, _P, _P, C.c_int, _P, _P,
                                               C.c_int64]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def ref_available():
    return os.path.exists(REF)


def load_ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF)
        sig = {
            "ref_pmr_from_spec": (C.c_int, [C.c_char_p, _P, C.POINTER(C.c_int)]),
            "ref_apply_spec": (C.c_int, [C.c_char_p, _P]),
            "ref_pmr_mul": (C.c_int, [C.c_char_p, C.c_char_p, _P, C.POINTER(C.c_int)]),
            "ref_basis_new": (_P, [_P, C.c_int64]),
            "ref_basis_free": (None, [_P]),
            "ref_basis_search": (None, [_P, _P, C.c_int64, _P, C.c_int]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _ref = lib
    return _ref


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def terms_from_opsum(H):
    """ctypes Term array from a qsi_subspace_ed_amd SymbolicOpSum, same order
    as the GPU library receives (off-diagonal first, then diagonal)."""
    allt = list(H.off_diag_terms) + list(H.diagonal_terms)
    arr = (Term * max(1, len(allt)))()
    M = (1 << 64) - 1
    for k, (c, op, slot) in enumerate(allt):
        t = arr[k]
        t.coeff = c
        for name in ("X", "Z", "up", "down"):
            v = getattr(op, name)
            getattr(t, name)[0] = v & M
            getattr(t, name)[1] = v >> 64
        t.sign = op.sign
        t.slot = slot
    return arr, len(allt)


def pmr_from_spec(spec, ref=False):
    """(X, Z, down, up, sign) as Python ints from an operator spec."""
    m = np.zeros(8, dtype=np.uint64)
    s = C.c_int(0)
    f = load_ref().ref_pmr_from_spec if ref else load().oracle_pmr_from_spec
    if f(spec.encode(), _ptr(m), C.byref(s)) != 0:
        raise ValueError("bad spec " + spec)
    v = [int(m[2 * k]) | (int(m[2 * k + 1]) << 64) for k in range(4)]
    return (*v, s.value)


def apply_spec(spec, state, ref=False):
    st = np.array([state & ((1 << 64) - 1), state >> 64], dtype=np.uint64)
    if ref:
        sg = load_ref().ref_apply_spec(spec.encode(), _ptr(st))
    else:
        X, Z, dn, up, sign = pmr_from_spec(spec)
        m = np.array([X & ((1 << 64) - 1), X >> 64, Z & ((1 << 64) - 1), Z >> 64,
                      dn & ((1 << 64) - 1), dn >> 64, up & ((1 << 64) - 1), up >> 64],
                     dtype=np.uint64)
        sg = load().oracle_apply_state(_ptr(m), sign, _ptr(st))
    return sg, int(st[0]) | (int(st[1]) << 64)


class Basis:
    """ZBasisInterp restatement (or, with ref=True, the reference's own)."""

    def __init__(self, lohi, ref=False):
        self.lohi = np.ascontiguousarray(lohi, dtype=np.uint64)
        self.ref = ref
        lib = load_ref() if ref else load()
        self._h = (lib.ref_basis_new if ref else lib.oracle_basis_new)(
            _ptr(self.lohi), self.lohi.shape[0])

    def search(self, q, bst=False):
        q = np.ascontiguousarray(q, dtype=np.uint64).reshape(-1, 2)
        out = np.empty(q.shape[0], dtype=np.int64)
        if self.ref:
            load_ref().ref_basis_search(self._h, _ptr(q), q.shape[0], _ptr(out), int(bst))
        else:
            load().oracle_basis_search(self._h, _ptr(q), q.shape[0], _ptr(out), int(bst))
        return out

    def __del__(self):
        try:
            if self.ref:
                load_ref().ref_basis_free(self._h)
            else:
                load().oracle_basis_free(self._h)
        except Exception:
            pass


def evaluate_add(basis: Basis, H, x, y):
    """y += H x with the reference's push algorithm (in place on y)."""
    arr, n = terms_from_opsum(H)
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    rc = load().oracle_evaluate_add(basis._h, arr, n, _ptr(x), _ptr(y))
    if rc != 0:
        raise RuntimeError("oracle apply: a connected state's high word has no bucket")
    return y


def evaluate_add_rows(basis: Basis, H, x, y, r0, r1):
    """The pushes of rows [r0, r1) of y += H x (in place on y)."""
    arr, n = terms_from_opsum(H)
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    rc = load().oracle_evaluate_add_rows(basis._h, arr, n, _ptr(x), _ptr(y),
                                         C.c_int64(int(r0)), C.c_int64(int(r1)))
    if rc != 0:
        raise RuntimeError("oracle apply: a connected state's high word has no bucket")
    return y


def sparse_triplets(basis: Basis, H, tol=1e-14):
    arr, n = terms_from_opsum(H)
    lib = load()
    k = lib.oracle_sparse_triplets(basis._h, arr, n, tol, 0, None, None, None)
    r = np.empty(k, dtype=np.int64)
    c = np.empty(k, dtype=np.int64)
    v = np.empty(k, dtype=np.float64)
    lib.oracle_sparse_triplets(basis._h, arr, n, tol, k, _ptr(r), _ptr(c), _ptr(v))
    return r, c, v


def expectations(basis: Basis, op_sums, psi):
    """compute_all_expectations (src/expectation_eval.cpp:28-53) for one or
    more vectors: [l, i, j] = psi_j . (O_l psi_i); psi (D,) or (D, n_vec).
    Each O_l psi_i is the push apply above (the reference materialises O_l as
    a sparse matrix, operator_matrix.hpp:79-105; same products)."""
    P = psi.reshape(psi.shape[0], -1)
    Y = _apply_all(basis, op_sums, P)
    return np.einsum("lir,rj->lij", Y, P)


def cross_correlations(basis: Basis, op_sums, psi):
    """compute_cross_correlation (src/expectation_eval.cpp:75-107):
    [l, m, i, j] = (O_l psi_i) . (O_m psi_j)."""
    P = psi.reshape(psi.shape[0], -1)
    Y = _apply_all(basis, op_sums, P)
    return np.einsum("lir,mjr->lmij", Y, Y)


def _apply_all(basis, op_sums, P):
    D, nv = P.shape
    Y = np.zeros((len(op_sums), nv, D))
    for l, H in enumerate(op_sums):
        for i in range(nv):
            Y[l, i] = evaluate_add(basis, H, np.ascontiguousarray(P[:, i]), np.zeros(D))
    return Y


def normal_stream(seed, n):
    v = np.empty(n, dtype=np.float64)
    load().oracle_normal_stream(seed, _ptr(v), n)
    return v


def random_unit(seed, n):
    v = np.empty(n, dtype=np.float64)
    load().oracle_random_unit(seed, _ptr(v), n)
    return v


def settings(krylov_dim=10, abs_tol=1e-8, rel_tol=1e-8, min_iterations=30,
             max_iterations=5000, x0_seed=10, calc_eigenvector=False, verify_eigenvector=False):
    s = LzSettings()
    s.krylov_dim, s.abs_tol, s.rel_tol = krylov_dim, abs_tol, rel_tol
    s.min_iterations, s.max_iterations, s.x0_seed = min_iterations, max_iterations, x0_seed
    s.calc_eigenvector, s.verify_eigenvector = int(calc_eigenvector), int(verify_eigenvector)
    return s


def lanczos_eigval0(basis: Basis, H, st: LzSettings, cap=0):
    arr, n = terms_from_opsum(H)
    D = basis.lohi.shape[0]
    ev = C.c_double(0.0)
    vec = np.zeros(D)
    res = LzResult()
    a = np.zeros(max(cap, 1))
    b = np.zeros(max(cap, 1))
    nab = C.c_int64(0)
    rc = load().oracle_lanczos_eigval0(basis._h, arr, n, C.byref(st), C.byref(ev), _ptr(vec),
                                       C.byref(res), _ptr(a) if cap else None,
                                       _ptr(b) if cap else None, cap, C.byref(nab))
    if rc != 0:
        raise RuntimeError(f"oracle lanczos failed rc={rc}")
    k = min(nab.value, cap)
    return ev.value, (vec if st.calc_eigenvector else None), res, a[:k], b[:k]


def lanczos_coeffs(basis: Basis, H, v0, M):
    """(alphas, betas) of M steps of the 3-term recurrence from v0."""
    arr, n = terms_from_opsum(H)
    v0 = np.ascontiguousarray(v0, dtype=np.float64)
    a = np.zeros(M)
    b = np.zeros(M)
    if load().oracle_lanczos_coeffs(basis._h, arr, n, _ptr(v0), M, _ptr(a), _ptr(b)) != 0:
        raise RuntimeError("oracle lanczos_coeffs failed")
    return a, b


def lanczos_dense(M, st: LzSettings):
    M = np.ascontiguousarray(M, dtype=np.float64)
    n = M.shape[0]
    ev = C.c_double(0.0)
    vec = np.zeros(n)
    res = LzResult()
    load().oracle_lanczos_dense(_ptr(M), n, C.byref(st), C.byref(ev), _ptr(vec), C.byref(res))
    return ev.value, vec, res


def jacobi_eig(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.empty(n)
    V = np.empty((n, n), order="F")
    load().oracle_jacobi_eig(A.ctypes.data_as(C.c_void_p), n, _ptr(w),
                             V.ctypes.data_as(C.c_void_p))
    return w, V


def tridiag_one(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    e0 = C.c_double(0)
    v = np.empty(d.shape[0])
    load().oracle_tridiag_one(_ptr(d), _ptr(e), d.shape[0], C.byref(e0), _ptr(v))
    return e0.value, v


def brute_basis(n_spins, tetra_members, order, n_pairs, cap=1 << 26):
    """Every state of a <= 32-spin lattice replayed along the visiting order
    with the basis builder's rule (brute_basis.cpp): (states ascending, count
    of disallowed positions per state) for all states with at most n_pairs."""
    lib = load()
    tet = np.ascontiguousarray(tetra_members, dtype=np.int32)
    order = np.ascontiguousarray(order, dtype=np.int32)
    out = np.zeros(cap, dtype=np.uint64)
    bad = np.zeros(cap, dtype=np.int8)
    n = lib.oracle_brute_basis(int(n_spins), tet.shape[0], _ptr(tet), _ptr(order), int(n_pairs),
                               _ptr(out), _ptr(bad), cap)
    if n < 0 or n > cap:
        raise ValueError("oracle_brute_basis: bad lattice or too many states")
    return out[:n].copy(), bad[:n].copy()
