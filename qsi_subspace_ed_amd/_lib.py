"""ctypes bindings to libqsi_gpu.so (the C ABI declared in include/qsi_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libqsi_gpu.so")

QSI_OK = 0
QSI_E_INVALID = -1
QSI_E_NOMEM = -2
QSI_E_HIP = -3
QSI_E_NOTFOUND = -4
QSI_E_NOCONV = -5
QSI_E_UNSUPPORTED = -6

QSI_APPLY_AUTO = 0
QSI_APPLY_WINDOW = 1
QSI_APPLY_LANE = 2
QSI_APPLY_LANE2 = 3
QSI_APPLY_LANE4 = 4
QSI_APPLY_LANE8 = 5
QSI_APPLY_HASH = 6
QSI_APPLY_HASH4 = 7
QSI_APPLY_CACHED = 8
QSI_APPLY_SLOT2 = 9
QSI_APPLY_SLOT4 = 10
QSI_APPLY_SLOT3 = 11
QSI_APPLY_LOW2 = 12
QSI_APPLY_LOW4 = 13
QSI_APPLY_LOW8 = 14
VARIANTS = {"auto": QSI_APPLY_AUTO, "window": QSI_APPLY_WINDOW, "lane": QSI_APPLY_LANE,
            "lane2": QSI_APPLY_LANE2, "lane4": QSI_APPLY_LANE4, "lane8": QSI_APPLY_LANE8,
            "hash": QSI_APPLY_HASH, "hash4": QSI_APPLY_HASH4, "cached": QSI_APPLY_CACHED,
            "slot2": QSI_APPLY_SLOT2, "slot3": QSI_APPLY_SLOT3, "slot4": QSI_APPLY_SLOT4,
            "low2": QSI_APPLY_LOW2, "low4": QSI_APPLY_LOW4, "low8": QSI_APPLY_LOW8}


class QsiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qsi error {code}: {msg}")
        self.code = code


class Term(C.Structure):
    """qsi_term_t: the SymbolicPMROperator state plus a coefficient."""
    _fields_ = [
        ("coeff", C.c_double),
        ("X", C.c_uint64 * 2),
        ("Z", C.c_uint64 * 2),
        ("up", C.c_uint64 * 2),
        ("down", C.c_uint64 * 2),
        ("sign", C.c_int32),
        ("slot", C.c_int32),
    ]


class LanczosSettings(C.Structure):
    """qsi_lanczos_settings_t = lanczos::Settings (lanczos.hpp:9-20)."""
    _fields_ = [
        ("krylov_dim", C.c_int64),
        ("abs_tol", C.c_double),
        ("rel_tol", C.c_double),
        ("min_iterations", C.c_int64),
        ("max_iterations", C.c_int64),
        ("x0_seed", C.c_uint32),
        ("calc_eigenvector", C.c_int),
        ("verify_eigenvector", C.c_int),
    ]


class LanczosResult(C.Structure):
    _fields_ = [
        ("eigval_converged", C.c_int),
        ("eigvec_converged", C.c_int),
        ("n_iterations", C.c_int64),
        ("eigval_error", C.c_double),
        ("eigvec_error", C.c_double),
    ]


_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double
_DP = C.POINTER(C.c_double)

# name: (restype, argtypes)
SIGNATURES = {
    "qsi_last_error": (C.c_char_p, []),
    "qsi_version": (C.c_char_p, []),
    "qsi_device_count": (C.c_int, []),
    "qsi_basis_create": (C.c_int, [_P, _I64, C.c_int, C.c_int, C.POINTER(_P)]),
    "qsi_basis_create_ex": (C.c_int, [_P, _I64, C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
    "qsi_basis_has_hash": (C.c_int, [_P]),
    "qsi_basis_build_slots": (C.c_int, [_P, C.c_int]),
    "qsi_basis_has_slots": (C.c_int, [_P]),
    "qsi_basis_build_lowkeys": (C.c_int, [_P, C.c_int]),
    "qsi_op_verify_lowkeys": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "qsi_basis_has_lowkeys": (C.c_int, [_P]),
    "qsi_basis_destroy": (C.c_int, [_P]),
    "qsi_basis_dim": (_I64, [_P]),
    "qsi_basis_n_spins": (C.c_int, [_P]),
    "qsi_basis_bucket_bits": (C.c_int, [_P]),
    "qsi_basis_key_fields": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    "qsi_basis_search": (C.c_int, [_P, _P, _I64, _P, _P]),
    "qsi_op_create": (C.c_int, [_P, C.POINTER(Term), C.c_int, C.c_int, C.POINTER(_P)]),
    "qsi_op_destroy": (C.c_int, [_P]),
    "qsi_op_dim": (_I64, [_P]),
    "qsi_op_n_offdiag": (C.c_int, [_P]),
    "qsi_op_n_diag": (C.c_int, [_P]),
    "qsi_op_set_variant": (C.c_int, [_P, C.c_int]),
    "qsi_op_get_variant": (C.c_int, [_P]),
    "qsi_op_set_cached_pipe": (C.c_int, [_P, C.c_int]),
    "qsi_op_get_cached_pipe": (C.c_int, [_P]),
    "qsi_op_set_block_cfg": (C.c_int, [_P, C.c_int]),
    "qsi_op_get_block_cfg": (C.c_int, [_P]),
    "qsi_op_apply": (C.c_int, [_P, _P, _P, C.c_int, _P]),
    "qsi_op_apply_host": (C.c_int, [_P, _P, _P, C.c_int]),
    "qsi_op_apply_block": (C.c_int, [_P, _P, _I64, _P, _I64, C.c_int, _P, C.c_int, _P]),
    "qsi_op_cache_connectivity": (C.c_int, [_P, C.c_int, C.POINTER(_I64)]),
    "qsi_op_connectivity_padding": (C.c_int, [_P]),
    "qsi_op_autotune": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, C.POINTER(C.c_int), _DP]),
    "qsi_block_autotune": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _DP]),
    "qsi_op_connectivity": (C.c_int, [_P, C.POINTER(_I64), _P, _P, _P, _P]),
    "qsi_lanczos_eigval0": (C.c_int, [_P, C.POINTER(LanczosSettings), _DP, _P,
                                      C.POINTER(LanczosResult)]),
    "qsi_lanczos_coeffs": (C.c_int, [_P, _P, C.c_uint64, C.c_int, _DP, _DP,
                                     C.POINTER(C.c_int)]),
    "qsi_lanczos_coeffs_noalloc": (C.c_int, [_P, _P, _P, C.c_int, _DP, _DP,
                                             C.POINTER(C.c_int)]),
    "qsi_eigsh": (C.c_int, [_P, C.c_int, C.c_int, _D, C.c_int, _P, C.c_uint64, _DP,
                            _P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "qsi_lanczos_block": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.POINTER(_P),
                                    C.POINTER(_P), C.c_int, _DP, _DP, C.c_int, _DP]),
    "qsi_ftlm_sample": (C.c_int, [_P, C.POINTER(_P), C.c_int, _P, C.c_int, _DP, C.c_int, _D,
                                  _DP, _DP, _DP, _DP, _DP, C.POINTER(C.c_int)]),
    "qsi_tridiag_eig": (C.c_int, [_DP, _DP, C.c_int, _DP, _DP]),
    "qsi_lanczos_ritz": (C.c_int, [_DP, _DP, C.c_int, _D, C.c_int, _DP, _DP,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "qsi_engine_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _D, C.c_int, C.POINTER(_P)]),
    "qsi_engine_destroy": (C.c_int, [_P]),
    "qsi_engine_footprint": (_I64, [_I64, C.c_int, C.c_int]),
    "qsi_engine_submit": (C.c_int, [_P, _DP, C.c_uint64, C.POINTER(_I64)]),
    "qsi_engine_active": (C.c_int, [_P]),
    "qsi_engine_queued": (C.c_int, [_P]),
    "qsi_engine_step": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "qsi_engine_step_begin": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "qsi_engine_step_end": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "qsi_engine_step_ready": (C.c_int, [_P]),
    "qsi_engine_inflight": (C.c_int, [_P]),
    "qsi_expectations": (C.c_int, [_P, C.c_int, _P, _I64, C.c_int, _DP]),
    "qsi_cross_correlations": (C.c_int, [_P, C.c_int, _P, _I64, C.c_int, _DP]),
    "qsi_engine_poll": (C.c_int, [_P, C.POINTER(_I64), _DP, C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]),
    "qsi_vec_dot": (C.c_int, [_P, _P, _I64, _DP, _P]),
    "qsi_vec_axpy": (C.c_int, [_D, _P, _P, _I64, _P]),
    "qsi_vec_scal": (C.c_int, [_D, _P, _I64, _P]),
    "qsi_vec_randn": (C.c_int, [_P, _I64, C.c_uint64, _P]),
    "qsi_enumerate_basis": (C.c_int, [C.c_int, C.c_int, _P, _P, C.c_int, C.c_int,
                                      C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(_I64)]),
    "qsi_free": (None, [_P]),
    "qsi_enumerate_basis_device": (C.c_int, [C.c_int, C.c_int, _P, _P, C.c_int, C.c_int,
                                             C.POINTER(_P), C.POINTER(_I64)]),
    "qsi_basis_download": (C.c_int, [_P, _P]),
}

_lib = None


def load():
    """Load libqsi_gpu.so (once) and bind every C-ABI entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP extension is required; there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code, what=""):
    if code != QSI_OK:
        msg = load().qsi_last_error().decode(errors="replace")
        raise QsiError(code, f"{what}: {msg}" if what else msg)
    return code
