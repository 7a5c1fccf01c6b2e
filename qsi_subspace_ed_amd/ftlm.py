"""Finite-temperature Lanczos (FTLM) on the GPU block-Lanczos recurrence.

The reference's FTLM drivers do not run (include/ftlm.hpp:56-91 calls an
undefined krylov_expv_inplace and divides its running sums on every sample;
jl/calc_Cv_ftlm.jl:77-81 evolves only the first sample after step one), so the
estimator here is the textbook one they aim at (Jaklic & Prelovsek): for R
normalised random start vectors r, M Lanczos steps give T_M = Y diag(e) Y^T;
with weights w_k = Y[0, k]^2,
    Z(T)   = (D / R) sum_r sum_k w_rk exp(-(e_rk - E_ref) / T)
    <E>(T) = (D / R) sum_r sum_k w_rk e_rk exp(-(e_rk - E_ref) / T) / Z
    C(T)   = (<E^2> - <E>^2) / T^2,
with E_ref the lowest Ritz value (a common shift that keeps the sums finite).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def random_sign_block(D, R, seed):
    """(D, R) float64 of +-1/sqrt(D) (normalised random-sign start vectors)."""
    rng = np.random.default_rng(seed)
    signs = rng.integers(0, 2, size=(D, R), dtype=np.int8)
    return (1.0 - 2.0 * signs.astype(np.float64)) / np.sqrt(D)


def lanczos_block(ops, V0s, M, coeffs=None, use_graph=True):
    """Run M block-Lanczos steps.  ops: list of GpuOp (one per sector);
    V0s: list of CUDA float64 (D_s, R) row-major start blocks; coeffs:
    optional list of CUDA (n_slots, R) per-column coefficients.  Returns
    (alpha, beta, device_ms) with alpha/beta shaped (n_sectors, R, M)."""
    S = len(ops)
    R = V0s[0].shape[1]
    op_arr = (C.c_void_p * S)(*[o._h.value for o in ops])
    v_arr = (C.c_void_p * S)(*[v.data_ptr() for v in V0s])
    c_arr = None
    if coeffs is not None:
        c_arr = (C.c_void_p * S)(*[(c.data_ptr() if c is not None else None) for c in coeffs])
    alpha = np.zeros(S * R * M)
    beta = np.zeros(S * R * M)
    ms = C.c_double(0.0)
    import torch
    torch.cuda.synchronize()
    check(_lib.load().qsi_lanczos_block(
        op_arr, S, R, v_arr, c_arr, M,
        alpha.ctypes.data_as(C.POINTER(C.c_double)), beta.ctypes.data_as(C.POINTER(C.c_double)),
        int(use_graph), C.byref(ms)), "qsi_lanczos_block")
    return alpha.reshape(S, R, M), beta.reshape(S, R, M), ms.value


def autotune(ops, R, reps=3):
    """Choose the block-apply configuration of lanczos_block(ops, ...) with R
    columns by timing every one on a few steps of the multi-sector block
    Lanczos recurrence (qsi_block_autotune).  Returns (configuration, ms per
    recurrence step)."""
    S = len(ops)
    op_arr = (C.c_void_p * S)(*[o._h.value for o in ops])
    best = C.c_int(-1)
    ms = C.c_double(0.0)
    check(_lib.load().qsi_block_autotune(op_arr, S, int(R), int(reps), C.byref(best),
                                         C.byref(ms)), "qsi_block_autotune")
    return best.value, ms.value


def tridiag_eig(alpha, beta):
    """Ritz values and first-row eigenvector components of one T_M
    (diagonal alpha[0:M], off-diagonal beta[0:M-1])."""
    n = len(alpha)
    d = np.ascontiguousarray(alpha, dtype=np.float64)
    e = np.ascontiguousarray(beta[: max(n - 1, 0)], dtype=np.float64)
    w = np.empty(n)
    f = np.empty(n)
    P = C.POINTER(C.c_double)
    check(_lib.load().qsi_tridiag_eig(d.ctypes.data_as(P), e.ctypes.data_as(P) if n > 1 else None,
                                      n, w.ctypes.data_as(P), f.ctypes.data_as(P)),
          "qsi_tridiag_eig")
    return w, f * f


def quadrature(alpha, beta):
    """Nodes and weights of every sample: lists of (e_k, w_k) arrays for the
    (R, M) coefficient arrays of one sector.  A sample whose recurrence ended
    (beta_j == 0: its Krylov space is invariant, later coefficients are 0)
    keeps only its first j + 1 steps, as the C++ FtlmComputer does."""
    out = []
    for r in range(alpha.shape[0]):
        b = np.asarray(beta[r])
        z = np.flatnonzero(b == 0.0)
        m = int(z[0]) + 1 if z.size else len(alpha[r])
        out.append(tridiag_eig(alpha[r][:m], b[:m]))
    return out


def thermodynamics(nodes, D, temps, e_ref=None, normalise=True):
    """FTLM Z, <E>, C on a temperature grid from per-sample (e, w) nodes.
    With normalise=True Z carries the D/R factor (random start vectors);
    with normalise=False the samples are summed as they are (a complete set
    of orthonormal start vectors gives the exact trace)."""
    temps = np.asarray(temps, dtype=float)
    if e_ref is None:
        e_ref = min(float(np.min(e)) for e, _ in nodes)
    R = len(nodes)
    fac = D / R if normalise else 1.0
    Z = np.zeros_like(temps)
    E1 = np.zeros_like(temps)
    E2 = np.zeros_like(temps)
    for e, w in nodes:
        boltz = np.exp(-np.outer(1.0 / temps, e - e_ref))  # (T, M)
        Z += fac * boltz @ w
        E1 += fac * boltz @ (w * e)
        E2 += fac * boltz @ (w * e * e)
    E = E1 / Z
    Cv = (E2 / Z - E * E) / temps ** 2
    return {"T": temps, "Z_shifted": Z, "E_ref": e_ref, "E": E, "C": Cv}


def sample_operators(H, ops, r, M, betas, e_shift=float("nan")):
    """One sample of the operator-expectation FTLM (qsi_ftlm_sample): from the
    CUDA float64 start vector r (normalised internally), M Lanczos steps and
    psi_b = exp(-beta_b (H - e_shift) / 2) r for every beta_b, built from the
    Lanczos vectors of a second identical pass.  Returns a dict with "norm"
    (<psi_b|psi_b>), "expect" ((n_ops, n_beta) <psi_b|O|psi_b>), "e_shift"
    (the shift used: the sample's lowest Ritz value when e_shift is NaN),
    "alpha" and "beta" (the coefficients)."""
    betas = np.ascontiguousarray(betas, dtype=np.float64)
    nb = betas.size
    no = len(ops)
    op_arr = (C.c_void_p * max(no, 1))(*[o._h.value for o in ops])
    norm = np.zeros(nb)
    expect = np.zeros(max(no, 1) * nb)
    al = np.zeros(int(M))
    be = np.zeros(int(M))
    used = C.c_double(0.0)
    n = C.c_int(0)
    P = C.POINTER(C.c_double)
    import torch
    torch.cuda.synchronize()
    check(_lib.load().qsi_ftlm_sample(H._h, op_arr, no, r.data_ptr(), int(M), betas.ctypes.data_as(P),
                                      nb, float(e_shift), C.byref(used), norm.ctypes.data_as(P),
                                      expect.ctypes.data_as(P), al.ctypes.data_as(P),
                                      be.ctypes.data_as(P), C.byref(n)), "qsi_ftlm_sample")
    m = n.value
    return {"norm": norm, "expect": expect[: no * nb].reshape(no, nb), "e_shift": used.value,
            "alpha": al[:m], "beta": be[:m]}
