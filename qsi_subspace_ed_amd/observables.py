"""Ring observables of eigenvectors on the GPU (the reference's
eval_observables, src/eval_observables.cpp:136-206).

  ring[l]            <psi|O_l|psi>           O_l = the ring operator of
                                             get_ring_ops (geometry.hpp:29-62)
  ring_ring[l, m]    (O_l psi) . (O_m psi)   compute_cross_correlation
  partial_vol[sl][v] <psi|OOO_v|psi>         OOO_v = the three rings of volume
                                             v other than its sublattice-sl
                                             ring, multiplied in the 6 orders
                                             of perm3 and summed
                                             (get_partial_vol_ops,
                                             geometry.hpp:98-142)

Every operator is one qsi_op_t on the eigenvectors' basis; the products and
dot products run on the device (qsi_expectations, qsi_cross_correlations).
For n_vec vectors the arrays carry two more (i, j) axes, as the reference's
datasets do (`ring` is (n_ops, n_vec, n_vec)).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .device import GpuOp
from .lattice import get_ring_ops
from .operators import SymbolicOpSum, SymbolicPMROperator

# the orders of the three factors (geometry.hpp:98-105)
PERM3 = ((0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2))


def ring_operators(basis, jdata):
    """One GpuOp per hexagonal ring: O_l with coefficient 1."""
    rings, _, _ = get_ring_ops(jdata)
    out = []
    for O in rings:
        H = SymbolicOpSum()
        H.add_term(1.0, O)
        out.append(GpuOp(basis, H))
    return out


def partial_vol_operators(basis, jdata, sl):
    """One GpuOp per volume: the sum over the six orders of the product of
    the volume's rings other than its sublattice-sl one."""
    rings, _, _ = get_ring_ops(jdata)
    out = []
    for vol in jdata["vols"]:
        plaq = vol["member_plaq_idx"]
        ops = [rings[plaq[i]] for i in range(4) if i != sl]
        assert len(ops) == 3
        H = SymbolicOpSum()
        for perm in PERM3:
            v = SymbolicPMROperator()
            for k in perm:
                v = v * ops[k]
            H.add_term(1.0, v)
        out.append(GpuOp(basis, H))
    return out


def _psi(psi, D):
    import torch
    if psi.dim() == 1:
        psi = psi.reshape(-1, 1)
    if psi.shape[0] != D:
        raise ValueError("psi rows must equal the basis dimension")
    if not (psi.is_cuda and psi.dtype == torch.float64):
        raise TypeError("psi must be CUDA float64")
    # column-major (D, n_vec): each vector contiguous
    return psi.t().contiguous(), psi.shape[1]


def expectations(ops, psi):
    """(n_ops, n_vec, n_vec) with [l, i, j] = psi_j . (O_l psi_i)
    (compute_all_expectations, src/expectation_eval.cpp:28-53)."""
    import torch
    D = ops[0].rows()
    P, nv = _psi(psi, D)
    arr = (C.c_void_p * len(ops))(*[o.handle.value for o in ops])
    out = np.zeros(len(ops) * nv * nv)
    torch.cuda.synchronize()
    check(_lib.load().qsi_expectations(arr, len(ops), P.data_ptr(), D, nv,
                                       out.ctypes.data_as(C.POINTER(C.c_double))),
          "qsi_expectations")
    return out.reshape(len(ops), nv, nv)


def cross_correlations(ops, psi):
    """(n_ops, n_ops, n_vec, n_vec) with [l, m, i, j] = (O_l psi_i) . (O_m psi_j)
    (compute_cross_correlation, src/expectation_eval.cpp:75-107)."""
    import torch
    D = ops[0].rows()
    P, nv = _psi(psi, D)
    arr = (C.c_void_p * len(ops))(*[o.handle.value for o in ops])
    out = np.zeros(len(ops) * len(ops) * nv * nv)
    torch.cuda.synchronize()
    check(_lib.load().qsi_cross_correlations(arr, len(ops), P.data_ptr(), D, nv,
                                             out.ctypes.data_as(C.POINTER(C.c_double))),
          "qsi_cross_correlations")
    return out.reshape(len(ops), len(ops), nv, nv)


def ring_observables(basis, jdata, psi, partial_vol=True):
    """The eval_observables datasets for eigenvectors psi ((D,) or (D, n_vec)
    CUDA float64): {'ring', 'ring_ring', 'partial_vol_sl0'..'partial_vol_sl3'}."""
    rops = ring_operators(basis, jdata)
    out = {"ring": expectations(rops, psi), "ring_ring": cross_correlations(rops, psi)}
    if partial_vol and "vols" in jdata:
        for sl in range(4):
            out[f"partial_vol_sl{sl}"] = expectations(partial_vol_operators(basis, jdata, sl), psi)
    return out
