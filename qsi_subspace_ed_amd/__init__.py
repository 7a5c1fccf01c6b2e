"""MI355X (gfx950) matrix-free ring-exchange H·v and Lanczos/FTLM for quantum
spin ice, behind the reference's LazyOpSum / Spectra / lanczos interfaces.

Public API:
  SymbolicPMROperator, SymbolicOpSum          host operator algebra
  load_lattice, build_hamiltonian_g/_jpm,
  add_local_field, greedy_spin_ordering       lattice + Hamiltonian builders
  GpuBasis, GpuOp, Settings, enumerate_basis  GPU objects and solvers
  ftlm, serve, observables, h5io, diag        batched FTLM, solve engine/server,
                                              ring observables, HDF5 files,
                                              diagonalise_real's algorithm
                                              dispatch (submodules)
"""
from .operators import SymbolicOpSum, SymbolicPMROperator  # noqa: F401
from .lattice import (add_local_field, build_hamiltonian_g,  # noqa: F401
                      build_hamiltonian_jpm, g_vals, get_loc_z, get_ring_ops,
                      greedy_spin_ordering, lattice_path, load_lattice)
from .device import (GpuBasis, GpuOp, Settings, as_lohi, enumerate_basis,  # noqa: F401
                     enumerate_basis_device, lanczos_ritz)

__version__ = "0.1.0"
