"""Lattice files and Hamiltonian construction (host side).

Mirrors the reference's builders so that the same lattice JSON and the same
parameters produce the same term list:
  get_ring_ops            include/physics/geometry.hpp:29-62
  build_hamiltonian (g)   include/hamiltonian_setup.hpp:105-121
  build_hamiltonian (Jpm, B)  include/hamiltonian_setup.hpp:14-101
  get_loc_z, g_vals       include/physics/Jring.hpp:6-54
  greedy_spin_ordering    components/common/src/tetra_graph_io.cpp:101-151
"""
import json
import math
import os

import numpy as np

from .operators import SymbolicOpSum, SymbolicPMROperator


def load_lattice(path):
    with open(path) as f:
        return json.load(f)


def n_spins(jdata):
    return len(jdata["atoms"])


def sublattice_of(jdata):
    return [int(a["sl"]) for a in jdata["atoms"]]


def get_ring_ops(jdata, incl_partial=False):
    """(O list, O^dagger list, sublattice list) for the hexagonal rings.
    O has '+' where the ring's sign is 1 and '-' elsewhere; only 6-member
    rings are used unless incl_partial."""
    ops, ops_h, sls = [], [], []
    for ring in jdata["rings"]:
        spins = ring["member_spin_idx"]
        if not incl_partial and len(spins) != 6:
            continue
        o = ["+" if s == 1 else "-" for s in ring["signs"]]
        oh = ["-" if s == 1 else "+" for s in ring["signs"]]
        ops.append(SymbolicPMROperator.from_ops(o, spins))
        ops_h.append(SymbolicPMROperator.from_ops(oh, spins))
        sls.append(int(ring["sl"]))
    return ops, ops_h, sls


def get_loc_z():
    """Local <111> easy axes of the four sublattices (unit vectors)."""
    r3 = math.sqrt(3.0)
    return np.array([[1.0, 1.0, 1.0],
                     [1.0, -1.0, -1.0],
                     [-1.0, 1.0, -1.0],
                     [-1.0, -1.0, 1.0]]) / r3


def g_vals(Jpm, B):
    """Ring-exchange amplitudes g_s(Jpm, B) of the four hexagon sublattices.

    Written with p_s = sqrt(3) B.z_s (p0 = Bx+By+Bz, p1 = Bx-By-Bz,
    p2 = -Bx+By-Bz, p3 = -Bx-By+Bz) this is
      g_s = [ -21/4 prod_{t != s} p_t^2 - 35/2 (p0 p1 p2 p3) Jpm
              - 15 p_s^2 Jpm^2 + 54 Jpm^3 ] / 36,
    the same polynomial as the reference's expanded form."""
    Bx, By, Bz = (float(v) for v in B)
    p = [Bx + By + Bz, Bx - By - Bz, -Bx + By - Bz, -Bx - By + Bz]
    P = p[0] * p[1] * p[2] * p[3]
    out = []
    for s in range(4):
        prod_others = 1.0
        for t in range(4):
            if t != s:
                prod_others *= p[t] * p[t]
        out.append((-21.0 / 4.0 * prod_others - 35.0 / 2.0 * P * Jpm
                    - 15.0 * p[s] * p[s] * Jpm * Jpm + 54.0 * Jpm ** 3) / 36.0)
    return out


def add_ring_terms(H, jdata, g, slots=False):
    """For each ring: (g[sl], O^dagger) then (g[sl], O), as in
    build_hamiltonian (hamiltonian_setup.hpp:112-119).  With slots=True the
    terms carry coefficient slot sl (per-column coefficients in block
    applies)."""
    ops, ops_h, sls = get_ring_ops(jdata)
    for o, oh, sl in zip(ops, ops_h, sls):
        H.add_term(g[sl], oh, sl if slots else -1)
        H.add_term(g[sl], o, sl if slots else -1)
    return H


def build_hamiltonian_g(jdata, g):
    """H = sum_rings g_sl (O + O^dagger) (hamiltonian_setup.hpp:105-121)."""
    return add_ring_terms(SymbolicOpSum(), jdata, list(g))


def build_hamiltonian_jpm(jdata, Jpm, B):
    """The --Jpm/--B Hamiltonian (hamiltonian_setup.hpp:14-101): ring terms
    with g = g_vals(Jpm, B), bond Ising terms -4 Jpm (B.z_si)(B.z_sj) Z_i Z_j,
    and the optional hand-written neighbour2 / neighbour4 terms."""
    ver = jdata.get("__version__")
    if ver is None:
        raise RuntimeError("__version__ field missing, suspect an old file")
    if float(ver) < 1.1 - 2e-5:
        raise RuntimeError("JSON file is old, API version 1.1 is required")
    g = g_vals(Jpm, B)
    H = add_ring_terms(SymbolicOpSum(), jdata, g)
    loc = get_loc_z()
    sl = sublattice_of(jdata)
    Bv = [float(v) for v in B]
    for bond in jdata["bonds"]:
        i, j = int(bond["from_idx"]), int(bond["to_idx"])
        zi = Bv[0] * loc[sl[i], 0] + Bv[1] * loc[sl[i], 1] + Bv[2] * loc[sl[i], 2]
        zj = Bv[0] * loc[sl[j], 0] + Bv[1] * loc[sl[j], 1] + Bv[2] * loc[sl[j], 2]
        H.add_term(-4.0 * Jpm * zi * zj, SymbolicPMROperator.from_ops("zz", [i, j]))
    for ss in jdata.get("neighbour2", []):
        H.add_term(Jpm / 2, SymbolicPMROperator.from_ops("+-", [ss[0], ss[1]]))
        H.add_term(Jpm / 2, SymbolicPMROperator.from_ops("-+", [ss[0], ss[1]]))
    h4 = Jpm * Jpm / 4
    for ss in jdata.get("neighbour4", []):
        # the reference adds the same operator twice (no conjugate)
        for _ in range(2):
            H.add_term(h4, SymbolicPMROperator.from_ops("+-+-", list(ss[:4])))
    return H


def add_local_field(H, jdata, B, slots=False, slot_base=4):
    """Zeeman term along the local easy axes, H_B = -sum_i (B.z_sl(i)) S^z_i
    with S^z_i = Z_i / 2, written as single-site 'z' terms (the reference's
    'z' operator semantics, operator.hpp:397-400).  With slots=True the four
    sublattice coefficients use slots slot_base..slot_base+3."""
    loc = get_loc_z()
    sl = sublattice_of(jdata)
    Bv = np.asarray(B, dtype=float)
    h = [float(Bv @ loc[s]) for s in range(4)]
    for i in range(n_spins(jdata)):
        H.add_term(-0.5 * h[sl[i]], SymbolicPMROperator.from_ops("z", [i]),
                   slot_base + sl[i] if slots else -1)
    return H, h


def greedy_spin_ordering(jdata, initial_id=0):
    """Spin visiting order of the enumerator: start at initial_id, then
    repeatedly take the first unvisited spin sharing the most tetrahedra with
    the visited set (tetra_graph_io.cpp:101-151)."""
    tet = [t["member_spin_idx"] for t in jdata["tetrahedra"]]
    ids = set()
    for t in tet:
        ids.update(t)
    for r in jdata["rings"]:
        ids.update(r["member_spin_idx"])
    N = len(ids)
    nbrs = [[] for _ in range(N)]
    for ti, t in enumerate(tet):
        for s in t:
            nbrs[s].append(ti)
    selected = [False] * N
    order = [initial_id]
    selected[initial_id] = True
    covered = set(nbrs[initial_id])
    for _ in range(1, N):
        best, best_ov = -1, -1
        for i in range(N):
            if selected[i]:
                continue
            ov = sum(1 for t in nbrs[i] if t in covered)
            if ov > best_ov:
                best_ov, best = ov, i
        if best == -1:
            best = next(i for i in range(N) if not selected[i])
        selected[best] = True
        order.append(best)
        covered.update(nbrs[best])
    return order


LATTICE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattices")


def lattice_path(name):
    """Path of a lattice file shipped with the package (pyrochlore supercells
    with fcc primitive vectors scaled by diag(a, b, c): 'pyro_a,0,0_0,b,0_0,0,c.json')."""
    return os.path.join(LATTICE_DIR, name)
