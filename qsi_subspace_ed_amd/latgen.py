"""Pyrochlore supercell lattice files (the reference's JSON v1.1 layout).

Produces the same content as the reference's generator
(scripts/make_latfile.py -> scripts/pyrochlore.py:export_json, with the
containers of scripts/lattice.py) for diagonal supercells diag(n0, n1, n2) of
the fcc primitive cell, in exact integer arithmetic:

  primitive vectors   a0 = (0,4,4), a1 = (4,0,4), a2 = (4,4,0)   (columns of A)
  sublattice offsets  (0,0,0), (0,2,2), (2,0,2), (2,2,0)        (pyrochlore.py:7-13)
  atoms               sublattice-major, then cells in (i0, i1, i2) row-major
                      order (lattice.py:433-437); index = i2 + n2 (i1 + n1 (i0 + n0 sl))
  bonds               per sublattice, per cell, the sublattice's bond list:
                      pairs (0,1),(0,2),(0,3),(1,2),(2,3),(3,1) with colour = pair
                      index, +delta then -delta (pyrochlore.py:200-212)
  tetrahedra          all up tetrahedra (cell + disp_j) then all down ones
                      (cell - disp_j), positions cell +- (1,1,1) unwrapped
                      (pyrochlore.py:95-110)
  rings               per cell, per hexagon sublattice mu: the 6 sites around
                      the wrapped ring centre, signs -1, +1, ... (pyrochlore.py:113-142)
  vols                per cell, eta = +1, -1: the 4 rings around the wrapped
                      volume centre (pyrochlore.py:145-170)

Reduced coordinates are kept in units of 1/8 of a primitive vector
(A^-1 = [[-1,1,1],[1,-1,1],[1,1,-1]] / 8), so wrapping into the supercell and
finding a site's index are integer operations.
"""
import json

import numpy as np

_A = np.array([[0, 4, 4], [4, 0, 4], [4, 4, 0]], dtype=np.int64).T  # columns a0, a1, a2
_AINV8 = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], dtype=np.int64)  # 8 A^-1
_DISP = [np.array(v, dtype=np.int64) for v in ([0, 0, 0], [0, 2, 2], [2, 0, 2], [2, 2, 0])]
# sublattice of a site from its fractional cell position (units of 1/8)
_SL_OF_FRAC = {(0, 0, 0): 0, (4, 0, 0): 1, (0, 4, 0): 2, (0, 0, 4): 3}
_TETRA_POS = np.array([1, 1, 1], dtype=np.int64)
_PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3), (3, 1)]
_PLAQ = [
    [[0, -2, 2], [2, -2, 0], [2, 0, -2], [0, 2, -2], [-2, 2, 0], [-2, 0, 2]],
    [[0, 2, -2], [2, 2, 0], [2, 0, 2], [0, -2, 2], [-2, -2, 0], [-2, 0, -2]],
    [[0, -2, -2], [-2, -2, 0], [-2, 0, 2], [0, 2, 2], [2, 2, 0], [2, 0, -2]],
    [[0, 2, 2], [-2, 2, 0], [-2, 0, -2], [0, -2, -2], [2, -2, 0], [2, 0, 2]],
]
_PLAQ_LOC = [[4, 4, 4], [4, 2, 2], [2, 4, 2], [2, 2, 4]]
_VOL_DELTA = [[-1, -1, -1], [-1, 1, 1], [1, -1, 1], [1, 1, -1]]
_VOL_LOC = [[3, 3, 3], [5, 5, 5]]


class Supercell:
    """diag(n0, n1, n2) supercell of the pyrochlore primitive cell."""

    def __init__(self, n):
        self.n = np.array([int(v) for v in n], dtype=np.int64)
        if self.n.shape != (3,) or np.any(self.n < 1):
            raise ValueError("need three positive repeats")

    def cells(self):
        n0, n1, n2 = (int(v) for v in self.n)
        for i0 in range(n0):
            for i1 in range(n1):
                for i2 in range(n2):
                    yield np.array([i0, i1, i2], dtype=np.int64)

    def _wrapped_m8(self, x):
        """Reduced coordinates (units of 1/8) wrapped into the supercell."""
        m8 = _AINV8 @ np.asarray(x, dtype=np.int64)
        return np.mod(m8, 8 * self.n)

    def wrap(self, x):
        """Position wrapped into the supercell (lattice.py:_wrap_coordinate)."""
        m8 = self._wrapped_m8(x)
        y8 = _A @ m8
        if np.any(y8 % 8):
            raise ValueError(f"{x} does not wrap to an integer position")
        return y8 // 8

    def index(self, x):
        """Site index of the lattice site at position x (as_linear_idx)."""
        m8 = self._wrapped_m8(x)
        cell = m8 // 8
        frac = tuple(int(v) for v in m8 % 8)
        sl = _SL_OF_FRAC.get(frac)
        if sl is None:
            raise ValueError(f"{x} is not a lattice site")
        n0, n1, n2 = (int(v) for v in self.n)
        return int(cell[2] + n2 * (cell[1] + n1 * (cell[0] + n0 * sl)))

    def cell_origin(self, c):
        return _A @ c

    def to_dict(self):
        f = lambda v: [float(x) for x in v]
        n = self.n
        out = {
            "primitive": {"lattice_vectors": {"a0": f(_A[:, 0]), "a1": f(_A[:, 1]),
                                              "a2": f(_A[:, 2])}},
            "lattice_vectors": {"A0": f(_A[:, 0] * n[0]), "A1": f(_A[:, 1] * n[1]),
                                "A2": f(_A[:, 2] * n[2])},
        }
        atoms = []
        for sl in range(4):
            for c in self.cells():
                atoms.append({"xyz": f(_DISP[sl] + self.cell_origin(c)), "sl": str(sl)})
        out["atoms"] = atoms
        bonds_of = {sl: [] for sl in range(4)}
        for colour, (i, j) in enumerate(_PAIRS):
            d = _DISP[j] - _DISP[i]
            bonds_of[i].append((d, colour))
            bonds_of[i].append((-d, colour))
        bonds = []
        for sl in range(4):
            for c in self.cells():
                x0 = self.cell_origin(c) + _DISP[sl]
                for d, colour in bonds_of[sl]:
                    bonds.append({"from_idx": self.index(x0), "to_idx": self.index(x0 + d),
                                  "bond_delta": f(d), "color": colour})
        out["bonds"] = bonds
        out["__version__"] = "1.1"
        up, dn = [], []
        for c in self.cells():
            x0 = self.cell_origin(c)
            up.append({"xyz": f(x0 + _TETRA_POS), "sl": 0,
                       "member_spin_idx": [self.index(x0 + d) for d in _DISP]})
            dn.append({"xyz": f(x0 - _TETRA_POS), "sl": 1,
                       "member_spin_idx": [self.index(x0 - d) for d in _DISP]})
        out["tetrahedra"] = up + dn
        rings = []
        for c in self.cells():
            x0 = self.cell_origin(c)
            for mu in range(4):
                pos = self.wrap(np.array(_PLAQ_LOC[mu]) + x0)
                members, signs, sign = [], [], 1
                for d in _PLAQ[mu]:
                    sign = -sign
                    members.append(self.index(pos + np.array(d)))
                    signs.append(sign)
                rings.append({"xyz": f(pos), "sl": mu, "member_spin_idx": members,
                              "signs": signs})
        out["rings"] = rings
        ring_at = {tuple(int(v) for v in r["xyz"]): k for k, r in enumerate(rings)}
        vols = []
        for c in self.cells():
            x0 = self.cell_origin(c)
            for i_eta, eta in enumerate((1, -1)):
                pos = self.wrap(np.array(_VOL_LOC[i_eta]) + x0)
                ids = [ring_at[tuple(int(v) for v in self.wrap(pos - eta * np.array(d)))]
                       for d in _VOL_DELTA]
                vols.append({"xyz": f(pos), "sl": i_eta, "member_plaq_idx": ids})
        out["vols"] = vols
        return out


def lattice_name(n):
    """File name the reference gives diag(n0, n1, n2) (make_latfile.py)."""
    n0, n1, n2 = (int(v) for v in n)
    return f"pyro_{n0},0,0_0,{n1},0_0,0,{n2}.json"


def write_lattice(n, path):
    with open(path, "w") as fh:
        json.dump(Supercell(n).to_dict(), fh)
    return path


if __name__ == "__main__":
    import os
    import sys
    reps = [int(v) for v in sys.argv[1:4]] if len(sys.argv) >= 4 else [2, 4, 3]
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattices")
    print(write_lattice(reps, os.path.join(here, lattice_name(reps))))
