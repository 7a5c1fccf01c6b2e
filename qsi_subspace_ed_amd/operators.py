"""Host-side operator algebra: the reference's SymbolicPMROperator /
SymbolicOpSum, producing the qsi_term_t records the GPU library consumes.

An operator is a generalised permutation of Z-basis states held as four bit
masks and a sign (components/operator/include/operator.hpp:215-467):
it maps |s> to sign * (-1)^popcount(~s' & Z) |s'>, s' = s ^ X, when
(s & down) == 0 and (s & up) == up, and to 0 otherwise.  Masks are Python
ints (bit i = spin i), split into [lo, hi] words for the C ABI.
"""
from dataclasses import dataclass, field
from typing import List, Tuple

from . import _lib

_M64 = (1 << 64) - 1


def _words(x):
    return (x & _M64, (x >> 64) & _M64)


@dataclass
class SymbolicPMROperator:
    X: int = 0
    Z: int = 0
    down: int = 0
    up: int = 0
    sign: int = 1

    @classmethod
    def from_ops(cls, ops, spin_ids):
        """Build from an op string and spin ids, processed right to left as in
        SymbolicPMROperator::_construct (operator.hpp:383-466).  Ops: x/X,
        z/Z, + (S+), - (S-), p/P (projector on up), q/Q (projector on down)."""
        if len(ops) != len(spin_ids):
            raise ValueError(f"bad_operator_spec: {len(ops)} ops vs {len(spin_ids)} ids")
        o = cls()
        for op, j in zip(reversed(list(ops)), reversed(list(spin_ids))):
            b = 1 << int(j)
            if op in "xX":
                o.X ^= b
                if o.Z & b:
                    o.sign = -o.sign
            elif op in "zZ":
                o.Z ^= b
            elif op == "+":
                if o.down & o.X & b:      # S+ S+ = 0
                    o.sign = 0
                    continue
                if o.Z & b:
                    o.sign = -o.sign
                o.X ^= b
                if o.X & b:
                    o.down |= b
                else:
                    o.up |= b
            elif op == "-":
                if o.up & o.X & b:        # S- S- = 0
                    o.sign = 0
                    continue
                if o.Z & b:
                    o.sign = -o.sign
                o.X ^= b
                if o.X & b:
                    o.up |= b
                else:
                    o.down |= b
            elif op in "pP":
                if o.up & o.X & b:
                    o.sign = 0
                    continue
                if o.X & b:
                    o.up |= b
                else:
                    o.down |= b
            elif op in "qQ":
                if o.down & o.X & b:
                    o.sign = 0
                    continue
                if o.Z & b:
                    o.sign = -o.sign
                o.X ^= b
                if o.X & b:
                    o.down |= b
                else:
                    o.up |= b
            else:
                raise ValueError(f"bad_operator_spec: op {op!r}")
        return o

    @classmethod
    def from_spec(cls, spec: str):
        """"0+ 1- 2Z" style spec (SymbolicPMROperator::parseSpec,
        operator.hpp:349-381)."""
        ops, ids = [], []
        for tok in spec.split():
            if len(tok) < 2:
                raise ValueError("Invalid token in operator string: " + tok)
            i = 0
            while i < len(tok) - 1 and tok[i].isdigit():
                i += 1
            op = tok[i]
            if op not in "xyzXYZ+-pqPQ":
                raise ValueError(f"Invalid op '{op}' in token: {tok}")
            ops.append(op)
            ids.append(int(tok[:i]) if i else 0)
        return cls.from_ops(ops, ids)

    def is_diagonal(self):
        return self.X == 0

    def apply_state(self, s: int) -> Tuple[int, int]:
        """applyState (operator.hpp:232-254): returns (sign, new_state)."""
        ok = (s & self.down) == 0 and (s & self.up) == self.up
        if not ok:
            return 0, s
        s2 = s ^ self.X
        par = bin(~s2 & self.Z & ((1 << 128) - 1)).count("1") & 1
        return self.sign * (1 - 2 * par), s2

    def __mul__(self, b):
        """operator*= (operator.hpp:312-330); the product's sign starts from
        the left factor's sign."""
        c = SymbolicPMROperator(self.X, self.Z, self.down, self.up, self.sign)
        c.sign *= 1 - 2 * (bin(c.X & b.Z).count("1") & 1)
        c.Z ^= b.Z
        c.X ^= b.X
        req_up = (c.down & b.X) | (c.up & ~b.X)
        req_dn = (c.up & b.X) | (c.down & ~b.X)
        c.up = b.up | req_up
        c.down = b.down | req_dn
        if c.up & c.down:
            c.sign = 0
        return c

    def to_term(self, coeff: float, slot: int = -1) -> _lib.Term:
        t = _lib.Term()
        t.coeff = float(coeff)
        t.X[0], t.X[1] = _words(self.X)
        t.Z[0], t.Z[1] = _words(self.Z)
        t.up[0], t.up[1] = _words(self.up)
        t.down[0], t.down[1] = _words(self.down)
        t.sign = int(self.sign)
        t.slot = int(slot)
        return t


@dataclass
class SymbolicOpSum:
    """SymbolicOpSum (operator.hpp:476-503): terms split into diagonal and
    off-diagonal lists in insertion order.  `slots` optionally tags each term
    with a coefficient slot for block applies with per-column coefficients."""
    diagonal_terms: List[Tuple[float, SymbolicPMROperator, int]] = field(default_factory=list)
    off_diag_terms: List[Tuple[float, SymbolicPMROperator, int]] = field(default_factory=list)

    def add_term(self, c, op: SymbolicPMROperator, slot: int = -1):
        if op.is_diagonal():
            self.diagonal_terms.append((float(c), op, slot))
        else:
            self.off_diag_terms.append((float(c), op, slot))

    def terms(self):
        """All terms, off-diagonal first then diagonal (each in insertion
        order), as a ctypes array of qsi_term_t."""
        allt = list(self.off_diag_terms) + list(self.diagonal_terms)
        arr = (_lib.Term * max(1, len(allt)))()
        for k, (c, op, slot) in enumerate(allt):
            arr[k] = op.to_term(c, slot)
        return arr, len(allt)

    def n_slots(self):
        s = [slot for (_, _, slot) in self.off_diag_terms + self.diagonal_terms]
        return max(s) + 1 if s else 0
