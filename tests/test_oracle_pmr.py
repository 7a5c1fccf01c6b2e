"""PMR operator masks and applyState: oracle restatement, the host operator
builder of the package, and the reference's own SymbolicPMROperator compiled
from its sources (oracle/_ref), against the known answers of
test/test_PMR_construct.cpp:81-191."""
import itertools
import random

import pytest

from oracle import pyoracle as orc
from qsi_subspace_ed_amd import SymbolicPMROperator

UP = 1            # Uint128{0, 1}: site 0 up
UP64 = 1 << 64    # Uint128{1, 0}: site 64 up
DOWN = 0

# (spec, expected sign, initial, final)  — test_PMR_construct.cpp:81-108
APPLY_CASES = [
    ("0+", 1, DOWN, UP),
    ("0-", 1, UP, DOWN),
    ("64-", 1, UP64, DOWN),
    ("0X", 1, DOWN, UP),
    ("0X", 1, UP, DOWN),
    ("0Z", 1, UP, UP),
    ("0Z", -1, DOWN, DOWN),
    ("0Z", 1, UP | 7263729, UP | 7263729),
    ("64Z", 1, UP64, UP64),
    ("64Z", -1, DOWN, DOWN),
    ("0+ 1- 2+ 3-", 1, 0b111010, 0b110101),
    ("0- 1+ 2- 3+", 1, 0b110101, 0b111010),
]

# (A, B, expected sign, initial): A|psi> == sign * B|psi>  — :110-158
EQUIV_CASES = [
    ("0Z 1X", "1X 0Z", +1, UP), ("0Z 0X", "0X 0Z", -1, UP), ("0Z 1Z", "1z 0z", +1, UP),
    ("1Z 1Z", "", +1, UP), ("0X 0X", "", +1, UP),
    ("1+ 1Z", "1Z 1+", -1, UP), ("1- 1Z", "1Z 1-", -1, DOWN),
    ("1+ 1X", "1+ 1-", +1, DOWN), ("1- 1X", "1- 1+", +1, UP),
    ("1X 1+", "1- 1+", +1, UP), ("1X 1-", "1+ 1-", +1, DOWN),
    ("0Z 0X 0Z", "0X", -1, UP), ("0X 0Z 0X", "0Z", -1, UP),
    ("0+ 0X 0+", "0+", +1, DOWN), ("0+ 0X 0+", "0+ 0- 0+", +1, DOWN),
    ("0+ 0X 0+", "0+", +1, UP), ("0+ 0X 0+", "0+ 0- 0+", +1, UP),
    ("0- 0X 0-", "0-", +1, UP), ("0- 0X 0-", "0- 0+ 0-", +1, UP),
    ("0+ 0Z", "0Z 0+", -1, UP), ("64+ 64Z", "64Z 64+", -1, UP64),
    ("0X 0+", "0- 0+", +1, UP), ("0X 0-", "0+ 0-", +1, DOWN),
]
ZERO_CASES = [("1+ 1+", UP), ("1- 1-", DOWN), ("0+ 0+", UP), ("0- 0-", DOWN)]


def _apply_py(spec, s):
    if not spec.strip():
        return 1, s
    return SymbolicPMROperator.from_spec(spec).apply_state(s)


def _apply_orc(spec, s, ref=False):
    if not spec.strip():
        return 1, s
    return orc.apply_spec(spec, s, ref=ref)


IMPLS = [("oracle", lambda sp, s: _apply_orc(sp, s)), ("package", _apply_py)]
if orc.ref_available():
    IMPLS.append(("reference", lambda sp, s: _apply_orc(sp, s, ref=True)))


@pytest.mark.parametrize("impl", IMPLS, ids=[i[0] for i in IMPLS])
def test_known_applications(impl):
    _, f = impl
    for spec, sign, init, final in APPLY_CASES:
        sg, out = f(spec, init)
        assert (sg, out) == (sign, final), spec


@pytest.mark.parametrize("impl", IMPLS, ids=[i[0] for i in IMPLS])
def test_known_equivalences(impl):
    _, f = impl
    for A, B, sign, init in EQUIV_CASES:
        sa, oa = f(A, init)
        sb, ob = f(B, init)
        assert oa == ob, (A, B)
        assert sa == sign * sb, (A, B)
    for spec, init in ZERO_CASES:
        assert f(spec, init)[0] == 0, spec


def _random_specs(n, seed, max_site=100):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        k = rng.randint(1, 6)
        toks = [f"{rng.randrange(max_site)}{rng.choice('xzXZ+-pqPQ')}" for _ in range(k)]
        out.append(" ".join(toks))
    return out


def test_masks_oracle_vs_package():
    for spec in _random_specs(2000, 1):
        o = SymbolicPMROperator.from_spec(spec)
        assert orc.pmr_from_spec(spec) == (o.X, o.Z, o.down, o.up, o.sign), spec


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
def test_masks_oracle_vs_reference():
    for spec in _random_specs(2000, 2, max_site=128):
        assert orc.pmr_from_spec(spec) == orc.pmr_from_spec(spec, ref=True), spec


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
def test_apply_state_oracle_vs_reference():
    rng = random.Random(3)
    for spec in _random_specs(500, 4, max_site=128):
        for _ in range(8):
            s = rng.getrandbits(128)
            assert orc.apply_spec(spec, s) == orc.apply_spec(spec, s, ref=True), (spec, s)
            assert _apply_py(spec, s) == orc.apply_spec(spec, s, ref=True), (spec, s)


def test_product_known_answers():
    # test_PMR_construct.cpp:164-191
    f = SymbolicPMROperator.from_spec
    key = lambda o: (o.X, o.Z, o.down, o.up, o.sign)  # noqa: E731
    assert key(f("0Z 1X") * f("1Z")) == key(f("0Z 1X 1Z"))
    assert (f("0+") * f("0-")).sign == 1
    assert (f("0+") * f("0+")).sign == 0
    assert (f("0-") * f("0-")).sign == 0
    assert key(f("0+ 1- 2+ 3-") * f("0- 1+ 4- 5+")) == key(f("0+ 0- 1- 1+ 2+ 3- 4- 5+"))


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
def test_product_vs_reference():
    specs = _random_specs(300, 5, max_site=12)
    import ctypes as C

    import numpy as np
    for a, b in itertools.islice(itertools.product(specs, specs), 3000):
        m = np.zeros(8, dtype=np.uint64)
        s = C.c_int(0)
        assert orc.load_ref().ref_pmr_mul(a.encode(), b.encode(),
                                          m.ctypes.data_as(C.c_void_p), C.byref(s)) == 0
        ref = tuple(int(m[2 * k]) | (int(m[2 * k + 1]) << 64) for k in range(4)) + (s.value,)
        p = SymbolicPMROperator.from_spec(a) * SymbolicPMROperator.from_spec(b)
        assert (p.X, p.Z, p.down, p.up, p.sign) == ref, (a, b)
