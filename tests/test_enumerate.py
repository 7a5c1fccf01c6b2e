"""Host basis enumerator (qsi_enumerate_basis) against the reference's basis
fixtures, bit for bit: 8 spins (18 states), 32 spins (2970), 48 spins
(87546).  The reference's own test checks the n=0 set is independent of the
spin visiting order (test/test_spinonfree_222.sh:66-85); so does this one."""
import numpy as np
import pytest

from conftest import load_basis
import qsi_subspace_ed_amd as q


@pytest.mark.parametrize("W", [8, 32, 48])
def test_enumerate_matches_fixture(W, lattices):
    got = q.enumerate_basis(lattices[W], 0, order="greedy", n_threads=4)
    assert np.array_equal(got, load_basis(W))


def test_enumerate_order_independent(lattices):
    jd = lattices[32]
    ref = load_basis(32)
    rng = np.random.default_rng(0)
    for order in ("none", list(rng.permutation(32)), list(rng.permutation(32))):
        got = q.enumerate_basis(jd, 0, order=order, n_threads=3)
        assert np.array_equal(got, ref)


def test_enumerate_threads_agree(lattices):
    jd = lattices[32]
    a = q.enumerate_basis(jd, 1, n_threads=1)
    b = q.enumerate_basis(jd, 1, n_threads=8)
    assert np.array_equal(a, b)
    # sorted, unique, contains the n=0 sector
    s = [int(x) | (int(y) << 64) for x, y in a]
    assert s == sorted(set(s))
    assert set(int(x) for x in load_basis(32)[:, 0]) <= set(s)


def test_greedy_order_is_permutation(lattices):
    for W, jd in lattices.items():
        o = q.greedy_spin_ordering(jd)
        assert sorted(o) == list(range(W)) and o[0] == 0
