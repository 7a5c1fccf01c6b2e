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


def _lat_arrays(jd):
    from qsi_subspace_ed_amd.device import _lattice_arrays
    return _lattice_arrays(jd, "greedy")


@pytest.mark.parametrize("W,n_max", [(8, 3), (32, 2)])
def test_spinon_sectors_brute_force(W, n_max, lattices):
    """Sectors n_spinon_pairs = 0 .. n_max pinned by brute force: every one of
    the 2^W states replayed along the greedy visiting order with the
    reference's rule (oracle/brute_basis.cpp: possible_spin_states and
    fork_state_impl, pyro_tree.cpp:20-81,168-189); the states with at most n
    disallowed positions must equal qsi_enumerate_basis(n), bit for bit."""
    from oracle import pyoracle as orc
    jd = lattices[W]
    n, tet, order = _lat_arrays(jd)
    states, bad = orc.brute_basis(n, tet, order, n_max)
    for k in range(n_max + 1):
        got = q.enumerate_basis(jd, k, n_threads=8)
        assert np.array_equal(got[:, 1], np.zeros(got.shape[0], dtype=np.uint64))
        assert np.array_equal(got[:, 0], states[bad <= k]), k
    assert np.array_equal(states[bad == 0], load_basis(W)[:, 0])


@pytest.mark.parametrize("W", [8, 32, 48])
def test_one_pair_sector_is_last_spin_flip(W, lattices):
    """Why D(n=1) = 2 D(n=0).  Every spin sits in two tetrahedra, so the up
    counts of all tetrahedra sum to twice the number of up spins: an even
    number.  A single tetrahedron with 1 or 3 up spins among tetrahedra that
    all hold 2 is therefore impossible, and one disallowed choice that breaks
    one tetrahedron dooms the branch (a second break would cost a second
    pair).  Only a spin that completes both of its tetrahedra at once can
    break two with one pair: the last visited spin does (in these lattices no
    other does), so the one-pair sector is the ice states plus the ice states
    with the last visited spin flipped.  Checked on the sets themselves."""
    jd = lattices[W]
    n, tet, order = _lat_arrays(jd)
    ice = q.enumerate_basis(jd, 0, n_threads=8)[:, 0]
    one = q.enumerate_basis(jd, 1, n_threads=8)[:, 0]
    flip = np.sort(ice ^ np.uint64(1 << int(order[-1])))
    assert one.size == 2 * ice.size
    assert np.array_equal(one, np.union1d(ice, flip))
