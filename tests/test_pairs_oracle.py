"""The oracle's brute-force pair finder against an independent k-d tree (scipy).

Ties at exactly the cutoff have measure zero for these random inputs, so the two sets
must agree exactly.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import pyoracle as po


def _kd_pairs(xyz, cutoff, box=None):
    if box is not None:
        # cKDTree's periodic mode wants coordinates in [0, L)
        t = cKDTree(np.mod(xyz + box / 2, box), boxsize=box)
    else:
        t = cKDTree(xyz)
    p = np.array(sorted(t.query_pairs(cutoff)), dtype=np.int64).reshape(-1, 2)
    return p


@pytest.mark.parametrize("N,cutoff", [(1, 3.0), (2, 100.0), (200, 5.0), (500, 12.0)])
def test_pairs_free_space(N, cutoff):
    rng = np.random.default_rng(N)
    xyz = rng.uniform(-30, 30, size=(N, 3))
    got = po.pairs_within(xyz, cutoff)
    assert np.array_equal(got, _kd_pairs(xyz, cutoff))


@pytest.mark.parametrize("N,box,cutoff", [(300, 50.0, 7.0), (100, 40.0, 15.0), (50, 20.0, 9.9)])
def test_pairs_periodic(N, box, cutoff):
    rng = np.random.default_rng(N + 1)
    xyz = rng.uniform(-box / 2, box / 2, size=(N, 3))
    xyz[:5] += box * rng.integers(-2, 3, size=(5, 3))   # a few beads outside the cell
    got = po.pairs_within(xyz, cutoff, pbc=True, box=box)
    assert np.array_equal(got, _kd_pairs(xyz, cutoff, box))


def test_pairs_planted_index_error_is_caught():
    """A pair finder off by one in the inner index differs from the k-d tree."""
    rng = np.random.default_rng(5)
    xyz = rng.uniform(-10, 10, size=(60, 3))
    good = po.pairs_within(xyz, 6.0)
    bad = good.copy()
    bad[:, 1] = np.minimum(bad[:, 1] + 1, 59)
    assert not np.array_equal(bad, _kd_pairs(xyz, 6.0))
    assert np.array_equal(good, _kd_pairs(xyz, 6.0))
