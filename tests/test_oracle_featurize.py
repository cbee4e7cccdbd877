"""Pins the FP32 distance oracle and the pair rules of the featurizer (no GPU)."""
import numpy as np
import pytest

from oracle.featurize import pair_distances
from pyemma_amd.coordinates.featurizer import Featurizer, _parse_pairwise_input

# integer distances: (a, b, c) with a^2 + b^2 + c^2 = d^2
QUADRUPLES = [(1, 2, 2, 3), (2, 3, 6, 7), (1, 4, 8, 9), (4, 4, 7, 9), (2, 6, 9, 11), (6, 6, 7, 11), (3, 4, 12, 13)]


def test_exact_integer_distances():
    xyz = np.zeros((len(QUADRUPLES), 2, 3), dtype=np.float32)
    for f, (a, b, c, _) in enumerate(QUADRUPLES):
        xyz[f, 0] = (0.5, -1.0, 2.0)
        xyz[f, 1] = (0.5 + a, -1.0 - b, 2.0 + c)
    d = pair_distances(xyz, [[0, 1]])
    np.testing.assert_array_equal(d[:, 0], [q[3] for q in QUADRUPLES])
    np.testing.assert_array_equal(pair_distances(xyz, [[1, 0]]), d)
    np.testing.assert_array_equal(pair_distances(xyz, [[0, 1]], inverse=True)[:, 0],
                                  np.float32(1.0) / np.array([q[3] for q in QUADRUPLES], dtype=np.float32))


def test_close_to_fp64():
    rng = np.random.RandomState(0)
    xyz = (rng.randn(500, 12, 3) * 3).astype(np.float32)
    pairs = np.array(list(zip(*np.triu_indices(12, 1))))
    d32 = pair_distances(xyz, pairs)
    x = xyz.astype(np.float64)
    d64 = np.linalg.norm(x[:, pairs[:, 1]] - x[:, pairs[:, 0]], axis=-1)
    # five roundings of values of the size of d: a few ulps of d
    assert np.all(np.abs(d32 - d64) <= 4 * np.spacing(d64.astype(np.float32)))


def test_pairs_rules():
    """featurizer.py:198-235 and util.py:91-133."""
    np.testing.assert_array_equal(Featurizer.pairs(np.array([0, 1, 2, 3]), excluded_neighbors=1),
                                  [[0, 2], [0, 3], [1, 3]])
    np.testing.assert_array_equal(Featurizer.pairs(np.array([3, 1, 0])), [[1, 3], [0, 3], [0, 1]])
    np.testing.assert_array_equal(_parse_pairwise_input([3, 1, 1, 2], None), [[1, 2], [1, 3], [2, 3]])
    np.testing.assert_array_equal(_parse_pairwise_input([2, 0], [5, 0, 4]), [[0, 4], [0, 5], [2, 4], [2, 5]])
    pairs = np.array([[0, 5], [2, 3]])
    np.testing.assert_array_equal(_parse_pairwise_input(pairs, None), pairs)


def test_index_checks():
    f = Featurizer(5)
    with pytest.raises(ValueError):
        f.add_distances(np.array([[0, 5]]))
    with pytest.raises(ValueError):
        f.add_distances(np.array([0, 1, 2]).reshape(1, 3))
    with pytest.raises(ValueError):  # indices2 minus indices1 is empty
        f.add_distances(np.arange(5), indices2=[1, 2])
    f.add_distances(np.arange(5))
    assert f.dimension() == 10
    f.add_distances(np.arange(5))  # the same feature is not added twice
    assert f.dimension() == 10
    f.add_inverse_distances([[0, 1]])
    assert f.dimension() == 11
    assert f.describe()[0] == "DIST: 0 - 1" and f.describe()[-1] == "INVDIST: 0 - 1"
