"""Pairwise-distance features of particle coordinates, in FP32.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates the non-periodic distance of mdtraj.compute_distances as
DistanceFeature.transform / InverseDistanceFeature.transform call it
(pyemma/coordinates/data/featurization/distances.py:53-54, 72-73): for pair
(i, j), d = x_j - x_i per component, then sqrt((dx*dx + dy*dy) + dz*dz), every
step a correctly rounded FP32 operation (mdtraj's pure-NumPy path
``(np.diff(xyz[:, pairs], axis=2)[:, :, 0] ** 2).sum(-1) ** 0.5`` on float32
coordinates sums the three squares left to right).  mdtraj is not present
here, so its SIMD path (same operations, horizontal adds in the same order)
is not run.

Pinned by (tests/test_oracle_featurize.py): Pythagorean quadruples whose
distances are exact integers, the FP64 distance (within a few FP32 ulps),
and the ordering and pair rules of Featurizer.pairs / _parse_pairwise_input
(featurizer.py:198-235, util.py:91-133) on small index sets.
"""
import numpy as np

__all__ = ["pair_distances"]


def pair_distances(xyz, pairs, inverse=False):
    xyz = np.asarray(xyz, dtype=np.float32)
    pairs = np.asarray(pairs, dtype=np.int64)
    a = xyz[:, pairs[:, 0], :]
    b = xyz[:, pairs[:, 1], :]
    d = b - a
    s = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    s = s + d[..., 2] * d[..., 2]
    r = np.sqrt(s)
    if inverse:
        r = np.float32(1.0) / r
    return r.astype(np.float32)
