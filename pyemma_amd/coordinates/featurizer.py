"""Distance features of particle coordinates on the GPU.

The producer of the distance features the moments and assignment kernels
consume (300 features are the pairwise distances of 25 particles).
Mirrors the distance part of MDFeaturizer
(pyemma/coordinates/data/featurization/featurizer.py:198-342):
`pairs`, `add_distances`, `add_inverse_distances`, `dimension`,
`describe`, `transform`.  Coordinates come as float32 arrays of shape
(n_frames, n_atoms, 3) (mdtraj's Trajectory.xyz layout) instead of mdtraj
trajectories, which are not part of this path; without unit-cell
information mdtraj computes non-periodic distances, which is what
`emma_pair_distances` computes, so `periodic` is accepted and has no effect.
"""
import itertools
import logging

import numpy as np

from .. import _native
from ._data import DeviceTrajectories

log = logging.getLogger("pyemma_amd.coordinates.featurizer")

__all__ = ["Featurizer", "featurizer"]


def _is_iterable_of_int(x):
    try:
        a = np.asarray(x)
    except Exception:
        return False
    return a.ndim == 1 and np.issubdtype(a.dtype, np.integer)


def _parse_pairwise_input(indices1, indices2, fname=""):
    """util.py:91-133."""
    if _is_iterable_of_int(indices1):
        log.warning("The 1D arrays input for %s have been sorted, and index duplicates have been eliminated.\n"
                    "Check the output of describe() to see the actual order of the features", fname)
        indices1 = np.unique(indices1)
        if indices2 is None:
            return np.array(list(itertools.combinations(indices1, 2)), dtype=np.int64).reshape(-1, 2)
        if _is_iterable_of_int(indices2):
            indices2 = np.unique(indices2)
            indices2 = indices2[np.isin(indices2, indices1, invert=True)]
            return np.array(list(itertools.product(indices1, indices2)), dtype=np.int64).reshape(-1, 2)
    return indices1


class _DistanceFeature(object):
    prefix_label = "DIST:"
    inverse = False

    def __init__(self, pairs, periodic):
        self.distance_indexes = np.asarray(pairs, dtype=np.int64)
        if len(self.distance_indexes) == 0:
            raise ValueError("empty indices")
        self.periodic = periodic

    @property
    def dimension(self):
        return len(self.distance_indexes)

    def describe(self):
        return ["%s %i - %i" % (self.prefix_label, p[0], p[1]) for p in self.distance_indexes]


class _InverseDistanceFeature(_DistanceFeature):
    prefix_label = "INVDIST:"
    inverse = True


class Featurizer(object):
    """Distance features of an n_atoms particle system (featurizer.py, distance part)."""

    def __init__(self, n_atoms):
        self.n_atoms = int(n_atoms)
        if self.n_atoms < 1:
            raise ValueError("n_atoms must be positive")
        self.active_features = []

    @staticmethod
    def pairs(sel, excluded_neighbors=0):
        """featurizer.py:198-235: non-redundant ordered pairs at least excluded_neighbors + 1 apart."""
        assert isinstance(excluded_neighbors, int)
        p = []
        for i in range(len(sel)):
            for j in range(i + 1, len(sel)):
                I, J = sel[i], sel[j]
                if I > J:
                    I, J = sel[j], sel[i]
                if J > I + excluded_neighbors:
                    p.append([I, J])
        return np.array(p)

    def _check_indices(self, pair_inds, pair_n=2):
        """featurizer.py:237-254; indices equal to n_atoms are rejected as well."""
        pair_inds = np.array(pair_inds).astype(dtype=int, casting="safe")
        if pair_inds.size == 0:
            raise ValueError("empty indices")
        if pair_inds.ndim != 2:
            raise ValueError("pair indices has to be a matrix.")
        if pair_inds.shape[1] != pair_n:
            raise ValueError("pair indices shape has to be (x, %i)." % pair_n)
        if pair_inds.max() >= self.n_atoms or pair_inds.min() < 0:
            raise ValueError("index out of bounds: %i. Maximum atom index available: %i"
                             % (pair_inds.max(), self.n_atoms - 1))
        return pair_inds

    def _add(self, f):
        if any(type(f) is type(g) and np.array_equal(f.distance_indexes, g.distance_indexes)
               for g in self.active_features):
            log.warning("tried to re-add the same feature %s", f.__class__.__name__)
            return
        self.active_features.append(f)

    def add_distances(self, indices, periodic=True, indices2=None):
        """featurizer.py:305-342."""
        pairs = self._check_indices(_parse_pairwise_input(indices, indices2, fname="add_distances()"))
        self._add(_DistanceFeature(pairs, periodic))

    def add_inverse_distances(self, indices, periodic=True, indices2=None):
        pairs = self._check_indices(_parse_pairwise_input(indices, indices2, fname="add_inverse_distances()"))
        self._add(_InverseDistanceFeature(pairs, periodic))

    def dimension(self):
        return int(sum(f.dimension for f in self.active_features))

    def describe(self):
        return [d for f in self.active_features for d in f.describe()]

    def transform_device(self, xyz):
        """(n_frames, n_atoms, 3) coordinates (torch device tensor or array) -> device tensor (n_frames, dim)."""
        t = _native.torch()
        if not self.active_features:
            raise ValueError("no features added")
        if not isinstance(xyz, t.Tensor):
            xyz = t.as_tensor(np.ascontiguousarray(xyz, dtype=np.float32), device=_native.device())
        if xyz.dtype != t.float32 or xyz.dim() != 3 or xyz.shape[1] != self.n_atoms or xyz.shape[2] != 3:
            raise ValueError("coordinates must be float32 of shape (n_frames, %d, 3)" % self.n_atoms)
        xyz = xyz.contiguous()
        n = xyz.shape[0]
        outs = []
        for f in self.active_features:
            pairs = t.as_tensor(f.distance_indexes.astype(np.int32), device=xyz.device).contiguous()
            Y = t.empty((n, f.dimension), dtype=t.float32, device=xyz.device)
            _native.check(_native.lib().emma_pair_distances(
                _native.ptr(xyz), n, self.n_atoms, _native.ptr(pairs), f.dimension,
                _native.EMMA_DIST_INVERSE if f.inverse else 0, _native.ptr(Y), _native.stream_ptr()),
                "emma_pair_distances")
            outs.append(Y)
        return outs[0] if len(outs) == 1 else t.cat(outs, dim=1)

    def transform_device_into(self, xyz, out):
        """Features of a contiguous (n_frames, n_atoms, 3) device tensor written
        into out, a C-contiguous (n_frames, dim) float32 device tensor (rows of a
        larger buffer); one feature kind is written in place, several are
        concatenated first."""
        if len(self.active_features) != 1:
            out.copy_(self.transform_device(xyz))
            return out
        f = self.active_features[0]
        t = _native.torch()
        if xyz.dtype != t.float32 or xyz.dim() != 3 or xyz.shape[1] != self.n_atoms or xyz.shape[2] != 3:
            raise ValueError("coordinates must be float32 of shape (n_frames, %d, 3)" % self.n_atoms)
        if out.dtype != t.float32 or tuple(out.shape) != (xyz.shape[0], f.dimension) or not out.is_contiguous():
            raise ValueError("out must be a contiguous float32 tensor of shape (%d, %d)" % (xyz.shape[0], f.dimension))
        pairs = getattr(f, "_pairs_dev", None)
        if pairs is None or pairs.device != xyz.device:
            # uploaded once: an upload from pageable host memory waits for the stream
            pairs = f._pairs_dev = t.as_tensor(f.distance_indexes.astype(np.int32), device=xyz.device).contiguous()
        _native.check(_native.lib().emma_pair_distances(
            _native.ptr(xyz), int(xyz.shape[0]), self.n_atoms, _native.ptr(pairs), f.dimension,
            _native.EMMA_DIST_INVERSE if f.inverse else 0, _native.ptr(out), _native.stream_ptr()),
            "emma_pair_distances")
        return out

    def transform(self, xyz):
        """Host float32 features of one (n_frames, n_atoms, 3) array, or a list of them."""
        if isinstance(xyz, (list, tuple)):
            return [self.transform(x) for x in xyz]
        return self.transform_device(xyz).cpu().numpy()

    def transform_trajectories(self, xyz_list):
        """Features of several trajectories kept on the device (DeviceTrajectories), the input
        of tica / covariance_lagged / cluster_kmeans / assign_to_centers."""
        t = _native.torch()
        Ys = [self.transform_device(x) for x in xyz_list]
        off = np.concatenate([[0], np.cumsum([y.shape[0] for y in Ys])]).astype(np.int64)
        return DeviceTrajectories(t.cat(Ys, dim=0).contiguous(), off)


def featurizer(n_atoms):
    """coordinates.featurizer(topfile) analogue for bare coordinate arrays."""
    return Featurizer(n_atoms)
