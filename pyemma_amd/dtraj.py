"""Transition counts of discrete trajectories on the GPU.

The consumer of the assignment labels: PyEMMA hands the dtrajs that
`assign_to_centers` / `cluster_kmeans(...).dtrajs` produce to its MSM layer,
whose DiscreteTrajectoryStats (msm/estimators/_dtraj_stats.py:25-167) takes
the state histogram, the number of states and the lagged transition count
matrix from deeptime (`count_states`, `number_of_states`, `count_matrix`,
_dtraj_stats.py:60-64, 138-145).  Those three run here as HIP kernels
(`emma_count_transitions`, `emma_max_state`); labels already on the device
(from `emma_assign`) are counted without a host round trip.

Counting rule: per trajectory, the pairs (s[t], s[t + lag]) for every
t < T - lag (count_mode 'sliding') or t = 0, lag, 2 lag, ... ('sample');
frames with state -1 (unassigned) are left out.  The dense count matrix is
returned as float64 (a scipy CSR matrix when sparse_return is True).
"""
import numpy as np

from . import _native

__all__ = ["DeviceDtrajs", "to_device_dtrajs", "number_of_states", "count_states", "count_matrix",
           "DiscreteTrajectoryStats"]


class DeviceDtrajs(object):
    """Row-concatenated int32 state trajectories on the GPU plus CSR offsets."""

    def __init__(self, S, offsets):
        t = _native.torch()
        self.S = S
        self.offsets = np.asarray(offsets, dtype=np.int64)
        if self.offsets.ndim != 1 or self.offsets.size < 2 or self.offsets[0] != 0 \
                or np.any(np.diff(self.offsets) < 0) or self.offsets[-1] != S.numel():
            raise ValueError("offsets must be non-decreasing CSR offsets from 0 to the number of frames")
        self.offsets_dev = t.as_tensor(self.offsets, device=S.device)

    @property
    def ntraj(self):
        return len(self.offsets) - 1


def to_device_dtrajs(dtrajs):
    """A list of integer arrays (or one array) -> DeviceDtrajs; DeviceDtrajs pass through."""
    if isinstance(dtrajs, DeviceDtrajs):
        return dtrajs
    t = _native.torch()
    if isinstance(dtrajs, np.ndarray) and dtrajs.ndim == 1:
        dtrajs = [dtrajs]
    arrs = []
    for d in dtrajs:
        d = np.asarray(d)
        if d.ndim != 1:
            d = d.reshape(-1)
        if d.size and not np.issubdtype(d.dtype, np.integer):
            raise ValueError("discrete trajectories must hold integers (got %s)" % d.dtype)
        if d.size and (d.min() < np.iinfo(np.int32).min or d.max() > np.iinfo(np.int32).max):
            raise ValueError("state index out of the int32 range")
        arrs.append(d.astype(np.int32, copy=False))
    if not arrs:
        raise ValueError("no discrete trajectories given")
    off = np.concatenate([[0], np.cumsum([a.size for a in arrs])]).astype(np.int64)
    S = t.as_tensor(np.concatenate(arrs) if off[-1] else np.zeros(0, np.int32), device=_native.device())
    return DeviceDtrajs(S, off)


def number_of_states(dtrajs):
    """Largest state + 1 (deeptime number_of_states, used at _dtraj_stats.py:64)."""
    dd = to_device_dtrajs(dtrajs)
    return int(_native.max_state(dd.S).item()) + 1


def _counts(dd, lag, sliding, nstates):
    C, H, bad = _native.count_transitions(dd.S, dd.offsets_dev, lag, sliding, nstates)
    if int(bad.item()) != 0:
        raise ValueError("Discrete trajectory contains elements < -1 or >= nstates (%d)." % nstates)
    return C, H


def count_states(dtrajs, ignore_negative=True, nstates=None):
    """Frames per state (deeptime count_states, _dtraj_stats.py:60)."""
    dd = to_device_dtrajs(dtrajs)
    n = number_of_states(dd) if nstates is None else int(nstates)
    if not ignore_negative and dd.S.numel() and int(_native.torch().min(dd.S).item()) < 0:
        raise ValueError("negative states with ignore_negative=False")
    if n <= 0:
        return np.zeros(0, dtype=np.int64)
    _, H = _counts(dd, 1, True, n)
    return H.cpu().numpy()


def count_matrix(dtrajs, lag, sliding=True, sparse_return=True, nstates=None):
    """Transition count matrix at `lag` (deeptime count_matrix, _dtraj_stats.py:138-145)."""
    import scipy.sparse

    lag = int(lag)
    if lag < 1:
        raise ValueError("lag must be a positive integer")
    dd = to_device_dtrajs(dtrajs)
    n = number_of_states(dd) if nstates is None else int(nstates)
    if n <= 0:
        raise ValueError("no states: every frame is unassigned")
    C, _ = _counts(dd, lag, sliding, n)
    C = C.cpu().numpy().astype(np.float64)
    return scipy.sparse.csr_matrix(C) if sparse_return else C


class DiscreteTrajectoryStats(object):
    """Count statistics of discrete trajectories (_dtraj_stats.py:25-167, counting part).

    Connectivity analysis, effective counts and core-set milestoning stay in
    the MSM package; this class provides the histogram, the number of states
    and the lagged count matrix in 'sliding' and 'sample' mode.
    """

    def __init__(self, dtrajs):
        self._dd = to_device_dtrajs(dtrajs)
        t = _native.torch()
        if self._dd.S.numel() and int(t.min(self._dd.S).item()) < -1:
            raise ValueError("Discrete trajectory contains elements < -1.")
        self._nstates = number_of_states(self._dd)
        self._hist = count_states(self._dd, nstates=max(self._nstates, 1))
        self._total_count = int(np.sum(self._hist))
        self._counted_at_lag = False
        self._C = None
        self._lag = None

    def count_lagged(self, lag, count_mode="sliding"):
        count_mode = count_mode.lower()
        if count_mode not in ("sliding", "sample"):
            raise NotImplementedError("count_mode %r: only 'sliding' and 'sample' are implemented on the GPU"
                                      % count_mode)
        self._C = count_matrix(self._dd, lag, sliding=count_mode == "sliding", nstates=self._nstates)
        self._lag = int(lag)
        self._counted_at_lag = True
        return self

    def count_matrix(self):
        assert self._counted_at_lag, \
            "You haven't run count_lagged yet. Do that first before accessing lag-based quantities"
        return self._C

    @property
    def lagtime(self):
        return self._lag

    @property
    def nstates(self):
        return self._nstates

    @property
    def total_count(self):
        return self._total_count

    @property
    def histogram(self):
        return self._hist

    @property
    def visited_set(self):
        return np.argwhere(self._hist > 0)[:, 0]
