"""Transition counts of discrete trajectories (restated count_matrix).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates deeptime's ``count_matrix(dtrajs, lag, sliding, nstates)`` and
``count_states`` as PyEMMA calls them (msm/estimators/_dtraj_stats.py:60-64,
138-145).  deeptime is not present here, so its published rule is restated:
per trajectory of length T, the pairs (s[t], s[t + lag]) for every
t < T - lag (sliding window) or for t = 0, lag, 2 lag, ... (sample mode);
trajectories with T <= lag contribute nothing; the matrix is nstates x
nstates with nstates = max state + 1 unless given.  Frames with state -1
(unassigned, _dtraj_stats.py:33-36) are left out of pairs and histogram.

Pinned by (tests/test_oracle_counts.py):
  * the reference's own count tensor of msm/tests/data/TestData_OOM_MSM.npz
    (C2t summed over its last axis = the lag-10 sliding counts of the 1000
    stored trajectories), exact;
  * the known answer of msm/tests/test_dtraj_stats.py:46-57
    (zeros(10) with the first and last frame 1, lag 1 -> [[7, 1], [1, 0]]);
  * explicit pair loops on small random inputs, and the pair-count closed
    forms sum(C) = sum_i (T_i - lag) (sliding) and
    sum_i ceil((T_i - lag) / lag) (sample).
Parity unpinned: the treatment of -1 frames (deeptime absent; this module
drops every pair that touches one) and the dtype of the returned matrix
(float64 here).
"""
import numpy as np

__all__ = ["number_of_states", "count_states", "count_matrix"]


def _as_list(dtrajs):
    if isinstance(dtrajs, np.ndarray) and dtrajs.ndim == 1:
        return [dtrajs]
    return [np.asarray(d).ravel() for d in dtrajs]


def number_of_states(dtrajs):
    return int(max(int(d.max()) if d.size else -1 for d in _as_list(dtrajs))) + 1


def count_states(dtrajs, nstates=None):
    dt = _as_list(dtrajs)
    n = number_of_states(dt) if nstates is None else nstates
    h = np.zeros(n, dtype=np.int64)
    for d in dt:
        d = d[d >= 0]
        np.add.at(h, d, 1)
    return h


def count_matrix(dtrajs, lag, sliding=True, nstates=None):
    dt = _as_list(dtrajs)
    n = number_of_states(dt) if nstates is None else nstates
    C = np.zeros((n, n), dtype=np.float64)
    for d in dt:
        if d.size <= lag:
            continue
        if sliding:
            r, c = d[:-lag], d[lag:]
        else:
            r, c = d[:-lag:lag], d[lag::lag]
        keep = (r >= 0) & (c >= 0)
        np.add.at(C, (r[keep], c[keep]), 1.0)
    return C
