"""Pins the count-matrix oracle (oracle/counts.py) without a GPU."""
import os

import numpy as np
import pytest

from oracle import counts as oc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def oom():
    f = np.load(os.path.join(GOLDEN, "oom_counts.npz"), allow_pickle=False)
    return [d.astype(np.int64) for d in f["dtrajs"]], f["C2t"]


def test_reference_count_tensor(oom):
    """TestData_OOM_MSM.npz: C2t summed over the last axis = lag-10 sliding counts."""
    dtrajs, C2t = oom
    np.testing.assert_array_equal(oc.count_matrix(dtrajs, 10, sliding=True), C2t.sum(axis=2))


def test_reference_mincount_known_answer():
    """msm/tests/test_dtraj_stats.py:46-57."""
    d = np.zeros(10, dtype=np.int64)
    d[0] = 1
    d[-1] = 1
    np.testing.assert_array_equal(oc.count_matrix(d, 1), np.array([[7, 1], [1, 0]]))
    np.testing.assert_array_equal(oc.count_states(d), [8, 2])


def _loop_counts(dtrajs, lag, sliding, n):
    C = np.zeros((n, n))
    for d in dtrajs:
        step = 1 if sliding else lag
        for t in range(0, len(d) - lag, step):
            if d[t] >= 0 and d[t + lag] >= 0:
                C[d[t], d[t + lag]] += 1
    return C


@pytest.mark.parametrize("sliding", [True, False])
@pytest.mark.parametrize("lag", [1, 3, 7])
def test_against_pair_loop(sliding, lag):
    rng = np.random.RandomState(lag)
    dtrajs = [rng.randint(-1, 6, size=n) for n in (1, 5, 7, 8, 40, 97)]
    np.testing.assert_array_equal(oc.count_matrix(dtrajs, lag, sliding=sliding, nstates=6),
                                  _loop_counts(dtrajs, lag, sliding, 6))


def test_pair_count_closed_forms(oom):
    dtrajs, _ = oom
    for lag in (1, 5, 13):
        T = np.array([len(d) for d in dtrajs])
        assert oc.count_matrix(dtrajs, lag, sliding=True).sum() == np.sum(np.maximum(T - lag, 0))
        assert oc.count_matrix(dtrajs, lag, sliding=False).sum() == np.sum(-(-np.maximum(T - lag, 0) // lag))
    assert oc.count_states(dtrajs).sum() == sum(len(d) for d in dtrajs)
    assert oc.number_of_states(dtrajs) == 5
