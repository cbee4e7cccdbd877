"""Transition counts on the GPU (emma_count_transitions) against the oracle.

Bit-exact integer work: every comparison is equality.
* the reference's own count tensor (msm/tests/data/TestData_OOM_MSM.npz,
  C2t summed over its last axis = the lag-10 sliding counts of the stored
  trajectories) and the known answer of msm/tests/test_dtraj_stats.py:46-57;
* sliding and sample mode at several lags against oracle/counts.py, with
  ragged trajectories, empty ones, trajectories shorter than the lag and -1
  (unassigned) frames;
* long metastable runs (one atomic per run of equal pairs) and a full-size
  case (200 x 50,000 frames, 1000 states) with the pair-count closed form;
* labels kept on the device from assignment give the same counts as the
  host dtrajs;
* entries below -1 or at/above nstates raise ValueError.
"""
import os

import numpy as np
import pytest

from oracle import counts as oc
from pyemma_amd import dtraj as pd

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def oom():
    f = np.load(os.path.join(GOLDEN, "oom_counts.npz"), allow_pickle=False)
    return [d.astype(np.int32) for d in f["dtrajs"]], f["C2t"]


def test_reference_count_tensor(oom):
    dtrajs, C2t = oom
    C = pd.count_matrix(dtrajs, 10, sliding=True, sparse_return=False)
    np.testing.assert_array_equal(C, C2t.sum(axis=2))
    Cs = pd.count_matrix(dtrajs, 10)
    np.testing.assert_array_equal(Cs.toarray(), C2t.sum(axis=2))


@pytest.mark.parametrize("sliding", [True, False])
@pytest.mark.parametrize("lag", [1, 2, 5, 13, 64, 199, 200])
def test_oom_trajectories_vs_oracle(oom, lag, sliding):
    dtrajs, _ = oom
    np.testing.assert_array_equal(pd.count_matrix(dtrajs, lag, sliding=sliding, sparse_return=False),
                                  oc.count_matrix(dtrajs, lag, sliding=sliding, nstates=5))


def test_reference_known_answer():
    d = np.zeros(10, dtype=np.int64)
    d[0] = 1
    d[-1] = 1
    np.testing.assert_array_equal(pd.count_matrix(d, 1, sparse_return=False), [[7, 1], [1, 0]])
    s = pd.DiscreteTrajectoryStats(d)
    assert s.nstates == 2 and s.total_count == 10
    np.testing.assert_array_equal(s.histogram, [8, 2])
    s.count_lagged(1)
    np.testing.assert_array_equal(s.count_matrix().toarray(), [[7, 1], [1, 0]])


@pytest.mark.parametrize("sliding", [True, False])
@pytest.mark.parametrize("lag", [1, 3, 7, 70])
def test_ragged_with_unassigned(lag, sliding):
    rng = np.random.RandomState(100 + lag)
    lens = [0, 1, 5, 7, 8, 63, 64, 65, 129, 1000, 0, 3, 4097]
    dtrajs = [rng.randint(-1, 9, size=n).astype(np.int32) for n in lens]
    got = pd.count_matrix(dtrajs, lag, sliding=sliding, sparse_return=False, nstates=9)
    np.testing.assert_array_equal(got, oc.count_matrix(dtrajs, lag, sliding=sliding, nstates=9))
    np.testing.assert_array_equal(pd.count_states(dtrajs, nstates=9), oc.count_states(dtrajs, nstates=9))


def test_metastable_runs():
    """Long runs of one pair: one atomic per run of equal lanes."""
    rng = np.random.RandomState(5)
    dtrajs = []
    for n in (1_000_003, 250_000, 77):
        runs = rng.geometric(1.0 / 2000.0, size=n // 10 + 10)
        states = rng.randint(0, 50, size=runs.size)
        dtrajs.append(np.repeat(states, runs)[:n].astype(np.int32))
    for lag in (1, 10, 100):
        for sliding in (True, False):
            np.testing.assert_array_equal(
                pd.count_matrix(dtrajs, lag, sliding=sliding, sparse_return=False),
                oc.count_matrix(dtrajs, lag, sliding=sliding))


def test_full_size_closed_form():
    """200 x 50,000 frames over 1000 states: total = sum(T - lag); rows match the oracle."""
    rng = np.random.RandomState(9)
    dtrajs = [rng.randint(0, 1000, size=50000).astype(np.int32) for _ in range(200)]
    for lag in (1, 20):
        C = pd.count_matrix(dtrajs, lag, sparse_return=False)
        assert C.sum() == 200 * (50000 - lag)
        np.testing.assert_array_equal(C, oc.count_matrix(dtrajs, lag))
    np.testing.assert_array_equal(pd.count_states(dtrajs), oc.count_states(dtrajs))


def test_device_labels_from_assignment():
    from pyemma_amd.coordinates import api

    rng = np.random.RandomState(2)
    X = [rng.randn(n, 8).astype(np.float32) for n in (3000, 1200, 17)]
    centers = rng.randn(12, 8).astype(np.float32)
    ac = api.assign_to_centers(X, centers, return_dtrajs=False)
    host = ac.dtrajs
    dev = ac.assign_device(X)
    for lag in (1, 4):
        np.testing.assert_array_equal(pd.count_matrix(dev, lag, sparse_return=False, nstates=12),
                                      oc.count_matrix(host, lag, nstates=12))


def test_out_of_range_raises():
    with pytest.raises(ValueError):
        pd.count_matrix([np.array([0, 1, -2, 1], dtype=np.int32)], 1)
    with pytest.raises(ValueError):
        pd.count_matrix([np.array([0, 1, 5, 1], dtype=np.int32)], 1, nstates=3)
    with pytest.raises(ValueError):
        pd.DiscreteTrajectoryStats([np.array([0, -3], dtype=np.int32)])
    assert pd.number_of_states([np.array([3, 1], dtype=np.int32), np.array([7], dtype=np.int32)]) == 8
