"""Batched uneven trajectories in one captured launch sequence (the config-4 shape, reduced).

96 linear-Langevin trajectories with log-uniform lengths 200 - 6,000 frames,
d = 128, lag 5, k = 512 centers: the eager pass matches the FP64 oracle
(moments) and the sequential FP32 oracle (labels); every graph replay, with
eager work and synchronizations in between, reproduces the eager results
bit for bit, also after new frames are copied into the captured input.
"""
import numpy as np
import pytest

from oracle import moments as om
from oracle import native as on
from pyemma_amd import _native, autotune
from pyemma_amd.batch import BatchedDiscretizer
from pyemma_amd.synthetic import langevin_trajectories

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=2e-5, atol_rel=1e-5):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol_rel * np.abs(b).max())


def test_batched_graph_replay_matches_eager_and_oracle():
    t = _native.torch()
    X, off = langevin_trajectories(96, 128, lmin=200, lmax=6000, seed=4)
    Xh = X.cpu().numpy()
    trajs = [Xh[a:b] for a, b in zip(off[:-1], off[1:])]
    rng = np.random.default_rng(0)
    C = Xh[np.sort(rng.choice(len(Xh), 512, replace=False))].copy()
    b = BatchedDiscretizer(X, off, 5, C)
    w, mean, C00, C0t, labels = b.run_eager()
    ref = om.covariance_lagged(trajs, 5, remove_data_mean=True, reversible=True, bessel=False)
    assert w == ref["w"]
    _close(mean, ref["mean"], rtol=1e-6)
    _close(C00, ref["C00"])
    _close(C0t, ref["C0t"])
    assert np.array_equal(labels, on.assign(Xh, C, nthreads=8))
    assert autotune.tuned(), "autotuner did not run"

    b.capture()
    for i in range(3):
        b.replay()
        t.cuda.synchronize()
        r = b.results()
        assert r[0] == w
        for x, y in zip(r[1:], (mean, C00, C0t, labels)):
            assert np.array_equal(x, y), "replay %d differs from the eager pass" % i
        # eager work of the program between replays
        _native.assign(b.X[:1000], b.C)
        t.cuda.synchronize()

    # new frames into the captured input: replay == eager on the new data
    X2, _ = langevin_trajectories(96, 128, lmin=200, lmax=6000, seed=5,
                                  lengths=np.diff(off))
    b.X.copy_(X2)
    b.replay()
    t.cuda.synchronize()
    r = b.results()
    e = BatchedDiscretizer(X2, off, 5, C).run_eager()
    assert r[0] == e[0]
    for x, y in zip(r[1:], e[1:]):
        assert np.array_equal(x, y)
