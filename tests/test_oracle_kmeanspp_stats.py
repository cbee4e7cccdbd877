"""Statistical pin of the oracle's kmeans++ seeding (oracle_kmeanspp).

The reference delegates kmeans++ to deeptime (absent from the tree; call
site coordinates/clustering/kmeans.py:254-258, entry observed at
clustering/tests/test_kmeans.py:304).  The published rule (Arthur and
Vassilvitskii 2007) is: the first center uniformly at random, every next
center x with probability D(x)^2 / sum_y D(y)^2, D the distance to the
nearest center chosen so far.  These tests check the oracle's pick
frequencies against exactly that distribution over many seeds, and show
that the plausible mistakes (weights D instead of D^2, farthest-first,
uniform) are rejected by the same test.  The GPU's picks equal the oracle's
bit for bit (tests/test_gpu_kmeans.py).
"""
import numpy as np
from scipy import stats

from oracle import native as on

# 8 frames on a line with uneven gaps: D^2 weights differ strongly from D
X8 = np.array([[0.0], [0.5], [1.5], [3.0], [3.25], [6.0], [10.0], [10.5]], dtype=np.float32)


def _picks(X, k, seeds):
    return np.array([on.kmeanspp(X, k, s)[1] for s in seeds], dtype=np.int64)


def _d2(X, centers):
    X = X.astype(np.float64)
    return np.min(((X[:, None, :] - X[None, centers, :]) ** 2).sum(-1), axis=1)


def test_uniform_stream_is_uniform():
    u = np.array([on.uniform(99, c) for c in range(100000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    # successive draws uncorrelated
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.02
    # different seeds give different streams
    v = np.array([on.uniform(100, c) for c in range(1000)])
    assert not np.allclose(u[:1000], v)


def test_first_pick_uniform():
    picks = _picks(X8, 1, range(16000))[:, 0]
    counts = np.bincount(picks, minlength=8)
    assert stats.chisquare(counts).pvalue > 1e-3


def test_second_pick_follows_d_squared():
    picks = _picks(X8, 2, range(24000))
    pvals_d2, pvals_d1 = [], []
    for first in range(8):
        sel = picks[picks[:, 0] == first, 1]
        obs = np.bincount(sel, minlength=8).astype(float)
        w2 = _d2(X8, [first])
        w1 = np.sqrt(w2)
        m = w2 > 0
        assert obs[~m].sum() == 0  # a chosen frame is never picked again
        exp2 = w2[m] / w2.sum() * obs.sum()
        exp1 = w1[m] / w1.sum() * obs.sum()
        pvals_d2.append(stats.chisquare(obs[m], exp2).pvalue)
        pvals_d1.append(stats.chisquare(obs[m], exp1).pvalue)
        # farthest-first would put every pick on the farthest frame
        assert obs.max() < obs.sum()
    # combined over the 8 first picks (Fisher); each group has ~3000 draws
    assert stats.combine_pvalues(pvals_d2).pvalue > 1e-3
    # the same test rejects D-weighting (a dropped square)
    assert stats.combine_pvalues(pvals_d1).pvalue < 1e-12


def test_third_pick_uses_nearest_chosen_center():
    """Pick 3 follows min over both chosen centers of D^2: pooled expected counts
    (sum of the conditional probabilities over the seeds) against the observed
    ones; weights from the distance to the first center only are rejected."""
    seeds = range(24000)
    picks = _picks(X8, 3, seeds)
    exp_min = np.zeros(8)
    exp_first = np.zeros(8)
    for a, b, _ in picks:
        w = _d2(X8, [a, b])
        exp_min += w / w.sum()
        w0 = _d2(X8, [a])
        w0[b] = 0.0
        exp_first += w0 / w0.sum()
    obs = np.bincount(picks[:, 2], minlength=8).astype(float)
    assert stats.chisquare(obs, exp_min).pvalue > 1e-3
    assert stats.chisquare(obs, exp_first).pvalue < 1e-12


def test_picks_on_larger_input_follow_d_squared():
    """20,000 frames (several blocks and superblocks of the pick's search
    hierarchy): the frame picked in round 1 lands in each D^2-mass decile with
    the decile's probability."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((20000, 3)).astype(np.float32) * np.array([1.0, 2.0, 0.5], np.float32)
    first = on.kmeanspp(X, 1, 0)[1][0]
    # all seeds with that first pick are not available; use the pick rule itself
    d2 = ((X.astype(np.float32) - X[first]) ** 2).sum(1).astype(np.float32)
    nb = (X.shape[0] + 4095) // 4096
    bsum = np.array([on.block_sum(d2, b * 4096, min((b + 1) * 4096, X.shape[0])) for b in range(nb)])
    u = rng.random(20000)
    picks = np.array([on.kmpp_pick(d2, bsum, x) for x in u])
    cum = np.cumsum(d2.astype(np.float64))
    cum /= cum[-1]
    # the pick is the first frame whose running prefix exceeds u * total
    np.testing.assert_array_equal(picks, np.searchsorted(cum, u, side="right"))
    dec = np.minimum((cum[picks] * 10).astype(int), 9)
    assert stats.chisquare(np.bincount(dec, minlength=10)).pvalue > 1e-3
