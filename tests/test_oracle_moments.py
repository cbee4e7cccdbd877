"""Pins the FP64 moment oracle (oracle/moments.py).

Checks come from the reference's own tests (closed forms) and from exact
rational arithmetic on tiny inputs, so that a dropped term, a wrong lag index
or a transposed operand in the restatement fails.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import moments as om


def _closed_forms(data, lag):
    """test_covar_estimator.py:19-90 and test_tica.py:203-230, written out."""
    X, Y = data[:-lag], data[lag:]
    w = X.shape[0]
    sym_s = X.sum(0) + Y.sum(0)
    m_sym = sym_s / (2.0 * w)
    X0s, Y0s = X - m_sym, Y - m_sym
    return dict(
        C00_sym=(X0s.T @ X0s + Y0s.T @ Y0s) / (2.0 * w),
        C0t_sym=(X0s.T @ Y0s + Y0s.T @ X0s) / (2.0 * w),
        mean_sym=m_sym,
        C00_nr=((X - X.mean(0)).T @ (X - X.mean(0))) / w,
        C0t_nr=((X - X.mean(0)).T @ (Y - Y.mean(0))) / w,
        Ctt_nr=((Y - Y.mean(0)).T @ (Y - Y.mean(0))) / w,
        mean_nr=X.mean(0),
        Mxx_raw=X.T @ X / w,
        Mxy_raw=X.T @ Y / w,
    )


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(5)
    return rng.random((5000, 3)) + np.array([3.0, -1.0, 0.5])


def test_reversible_mean_free_matches_closed_form(data):
    ref = _closed_forms(data, 10)
    m = om.covariance_lagged([data], 10, remove_data_mean=True, reversible=True, bessel=False, chunksize=500)
    np.testing.assert_allclose(m["C00"], ref["C00_sym"], rtol=1e-10)
    np.testing.assert_allclose(m["C0t"], ref["C0t_sym"], rtol=1e-10)
    np.testing.assert_allclose(m["mean"], ref["mean_sym"], rtol=1e-12)


def test_nonreversible_mean_free_matches_closed_form(data):
    ref = _closed_forms(data, 10)
    m = om.covariance_lagged([data], 10, ctt=True, remove_data_mean=True, reversible=False, bessel=False,
                             chunksize=500)
    np.testing.assert_allclose(m["C00"], ref["C00_nr"], rtol=1e-10)
    np.testing.assert_allclose(m["C0t"], ref["C0t_nr"], rtol=1e-10)
    np.testing.assert_allclose(m["Ctt"], ref["Ctt_nr"], rtol=1e-10)
    np.testing.assert_allclose(m["mean"], ref["mean_nr"], rtol=1e-12)


def test_raw_moments_and_bessel(data):
    ref = _closed_forms(data, 10)
    m = om.covariance_lagged([data], 10, remove_data_mean=False, reversible=False, bessel=False, chunksize=777)
    np.testing.assert_allclose(m["C00"], ref["Mxx_raw"], rtol=1e-10)
    np.testing.assert_allclose(m["C0t"], ref["Mxy_raw"], rtol=1e-10)
    mb = om.covariance_lagged([data], 10, remove_data_mean=False, reversible=False, bessel=True)
    w = data.shape[0] - 10
    np.testing.assert_allclose(mb["C00"], ref["Mxx_raw"] * w / (w - 1), rtol=1e-10)


def test_chunk_size_invariance():
    """test_tica.py:143-163: mean and covariance independent of the chunking."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((23000, 3))
    a = om.covariance_lagged([X], 100, remove_data_mean=True, reversible=True, bessel=False)
    b = om.covariance_lagged([X], 100, remove_data_mean=True, reversible=True, bessel=False, chunksize=40)
    np.testing.assert_allclose(a["mean"], b["mean"], rtol=1e-12)
    np.testing.assert_allclose(a["C00"], b["C00"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(a["C0t"], b["C0t"], rtol=1e-10, atol=1e-14)


def _exact_pairs_moments(trajs, lag, stride, skip, reversible):
    """Brute force over every (x_t, x_{t+lag}) pair in rational arithmetic."""
    pairs = []
    for tr in trajs:
        T = len(tr)
        i = 0
        while skip + lag + i * stride < T:
            pairs.append(([Fraction(float(v)) for v in tr[skip + i * stride]],
                          [Fraction(float(v)) for v in tr[skip + lag + i * stride]]))
            i += 1
    d = len(trajs[0][0])
    P = len(pairs)
    if reversible:
        w = 2 * P
        mean = [sum(p[0][a] + p[1][a] for p in pairs) / w for a in range(d)]
        C00 = [[sum((p[0][a] - mean[a]) * (p[0][b] - mean[b]) + (p[1][a] - mean[a]) * (p[1][b] - mean[b])
                    for p in pairs) / w for b in range(d)] for a in range(d)]
        C0t = [[sum((p[0][a] - mean[a]) * (p[1][b] - mean[b]) + (p[1][a] - mean[a]) * (p[0][b] - mean[b])
                    for p in pairs) / w for b in range(d)] for a in range(d)]
    else:
        w = P
        mean = [sum(p[0][a] for p in pairs) / w for a in range(d)]
        my = [sum(p[1][a] for p in pairs) / w for a in range(d)]
        C00 = [[sum((p[0][a] - mean[a]) * (p[0][b] - mean[b]) for p in pairs) / w for b in range(d)]
               for a in range(d)]
        C0t = [[sum((p[0][a] - mean[a]) * (p[1][b] - my[b]) for p in pairs) / w for b in range(d)]
               for a in range(d)]
    f = lambda M: np.array([[float(v) for v in row] for row in M])
    return np.array([float(v) for v in mean]), f(C00), f(C0t)


@pytest.mark.parametrize("reversible", [True, False])
@pytest.mark.parametrize("lag,stride,skip", [(1, 1, 0), (3, 2, 1), (2, 3, 4)])
def test_exact_pairs_multi_trajectory(reversible, lag, stride, skip):
    rng = np.random.default_rng(17)
    trajs = [rng.integers(-8, 9, size=(T, 3)).astype(np.float64) / 4.0 for T in (13, 9, 4, 11)]
    mean, C00, C0t = _exact_pairs_moments(trajs, lag, stride, skip, reversible)
    m = om.covariance_lagged(trajs, lag, remove_data_mean=True, reversible=reversible, bessel=False,
                             stride=stride, skip=skip, chunksize=2)
    np.testing.assert_allclose(m["mean"], mean, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(m["C00"], C00, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(m["C0t"], C0t, rtol=1e-12, atol=1e-13)


def test_pairing_rule():
    """iterable.py:152-160: X = traj[skip::stride], Y = traj[skip+lag::stride], equal length."""
    tr = np.arange(20)[:, None]
    X, Y = om.lagged_pairs(tr, 3, stride=4, skip=2)
    assert X.ravel().tolist() == [2, 6, 10, 14]
    assert Y.ravel().tolist() == [5, 9, 13, 17]
    assert om.n_pairs(5, 5) == 0 and om.n_pairs(6, 5) == 1


def test_too_short_raises():
    """covariance.py:181-190 / test_tica.py:173-176."""
    with pytest.raises(ValueError):
        om.covariance_lagged([np.zeros((100, 1))], 100)
