"""Nystroem TICA on the GPU moments (the reference's test_nystroem_tica.py:27-95 cases)."""
import logging

import numpy as np
import pytest

from pyemma_amd.coordinates import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def simple():
    rng = np.random.RandomState(0)
    data = np.ones((10000, 100))
    variable = rng.choice(100, 10, replace=False)
    data[:, variable] = rng.rand(10000, 10)
    initial = np.setdiff1d(np.arange(100), variable)[0:1]
    return data, variable, initial, api.tica(data, lag=1, precision="fp64")


def _compare(t, ref):
    np.testing.assert_allclose(t._oasis.error, 0, rtol=1e-05, atol=1e-08)
    np.testing.assert_allclose(t.cov, ref.cov[:, t.column_indices])
    np.testing.assert_allclose(t.cov_tau, ref.cov_tau[:, t.column_indices])
    np.testing.assert_allclose(t.timescales, ref.timescales)


def test_single(simple):
    data, variable, initial, ref = simple
    t = api.tica_nystroem(data=data, lag=1, max_columns=11, initial_columns=initial)
    _compare(t, ref)
    assert set(variable) <= set(t.column_indices)


def test_single_issues_warning(simple, caplog):
    data, variable, initial, ref = simple
    with caplog.at_level(logging.WARNING):
        t = api.tica_nystroem(data=data, lag=1, max_columns=12, initial_columns=initial)
    assert any("prematurely" in r.message for r in caplog.records)
    _compare(t, ref)


def test_multiple(simple):
    data, variable, initial, ref = simple
    t = api.tica_nystroem(data=data, lag=1, max_columns=11, nsel=3, initial_columns=initial)
    _compare(t, ref)


def test_multiple_issues_warning():
    """Fewer columns left (2) than requested per round (3): a UserWarning, then the rest are taken."""
    with pytest.warns(UserWarning):
        t = api.tica_nystroem(data=np.random.RandomState(2).rand(100, 10), lag=1, max_columns=11, nsel=3,
                              initial_columns=np.arange(8))
    assert sorted(t.column_indices) == list(range(10))


def test_describe():
    assert "NystroemTICA" in api.tica_nystroem(max_columns=2).describe()
    assert "NystroemTICA" in api.tica_nystroem(max_columns=2, data=np.random.RandomState(3).random((100, 10))) \
        .describe()


def _double_well(T=10000, nstates=100, seed=4):
    """Metropolis walk (kT = 0.5) on 100 grid points of the double well U(b) = 2 (b^2 - 1)^2."""
    rng = np.random.RandomState(seed)
    b = np.linspace(-1, 1, nstates)
    U = 2.0 * (b ** 2 - 1) ** 2
    v = np.zeros(T, dtype=int)
    s = nstates // 4
    for t in range(T):
        v[t] = s
        p = s + (1 if rng.rand() < 0.5 else -1)
        if 0 <= p < nstates and rng.rand() < np.exp(-(U[p] - U[s]) / 0.5):
            s = p
    return b, v


def test_double_well():
    """test_nystroem_tica.py:72-95 on a simulated double-well chain with Gaussian basis features
    (the reference's 2-well dataset is not part of this path): 10 of 100 columns give the three
    leading eigenvalues to 1e-3."""
    b, v = _double_well()
    sigma = 0.15
    Z = np.exp(-(b[v][:, None] - b[None, :]) ** 2 / (2 * sigma ** 2))
    lag = 10
    ref = api.tica(Z, lag=lag, precision="fp64")
    t = api.tica_nystroem(data=Z, lag=lag, max_columns=10, initial_columns=np.array([91]))
    np.testing.assert_allclose(t.cov, ref.cov[:, t.column_indices])
    np.testing.assert_allclose(t.cov_tau, ref.cov_tau[:, t.column_indices])
    assert t.column_indices.size == 10
    assert np.all(np.abs(ref.eigenvalues[:3] - t.eigenvalues[:3]) < 1e-3)
    # projection onto the selected features
    Y = t.transform(Z)
    assert Y.shape == (len(Z), t.dimension()) and Y.dtype == np.float32
    cols = t.column_indices
    Yh = ((Z - t.mean)[:, cols]).dot(np.real(t.eigenvectors[:, :t.dimension()]))
    np.testing.assert_allclose(Y, Yh, rtol=1e-4, atol=1e-5)
