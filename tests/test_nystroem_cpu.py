"""oASIS column selection of Nystroem TICA on the host (no GPU).

Restates the reference's oASIS tests (pyemma/coordinates/tests/
test_nystroem_tica.py:97-168): with 5 variable and 5 constant columns, every
strategy drives the Nystroem error of X^T X to zero once the variable
columns are in; the known error [0, 0, 3] of a matrix whose two initial
columns are zero; the constant-eigenvector fix of spectral oASIS.
"""
import numpy as np
import pytest

from pyemma_amd.coordinates.nystroem import SelectionStrategySpectralOasis, oASIS_Nystroem


@pytest.fixture(scope="module")
def case():
    rng = np.random.RandomState(3)
    X = np.ones((10000, 10))
    variable = rng.choice(10, 5, replace=False)
    X[:, variable] = rng.randn(10000, 5)
    C0 = X.T.dot(X)
    initial = np.setdiff1d(np.arange(10), variable)[0:1]
    return X, variable, C0, np.diag(C0), initial


def test_random(case):
    X, variable, C0, d, initial = case
    np.random.seed(5)
    oasis = oASIS_Nystroem(d, C0[:, initial], initial)
    oasis.set_selection_strategy(strategy="random", nsel=1)
    while oasis.k < 6:
        newcol = oasis.select_columns()
        added = oasis.add_columns(C0[:, newcol], newcol)
        assert ((len(added) > 0 and np.any(np.isin(variable, added)))
                or (len(added) == 0 and not np.all(np.isin(variable, newcol))))
    assert oasis.select_columns() is None
    np.testing.assert_allclose(oasis.error, 0, rtol=1e-08, atol=1e-10)


def _single(case, strategy):
    X, variable, C0, d, initial = case
    oasis = oASIS_Nystroem(d, C0[:, initial], initial)
    oasis.set_selection_strategy(strategy=strategy, nsel=1)
    for _ in range(5):
        newcol = oasis.select_columns()
        assert oasis.add_columns(C0[:, newcol], newcol) == newcol
    assert oasis.select_columns() is None
    np.testing.assert_allclose(oasis.error, 0, rtol=1e-08, atol=1e-10)
    # the Nystroem approximation is exact once the column space is spanned
    np.testing.assert_allclose(oasis.approximate_matrix(), C0, rtol=1e-8, atol=1e-6)
    return oasis


def test_oasis_and_spectral_agree(case):
    a = _single(case, "oasis")
    b = _single(case, "spectral-oasis")
    assert np.all(a.column_indices == b.column_indices)
    assert set(a.column_indices[1:]) == set(case[1])


def test_spectral_oasis_multiple(case):
    X, variable, C0, d, initial = case
    oasis = oASIS_Nystroem(d, C0[:, initial], initial)
    oasis.set_selection_strategy(strategy="spectral-oasis", nsel=5)
    newcols = oasis.select_columns()
    assert np.all(np.equal(oasis.add_columns(C0[:, newcols], newcols), newcols))
    assert oasis.select_columns() is None
    np.testing.assert_allclose(oasis.error, 0, rtol=1e-08, atol=1e-10)


def test_initial_zero():
    X = np.zeros((3, 3))
    X[:, 2:3] = 1.0
    C0 = X.T.dot(X)
    initial = np.array([0, 1])
    oasis = oASIS_Nystroem(np.diag(C0), C0[:, initial], initial)
    np.testing.assert_allclose(np.abs(oasis.error), [0, 0, 3])
    np.testing.assert_allclose(oasis.approximate_matrix(), np.zeros((3, 3)))
    for strategy in ("random", "oasis", "spectral-oasis"):
        oasis.set_selection_strategy(strategy=strategy, nsel=1)
        assert oasis.select_columns() == np.array([2])


def test_bordered_inverse_matches_pinv(case):
    """The rank-one updates of add_column reproduce the directly computed W_k^-1 and R_k."""
    X, variable, C0, d, initial = case
    cols = np.array(list(variable[:3]))
    oasis = oASIS_Nystroem(d, C0[:, cols[:1]], cols[:1])
    oasis.add_columns(C0[:, cols[1:]], cols[1:])
    np.testing.assert_allclose(oasis.Wk_inv, np.linalg.inv(C0[np.ix_(cols, cols)]), rtol=1e-8)
    R = oasis.Rk.copy()
    oasis.update_inverse()
    np.testing.assert_allclose(R, oasis.Rk, rtol=1e-8, atol=1e-10)


def test_constant_evec():
    rng = np.random.RandomState(1)
    U = np.vstack([1 + rng.randn(1, 4) * 1e-6, np.array([1, -1, 0, 0]), np.array([0, 0, 1, -1])]).T
    M = U.dot(np.diag([1, 0.95, 0.7])).dot(U.T)
    oasis = oASIS_Nystroem(np.diag(M), M[:, 0:2], np.array([0, 1]))
    sel = SelectionStrategySpectralOasis(oasis, strategy="spectral-oasis")
    evecs = oasis.approximate_eig()[1]
    fixed = sel._fix_constant_evec(evecs)
    assert np.all(fixed[:, 0] == 1)
    assert evecs.shape == fixed.shape
