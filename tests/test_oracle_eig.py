"""Pins the FP64 eigen oracle and the product's host eig_corr restatement.

Known answers: the 3x3 matrices of solvers/tests/test_direct.py:18-71; the
identity C0^-1 Ct = R diag(l) R^-1.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import eig as oeig
from pyemma_amd import _solvers
from pyemma_amd.exceptions import ZeroRankError

C0 = np.array([[1.0, 0.3, 0.2], [0.3, 0.8, 0.5], [0.2, 0.5, 0.9]])
CT_SYM = np.array([[0.5, 0.1, 0.0], [0.1, 0.3, 0.3], [0.0, 0.3, 0.2]])
CT_NONSYM = np.array([[0.5, 0.1, 0.3], [0.1, 0.3, 0.3], [0.0, 0.3, 0.2]])


def _sort(evals, evecs):
    arr = np.zeros(len(evals), dtype=[("mag", np.float64), ("sign", np.float64)])
    arr["mag"] = np.abs(evals)
    arr["sign"] = np.sign(np.imag(evals))
    I = np.argsort(arr, order=["mag", "sign"])[::-1]
    return evals[I], evecs[:, I]


@pytest.mark.parametrize("Ct", [CT_SYM, CT_NONSYM])
def test_oracle_matches_scipy_generalized(Ct):
    v0, _ = _sort(*scipy.linalg.eig(Ct, C0))
    v, R = oeig.generalized_eig(C0, Ct)
    v, R = _sort(v, R)
    np.testing.assert_allclose(v, v0, rtol=1e-12)
    K = np.linalg.inv(C0) @ Ct
    np.testing.assert_allclose(K, R @ np.diag(v) @ np.linalg.inv(R), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("Ct", [CT_SYM, CT_NONSYM])
@pytest.mark.parametrize("method", ["QR", "schur"])
def test_eig_corr_known_answer(Ct, method):
    """test_direct.py:36-71."""
    v0, R0 = _sort(*scipy.linalg.eig(Ct, C0))
    v, R = _sort(*_solvers.eig_corr(C0, Ct, method=method))
    try:
        np.testing.assert_allclose(v0, v)
    except AssertionError:
        np.testing.assert_allclose(v0, v.conj())
    for i in range(R0.shape[1]):
        try:
            np.testing.assert_allclose(R0[:, i] / R0[0, i], R[:, i] / R[0, i])
        except AssertionError:
            np.testing.assert_allclose((R0[:, i] / R0[0, i]).conj(), R[:, i] / R[0, i])
    K = np.linalg.inv(C0) @ Ct
    np.testing.assert_allclose(K, R @ np.diag(v) @ np.linalg.inv(R))


def test_spd_inv_split():
    """test_direct.py:18-33."""
    for method in ["QR", "schur"]:
        L = _solvers.spd_inv_split(C0, method=method)
        assert np.allclose(L @ L.T, np.linalg.inv(C0))
        M = L.T @ L
        assert np.max(np.abs(M - np.diag(np.diag(M)))) < 1e-12
    with pytest.raises(AssertionError):
        _solvers.spd_inv_split(np.array([[1.0, 0.2], [0.3, 0.8]]))


def test_zero_rank():
    """direct.py:89-90 / test_tica.py:613-632."""
    with pytest.raises(ZeroRankError):
        _solvers.eig_corr(np.zeros((3, 3)), np.zeros((3, 3)), epsilon=1e-6)


def test_rank_truncation_drops_constant_direction():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2000, 2))
    X = np.hstack([X, X[:, :1]])  # duplicated column (test_tica.py:111-121)
    C = X.T @ X / len(X)
    l, R = _solvers.eig_corr(C, C * 0.5, epsilon=1e-6)
    assert len(l) == 2 and R.shape == (3, 2)


def _whitened(C0, Ct, epsilon):
    """The reference's route (direct.py:225-295): whitening with C0's
    eigendecomposition, then the symmetric eigenproblem, max element positive."""
    s, V = np.linalg.eigh(C0)
    keep = np.abs(s) >= epsilon
    L = V[:, keep] / np.sqrt(s[keep])
    l, Rt = np.linalg.eigh(L.T @ Ct @ L)
    I = np.argsort(np.abs(l))[::-1]
    l, R = l[I], L @ Rt[:, I]
    for j in range(R.shape[1]):
        R[:, j] *= np.sign(R[np.argmax(np.abs(R[:, j])), j])
    return l, R


@pytest.mark.parametrize("smin", [1e-2, 3e-6, 1.5e-6, 0.5e-6])
def test_cholesky_route_equals_whitening(smin):
    """Full-rank C0 takes the Cholesky (dsygvd) route, C0 with an eigenvalue
    near or below the cutoff the whitening route; both give the whitening's
    eigenvalues and eigenvectors, and the cutoff drops the same directions."""
    rng = np.random.default_rng(7)
    d, eps = 40, 1e-6
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    s0 = np.concatenate([np.linspace(1.0, 2.0, d - 1), [smin]])
    C0 = (Q * s0) @ Q.T
    C0 = 0.5 * (C0 + C0.T)
    A = rng.standard_normal((d, d)) * 0.1
    Ct = 0.5 * (A + A.T) + 0.3 * C0
    assert _solvers._no_truncation(C0, eps) == (smin > 2 * eps)
    l, R = _solvers.eig_corr(C0, Ct, epsilon=eps, sign_maxelement=True)
    l0, R0 = _whitened(C0, Ct, eps)
    assert len(l) == len(l0) == (d if smin >= eps else d - 1)
    np.testing.assert_allclose(l, l0, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(R, R0, rtol=1e-6, atol=1e-6 * np.abs(R0).max())
    # C0-orthonormal eigenvectors of Ct r = C0 r l
    np.testing.assert_allclose(R.T @ C0 @ R, np.eye(len(l)), atol=1e-8)
    np.testing.assert_allclose(R.T @ Ct @ R, np.diag(l), atol=1e-8)
