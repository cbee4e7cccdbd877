"""VAMP and PCA on the GPU moments kernels.

* VAMP on indicator features of a 3-state Markov chain against the MSM
  closed forms (pyemma/coordinates/tests/test_vamp.py:176-300): Koopman matrix
  = row-normalized count matrix, singular values, VAMP1/2/E scores;
* VAMP self-consistency of test_vamp.py:71-155 (whitened left/right singular
  functions, their cross-covariance = diag(singular values), partial_fit ==
  estimate), full rank and low rank, at the reference test's tolerances;
* VAMP covariances against the FP64 oracle at every precision;
* PCA: covariance (Bessel-corrected) against the FP64 oracle, eigenvalues
  against numpy's eigh, projected variances = eigenvalues, the dimension
  rules of test_pca.py:94-104.
"""
import numpy as np
import pytest

from oracle.moments import covariance_lagged as oracle_cov
from pyemma_amd.coordinates import api

from test_vamp_pca_cpu import one_hot_msm, markov_chain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def msm_vamp():
    trajs, P, p0, p1 = one_hot_msm()
    v = api.vamp(trajs, lag=1, scaling=None, dim=1.0)
    return trajs, P, p0, p1, v


def test_vamp_koopman_is_count_matrix(msm_vamp):
    trajs, P, p0, p1, v = msm_vamp
    m = v.model
    K = np.linalg.inv(m.C00 + np.outer(m.mean_0, m.mean_0)).dot(m.C0t + np.outer(m.mean_0, m.mean_t))
    np.testing.assert_allclose(K, P, atol=1e-10)
    Tsym = np.diag(p0 ** 0.5).dot(P).dot(np.diag(p1 ** -0.5))
    np.testing.assert_allclose(v.singular_values[0:2], np.linalg.svd(Tsym)[1][1:], atol=1e-10)


def test_vamp_scores_vs_msm(msm_vamp):
    trajs, P, p0, p1, v = msm_vamp
    Padj = np.diag(1.0 / p1).dot(P.T).dot(np.diag(p0))
    nfro = np.trace(P.dot(Padj))
    Tsym = np.diag(p0 ** 0.5).dot(P).dot(np.diag(p1 ** -0.5))
    np.testing.assert_allclose(v.score(score_method="VAMP2"), nfro)
    np.testing.assert_allclose(v.score(score_method="VAMP1"), np.linalg.norm(Tsym, ord="nuc"))
    np.testing.assert_allclose(v.score(score_method="VAMPE"), nfro)


def test_vamp_cross_validation_score(msm_vamp):
    """score(test_data) = model score against a model estimated on the test trajectories."""
    trajs, P, p0, p1, v = msm_vamp
    train, test = trajs[:10], trajs[10:]
    vt = api.vamp(train, lag=1, dim=1.0)
    ve = api.vamp(test, lag=1, dim=1.0)
    for method in ("VAMP1", "VAMP2", "VAMPE"):
        np.testing.assert_allclose(vt.score(test_data=test, score_method=method),
                                   vt.model.score(ve.model, score_method=method), rtol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x4", "bf16x4-acc"])
def test_vamp_covariances_vs_oracle(precision):
    rng = np.random.RandomState(3)
    trajs = [np.cumsum(rng.randn(n, 12), axis=0).astype(np.float32) * 0.1 + rng.randn(12).astype(np.float32)
             for n in (300, 517, 1024)]
    v = api.vamp(trajs, lag=7, precision=precision)
    r = oracle_cov([t.astype(np.float64) for t in trajs], 7, c00=True, c0t=True, ctt=True, remove_data_mean=True,
                   reversible=False, bessel=False)
    rtol = 1e-10 if precision == "fp64" else 3e-5
    for a, b in ((v.model.mean_0, r["mean"]), (v.model.mean_t, r["mean_tau"])):
        np.testing.assert_allclose(a, b, rtol=1e-10 if precision == "fp64" else 1e-6, atol=1e-12)
    for a, b in ((v.model.C00, r["C00"]), (v.model.C0t, r["C0t"]), (v.model.Ctt, r["Ctt"])):
        np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * np.abs(b).max())


def _random_matrix(n, rank, rng, eps=0.01):
    """test_vamp.py:34-42."""
    m = rng.randn(n, n)
    u, s, v = np.linalg.svd(m)
    s = np.concatenate((np.maximum(s, eps)[0:rank], np.zeros(n - rank)))
    return u.dot(np.diag(s)).dot(v)


def _allclose_ignore_phase(A, B, atol, rtol):
    A, B = np.atleast_2d(A), np.atleast_2d(B)
    assert A.shape == B.shape
    for i in range(B.shape[1]):
        assert (np.allclose(A[:, i], B[:, i], atol=atol, rtol=rtol)
                or np.allclose(A[:, i], -B[:, i], atol=atol, rtol=rtol))


@pytest.mark.parametrize("dim,rank", [(20, 20), (30, 15)])
def test_vamp_self_consistency(dim, rank):
    """test_vamp.py:71-155 at its tolerances (float32 output: atol 10 eps, rtol resolution)."""
    rng = np.random.RandomState(11 + dim)
    A = _random_matrix(dim, rank, rng)
    mean = rng.randn(dim)
    trajs = [np.dot(np.cumsum(rng.randn(n, dim), axis=0), A) + mean for n in (123, 456, 789)]
    tau = 50
    v = api.vamp(trajs, lag=tau, scaling=None)
    assert v.dimension() <= rank
    atol = np.finfo(np.float32).eps * 10.0
    rtol = np.finfo(np.float32).resolution
    v.right = True
    phi_trajs = [sf[tau:, :] for sf in v.get_output()]
    phi = np.concatenate(phi_trajs)
    np.testing.assert_allclose(phi.sum(axis=0) / phi.shape[0], 0.0, rtol=rtol, atol=atol)
    np.testing.assert_allclose(phi.T.dot(phi) / phi.shape[0], np.eye(v.dimension()), rtol=rtol, atol=atol)
    v.right = False
    psi_trajs = [sf[0:-tau, :] for sf in v.get_output()]
    psi = np.concatenate(psi_trajs)
    np.testing.assert_allclose(psi.sum(axis=0) / psi.shape[0], 0.0, rtol=rtol, atol=atol)
    np.testing.assert_allclose(psi.T.dot(psi) / psi.shape[0], np.eye(v.dimension()), rtol=rtol, atol=atol)
    C01 = psi.T.dot(phi) / phi.shape[0]
    np.testing.assert_allclose(C01, np.diag(v.singular_values[0:v.dimension()]), rtol=rtol, atol=atol)

    # partial_fit, one trajectory at a time, gives the same model
    v2 = api.vamp(lag=tau, scaling=None)
    for t in trajs:
        v2.partial_fit(t)
    for name in ("mean_0", "mean_t", "C00", "C0t", "Ctt"):
        np.testing.assert_allclose(getattr(v.model, name), getattr(v2.model, name), rtol=1e-5, atol=1e-14,
                                   err_msg=name)
    np.testing.assert_allclose(v.singular_values, v2.singular_values, rtol=1e-5, atol=1e-14)
    v2.right = True
    for t, ref in zip(trajs, phi_trajs):
        _allclose_ignore_phase(v2.transform(t[tau:]), ref, atol=1e-6, rtol=1e-5)
    v2.right = False
    for t, ref in zip(trajs, psi_trajs):
        _allclose_ignore_phase(v2.transform(t[0:-tau]), ref, atol=1e-6, rtol=1e-5)


def test_vamp_kinetic_map(msm_vamp):
    """test_vamp.py:320-325: with scaling='km', the std of the left functions is the singular values."""
    trajs = msm_vamp[0]
    lag = 10
    v = api.vamp(trajs, lag=lag, scaling="km", right=False)
    transformed = [t[:-lag] for t in v.get_output()]
    std = np.std(np.concatenate(transformed), axis=0)
    np.testing.assert_allclose(std, v.singular_values[:v.dimension()], atol=1e-4, rtol=1e-4)


@pytest.fixture(scope="module")
def two_gaussian_hmm():
    """test_pca.py:38-61: 2-state HMM with two anisotropic Gaussians in 2-D."""
    rng = np.random.RandomState(0)
    P = np.array([[0.99, 0.01], [0.01, 0.99]])
    n = 10000
    dtraj = markov_chain(P, n, rng)
    means = np.array([[-1.0, 1.0], [1.0, -1.0]])
    widths = np.array([[0.3, 2.0], [0.3, 2.0]])
    X = widths[dtraj] * rng.randn(n, 2) + means[dtraj]
    return X


def test_pca_against_oracle(two_gaussian_hmm):
    X = two_gaussian_hmm
    p = api.pca(data=X, dim=1)
    r = oracle_cov([X.astype(np.float32).astype(np.float64)], 0, c00=True, c0t=False, remove_data_mean=True,
                   bessel=True)
    np.testing.assert_allclose(p.cov, r["C00"], rtol=1e-10)
    np.testing.assert_allclose(p.mean, r["mean"], rtol=1e-10)
    w, V = np.linalg.eigh(r["C00"])
    np.testing.assert_allclose(p.eigenvalues, w[::-1], rtol=1e-10)
    _allclose_ignore_phase(p.eigenvectors, V[:, ::-1], atol=1e-10, rtol=1e-8)
    assert len(p.cumvar) == 2 and np.allclose(p.cumvar[-1], 1.0)
    # test_pca.py:82-85
    cov_ref = np.dot(X.T, X) / float(len(X))
    assert np.max(p.cov - cov_ref) < 3e-1


def test_pca_variances_and_outputs(two_gaussian_hmm):
    X = two_gaussian_hmm
    obj = api.pca(data=X)
    O = obj.get_output()[0]
    assert np.max(np.abs(np.var(O, axis=0) - obj.eigenvalues[:O.shape[1]])) < 0.01
    p = api.pca(data=X, dim=1)
    out = p.get_output()
    assert isinstance(out, list) and len(out) == 1
    assert out[0].dtype == np.float32 and out[0].shape == (len(X), 1)
    assert p.eigenvectors.shape == (2, 2)


def test_pca_dimension_rules(two_gaussian_hmm):
    """test_pca.py:94-104."""
    X = two_gaussian_hmm
    assert api.pca(data=X, dim=1).dimension() == 1
    assert api.pca(data=X, dim=-1, var_cutoff=1.0).dimension() == 2
    assert api.pca(data=X, dim=-1, var_cutoff=0.8).dimension() == 1
    with pytest.raises(ValueError):
        api.pca(data=X, dim=1, var_cutoff=0.8)


def test_pca_partial_fit_matches_estimate(two_gaussian_hmm):
    X = two_gaussian_hmm
    p = api.pca(data=X, dim=1)
    q = api.pca(dim=1)
    for lo in range(0, len(X), 2500):
        q.partial_fit(X[lo:lo + 2500])
    np.testing.assert_allclose(q.cov, p.cov, rtol=1e-10)
    np.testing.assert_allclose(q.eigenvalues, p.eigenvalues, rtol=1e-10)
