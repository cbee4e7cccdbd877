"""Host algebra of VAMP and PCA, pinned by closed forms (no GPU).

The covariances come from the FP64 oracle.  For one-hot (indicator) features
of a discrete Markov chain, VAMP has closed forms in terms of the transition
count matrix (the reference checks the same identities against an MSM,
pyemma/coordinates/tests/test_vamp.py:176-300):
  K = (C00 + m0 m0^T)^-1 (C0t + m0 mt^T) is the row-normalized count matrix,
  the singular values are those of diag(p0)^1/2 P diag(p1)^-1/2 without the
  leading 1, VAMP2 = trace(P Padj), VAMP1 = nuclear norm, VAMPE = VAMP2.
"""
import numpy as np
import pytest

from oracle.moments import covariance_lagged as oracle_cov
from pyemma_amd.coordinates.pca import PCA
from pyemma_amd.coordinates.vamp import VAMPModel, output_dimension

T3 = np.array([[0.7, 0.2, 0.1], [0.1, 0.8, 0.1], [0.1, 0.1, 0.8]])


def markov_chain(T, n_steps, rng, s0=0):
    """dtraj of a Markov chain with transition matrix T (test_vamp.py:157-164)."""
    cdf = T.cumsum(axis=1)
    u = rng.rand(n_steps)
    out = np.zeros(n_steps, dtype=np.int64)
    s = s0
    for t in range(n_steps):
        out[t] = s
        s = min(int(np.searchsorted(cdf[s], u[t])), T.shape[0] - 1)
    return out


def one_hot_msm(n_traj=20, n_steps=10000, lag=1, seed=7):
    rng = np.random.RandomState(seed)
    dtrajs = [markov_chain(T3, n_steps, rng) for _ in range(n_traj)]
    trajs, C = [], np.zeros((3, 3))
    p0, p1 = np.zeros(3), np.zeros(3)
    for dt in dtrajs:
        x = np.zeros((n_steps, 3))
        x[np.arange(n_steps), dt] = 1.0
        trajs.append(x)
        np.add.at(C, (dt[:-lag], dt[lag:]), 1.0)
        p0 += x[:-lag].sum(axis=0)
        p1 += x[lag:].sum(axis=0)
    P = C / C.sum(axis=1)[:, None]
    return trajs, P, p0 / p0.sum(), p1 / p1.sum()


def vamp_model_from_oracle(trajs, lag, dim=1.0, scaling=None):
    r = oracle_cov(trajs, lag, c00=True, c0t=True, ctt=True, remove_data_mean=True, reversible=False,
                   bessel=False)
    return VAMPModel(mean_0=r["mean"], mean_t=r["mean_tau"], C00=r["C00"], C0t=r["C0t"], Ctt=r["Ctt"], dim=dim,
                     epsilon=1e-6, scaling=scaling)


@pytest.fixture(scope="module")
def msm_case():
    trajs, P, p0, p1 = one_hot_msm()
    return trajs, P, p0, p1, vamp_model_from_oracle(trajs, 1)


def test_vamp_koopman_matrix_is_count_matrix(msm_case):
    trajs, P, p0, p1, m = msm_case
    m0, mt = m.mean_0, m.mean_t
    K = np.linalg.inv(m.C00 + np.outer(m0, m0)).dot(m.C0t + np.outer(m0, mt))
    np.testing.assert_allclose(K, P, atol=1e-10)


def test_vamp_singular_values_and_cumvar(msm_case):
    trajs, P, p0, p1, m = msm_case
    Tsym = np.diag(p0 ** 0.5).dot(P).dot(np.diag(p1 ** -0.5))
    S = np.linalg.svd(Tsym)[1]
    np.testing.assert_allclose(S[0], 1.0, rtol=1e-12)
    np.testing.assert_allclose(m.singular_values[0:2], S[1:], atol=1e-10)
    cs = np.cumsum(S[1:] ** 2)
    np.testing.assert_allclose(m.cumvar, cs / cs[-1], rtol=1e-10)
    assert m.dimension() == 2  # one indicator is linearly dependent on the others


def test_vamp_scores_against_msm(msm_case):
    trajs, P, p0, p1, m = msm_case
    Padj = np.diag(1.0 / p1).dot(P.T).dot(np.diag(p0))
    nfro = np.trace(P.dot(Padj))
    Tsym = np.diag(p0 ** 0.5).dot(P).dot(np.diag(p1 ** -0.5))
    nnuc = np.linalg.norm(Tsym, ord="nuc")
    np.testing.assert_allclose(m.score(score_method="VAMP2"), nfro, rtol=1e-10)
    np.testing.assert_allclose(m.score(score_method="VAMP1"), nnuc, rtol=1e-10)
    np.testing.assert_allclose(m.score(score_method="VAMPE"), nfro, rtol=1e-10)
    with pytest.raises(ValueError):
        m.score(score_method="VAMP3")


def test_vamp_singular_functions_against_msm(msm_case):
    """Left/right singular functions evaluated on the indicator basis equal the MSM ones."""
    trajs, P, p0, p1, m = msm_case
    Tsym = np.diag(p0 ** 0.5).dot(P).dot(np.diag(p1 ** -0.5))
    Up, S, Vhp = np.linalg.svd(Tsym)
    U = (Up * (p0 ** -0.5)[:, None])[:, 1:]
    V = (Vhp.T * (p1 ** -0.5)[:, None])[:, 1:]
    psi = (np.eye(3) - m.mean_0).dot(m.U)
    phi = (np.eye(3) - m.mean_t).dot(m.V)
    for A, B in ((psi, U), (phi, V)):
        for i in range(B.shape[1]):
            assert np.allclose(A[:, i], B[:, i], atol=1e-8) or np.allclose(A[:, i], -B[:, i], atol=1e-8)


def test_vamp_expectation_one_step_is_transition_matrix(msm_case):
    """E[f(x_{t+tau}) | x_t = i] for indicator observables equals row i of P."""
    trajs, P, p0, p1, m = msm_case
    # statistics = indicators: returns (observables x statistics) lagged covariances
    cov = m.expectation(np.eye(3), np.eye(3), lag_multiple=1)
    # cov[j, i] = E[1_j(x_{t+tau}) 1_i(x_t)] = p0_i P_ij
    np.testing.assert_allclose(cov.T, np.diag(p0).dot(P), atol=1e-10)
    # two lag multiples: the rank-2 model of a 3-state chain is exact, so P^2 appears
    cov2 = m.expectation(np.eye(3), np.eye(3), lag_multiple=2)
    np.testing.assert_allclose(cov2.T, np.diag(p0).dot(P).dot(P), atol=1e-10)
    # no statistics: the future expectation of the indicators is the lagged-frame distribution
    np.testing.assert_allclose(m.expectation(np.eye(3), None), p1, atol=1e-10)


def test_vamp_dimension_rules():
    m = VAMPModel(dim=5)
    with pytest.warns(UserWarning):
        assert m.dimension() == 5
    with pytest.raises(RuntimeError):
        VAMPModel(dim=None).dimension()
    with pytest.raises(ValueError):
        VAMPModel(scaling="bogus")
    assert output_dimension(4, 3, None, np.ones(3)) == 3
    assert output_dimension(4, 3, 2, np.ones(3)) == 2
    # float dim: searchsorted into the cumulative kinetic variance
    sv = np.array([0.9, 0.5, 0.1])
    cv = np.cumsum(sv ** 2) / np.sum(sv ** 2)
    assert output_dimension(3, 3, 0.5, sv) == np.searchsorted(cv, 0.5) + 1 == 1
    assert output_dimension(3, 3, 0.9, sv) == 2


def test_vamp_kinetic_map_scaling(msm_case):
    trajs, P, p0, p1, m = msm_case
    mk = vamp_model_from_oracle(trajs, 1, scaling="km")
    np.testing.assert_allclose(np.abs(mk.U), np.abs(m.U) * m.singular_values[None, :2], rtol=1e-10)


def test_pca_parameter_validation():
    with pytest.raises(ValueError):
        PCA(dim=1, var_cutoff=0.8)
    p = PCA(dim=-1, var_cutoff=0.8)
    with pytest.raises(RuntimeError):
        p.dimension()
