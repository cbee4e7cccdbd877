"""Frame-weighted moments and Koopman reweighting on the GPU.

* weighted moments (every precision, reversible and not) against the FP64
  oracle's weighted estimator (moments.py:267-303, 540-553 restated);
* the reference's TestKoopman case (coordinates/tests/test_tica.py:440-611)
  on its own input trajectories: the expected means, C0, Ct and eigenvalues
  are computed in the test exactly as the reference test computes them, and
  the GPU path (precision 'fp64') must match them at the reference's
  tolerance (assert_allclose defaults, rtol 1e-7).
"""
import numpy as np
import pytest
import scipy.linalg as scl

from oracle import moments as om
from pyemma_amd import _native
from pyemma_amd.coordinates import api
from pyemma_amd.coordinates.koopman import KoopmanWeights
from pyemma_amd._solvers import sort_by_norm

pytestmark = pytest.mark.gpu

PRECS = [_native.EMMA_PREC_BF16X4, _native.EMMA_PREC_FP32, _native.EMMA_PREC_FP64, _native.EMMA_PREC_BF16X4_ACC]


def _gpu(trajs, weights, lag, prec, **kw):
    t = _native.torch()
    X = np.concatenate(trajs).astype(np.float32)
    off = np.concatenate([[0], np.cumsum([len(a) for a in trajs])]).astype(np.int64)
    W = t.as_tensor(np.concatenate(weights).astype(np.float64), device=_native.device())
    out = _native.moments_lagged(_native.as_device_f32(X), off, lag, precision=prec, weights=W, **kw)
    t.cuda.synchronize()
    return [None if o is None else o.cpu().numpy() for o in out]


def _close(a, b, prec):
    rtol, atol = (1e-10, 1e-12) if prec == _native.EMMA_PREC_FP64 else (2e-5, 1e-5)
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol * np.abs(b).max())


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("reversible", [True, False])
@pytest.mark.parametrize("d", [37, 300])
def test_weighted_moments_vs_oracle(prec, reversible, d):
    rng = np.random.default_rng(11 + d)
    trajs = [(rng.standard_normal((T, d)).cumsum(0) * 0.1 + 2.0).astype(np.float32) for T in (900, 41, 333)]
    weights = [rng.uniform(-0.2, 1.5, size=len(a)) for a in trajs]
    lag = 7
    ref = om.covariance_lagged(trajs, lag, remove_data_mean=True, reversible=reversible, bessel=False,
                               weights=weights)
    w, sx, sy, Mxx, Mxy, _ = _gpu(trajs, weights, lag, prec, reversible=reversible, remove_mean=True, c0t=True)
    np.testing.assert_allclose(w[0], ref["w"], rtol=1e-12)
    _close(sx / w[0], ref["mean"], prec)
    if not reversible:
        _close(sy / w[0], ref["mean_tau"], prec)
    _close(Mxx / w[0], ref["C00"], prec)
    _close(Mxy / w[0], ref["C0t"], prec)
    np.testing.assert_array_equal(Mxx, Mxx.T)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d", [37, 300])
def test_weighted_lag0_vs_oracle(prec, d):
    """Weighted instantaneous moments (moments_XX with weights, moments.py:601-698)."""
    rng = np.random.default_rng(17 + d)
    trajs = [(rng.standard_normal((T, d)) + 1.5).astype(np.float32) for T in (700, 90)]
    weights = [rng.uniform(0.1, 2.0, size=len(a)) for a in trajs]
    ref = om.covariance_lagged(trajs, 0, c00=True, c0t=False, remove_data_mean=True, bessel=False,
                               weights=weights, stride=2, skip=1)
    w, sx, _, Mxx, _, _ = _gpu(trajs, weights, 0, prec, stride=2, skip=1, reversible=False, remove_mean=True,
                               c0t=False)
    np.testing.assert_allclose(w[0], ref["w"], rtol=1e-12)
    _close(sx / w[0], ref["mean"], prec)
    _close(Mxx / w[0], ref["C00"], prec)
    np.testing.assert_array_equal(Mxx, Mxx.T)


def _transform_C0(C, epsilon):
    """test_tica.py:66-76."""
    d, V = scl.eigh(C)
    evmin = np.minimum(0, np.min(d))
    ep = np.maximum(-evmin, epsilon)
    d, V = sort_by_norm(d, V)
    ind = np.where(np.abs(d) > ep)[0]
    d = d[ind]
    V = V[:, ind]
    for j in range(V.shape[1]):
        jj = np.argmax(np.abs(V[:, j]))
        V[:, j] *= np.sign(V[jj, j])
    return np.dot(V, np.diag(d ** (-0.5)))


def test_koopman_reference_case(koopman_data):
    """test_tica.py:440-611 (TestKoopman), expected values computed as there."""
    data = koopman_data
    tau, eps, nf = 10, 1e-6, 10
    mean_x = np.zeros(nf)
    mean_y = np.zeros(nf)
    frames = 0
    for traj in data:
        mean_x += np.sum(traj[:-tau, :], axis=0)
        mean_y += np.sum(traj[tau:, :], axis=0)
        frames += traj[:-tau, :].shape[0]
    mean_x /= frames
    mean_y /= frames
    C0 = np.zeros((nf, nf))
    Ct = np.zeros((nf, nf))
    for traj in data:
        itraj = traj - mean_x[None, :]
        C0 += itraj[:-tau].T.dot(itraj[:-tau])
        Ct += itraj[:-tau].T.dot(itraj[tau:])
    C0 /= frames
    Ct /= frames
    R = _transform_C0(C0, eps)
    K = np.dot(R.T, np.dot(Ct, R))
    K = np.vstack((K, np.dot((mean_y - mean_x), R)))
    K = np.hstack((K, np.eye(K.shape[0], 1, k=-K.shape[0] + 1)))
    N1 = K.shape[0]
    ln, Un = scl.eig(K.T)
    ln, Un = sort_by_norm(ln, Un)
    u = np.real(Un[:, 0])
    u *= 1.0 / np.dot(u, np.eye(N1, 1, k=-N1 + 1)[:, 0])
    N = R.shape[0]
    u_input = np.zeros(N + 1)
    u_input[0:N] = R.dot(u[0:-1])
    u_input[N] = u[-1] - mean_x.dot(R.dot(u[0:-1]))
    weight_obj = KoopmanWeights(u_input[:-1], u_input[-1])
    wtraj = []
    for traj in data:
        tr = np.hstack((np.dot(traj - mean_x[None, :], R), np.ones((traj.shape[0], 1))))
        wtraj.append(tr.dot(u))
    mean_eq = np.zeros(nf)
    for q, traj in enumerate(data):
        mean_eq += np.sum(wtraj[q][:-tau, None] * traj[:-tau], axis=0) + np.sum(wtraj[q][:-tau, None] * traj[tau:],
                                                                                 axis=0)
    mean_eq /= 2 * frames
    C0_eq = np.zeros((nf, nf))
    Ct_eq = np.zeros((nf, nf))
    for q, traj in enumerate(data):
        wq = wtraj[q][:-tau, None]
        tr = traj - mean_eq[None, :]
        C0_eq += (wq * tr[:-tau]).T.dot(tr[:-tau]) + (wq * tr[tau:]).T.dot(tr[tau:])
        Ct_eq += (wq * tr[:-tau]).T.dot(tr[tau:]) + (wq * tr[tau:]).T.dot(tr[:-tau])
    C0_eq /= 2 * frames
    Ct_eq /= 2 * frames
    S = _transform_C0(C0_eq, eps)
    lr, _ = scl.eigh(np.dot(S.T, np.dot(Ct_eq, S)))
    lr, _ = sort_by_norm(lr, np.eye(len(lr)))

    koop_eq = api.tica(data, lag=tau, kinetic_map=False, weights="koopman", precision="fp64")
    koop_direct = api.tica(data, lag=tau, kinetic_map=False, weights=weight_obj, precision="fp64")
    for est in (koop_eq, koop_direct):
        np.testing.assert_allclose(est.mean, mean_eq)
        np.testing.assert_allclose(est.cov, C0_eq)
        np.testing.assert_allclose(est.cov_tau, Ct_eq)
        np.testing.assert_allclose(est.eigenvalues, lr)
