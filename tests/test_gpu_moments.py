"""GPU lagged moments through the C ABI vs the FP64 oracle.

Tolerances: BF16X4 and FP32 tiles carry products with ~2^-16 relative error
(bf16 hi/lo split; gfx950's FP32 MFMA behaves alike) and accumulate in FP32
over at most 1024 pairs before an FP64 merge: entries agree with the FP64
oracle to rtol 2e-5 with atol 1e-5 * max|C| on the small test inputs (few
pairs, so little averaging).  FP64 tiles agree to rtol 1e-10.
"""
import numpy as np
import pytest

from oracle import eig as oeig
from oracle import moments as om
from pyemma_amd import _native

pytestmark = pytest.mark.gpu

PRECS = [_native.EMMA_PREC_BF16X4, _native.EMMA_PREC_FP32, _native.EMMA_PREC_FP64, _native.EMMA_PREC_BF16X4_ACC]


def _gpu(trajs, lag, prec, **kw):
    t = _native.torch()
    X = np.concatenate([np.asarray(a, dtype=np.float32).reshape(len(a), -1) for a in trajs])
    off = np.concatenate([[0], np.cumsum([len(a) for a in trajs])]).astype(np.int64)
    Xt = _native.as_device_f32(X)
    out = _native.moments_lagged(Xt, off, lag, precision=prec, **kw)
    t.cuda.synchronize()
    return [None if o is None else o.cpu().numpy() for o in out]


def _close(a, b, rtol=2e-5, atol_rel=1e-5, prec=None):
    if prec == _native.EMMA_PREC_FP64:
        rtol, atol_rel = 1e-10, 1e-12
    scale = max(np.max(np.abs(b)), 1e-300)
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol_rel * scale)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d", [1, 3, 30, 130, 300])
def test_reversible_mean_free(prec, d):
    rng = np.random.default_rng(d)
    trajs = [rng.standard_normal((T, d)).cumsum(0) * 0.1 + 3.0 for T in (3000, 57, 10, 1200)]
    trajs = [t.astype(np.float32) for t in trajs]
    lag = 10
    ref = om.covariance_lagged(trajs, lag, remove_data_mean=True, reversible=True, bessel=False)
    w, sx, sy, Mxx, Mxy, _ = _gpu(trajs, lag, prec, reversible=True, remove_mean=True, c0t=True)
    assert w[0] == ref["w"]
    _close(sx / w[0], ref["mean"], rtol=1e-6, prec=prec)
    _close(Mxx / w[0], ref["C00"], prec=prec)
    _close(Mxy / w[0], ref["C0t"], prec=prec)
    np.testing.assert_array_equal(Mxx, Mxx.T)
    np.testing.assert_array_equal(Mxy, Mxy.T)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("remove_mean", [True, False])
@pytest.mark.parametrize("d,ctt", [(7, True), (130, True), (300, True), (300, False)])
def test_nonreversible_with_ctt(prec, remove_mean, d, ctt):
    rng = np.random.default_rng(4 + d)
    trajs = [(rng.standard_normal((T, d)).cumsum(0) * 0.2 + 1.0).astype(np.float32) for T in (800, 333, 5)]
    lag = 3
    ref = om.covariance_lagged(trajs, lag, c00=True, c0t=True, ctt=True, remove_data_mean=remove_mean,
                               reversible=False, bessel=False)
    w, sx, sy, Mxx, Mxy, Myy = _gpu(trajs, lag, prec, reversible=False, remove_mean=remove_mean, c0t=True,
                                   ctt=ctt)
    assert w[0] == ref["w"]
    _close(sx / w[0], ref["mean"], rtol=1e-6, prec=prec)
    _close(sy / w[0], ref["mean_tau"], rtol=1e-6, prec=prec)
    _close(Mxx / w[0], ref["C00"], prec=prec)
    _close(Mxy / w[0], ref["C0t"], prec=prec)
    if ctt:
        _close(Myy / w[0], ref["Ctt"], prec=prec)
        np.testing.assert_array_equal(Myy, Myy.T)
    else:
        assert Myy is None
    np.testing.assert_array_equal(Mxx, Mxx.T)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d", [5, 300])
def test_lag0_stride_skip(prec, d):
    rng = np.random.default_rng(6)
    trajs = [(rng.standard_normal((T, d)) + 2).astype(np.float32) for T in (100, 37)]
    ref = om.covariance_lagged(trajs, 0, c00=True, c0t=False, remove_data_mean=True, bessel=False,
                               stride=3, skip=2)
    w, sx, _, Mxx, _, _ = _gpu(trajs, 0, prec, stride=3, skip=2, reversible=False, remove_mean=True, c0t=False)
    assert w[0] == ref["w"]
    _close(sx / w[0], ref["mean"], rtol=1e-6, prec=prec)
    _close(Mxx / w[0], ref["C00"], prec=prec)
    np.testing.assert_array_equal(Mxx, Mxx.T)
    for lag, stride, skip in [(4, 3, 2), (1, 2, 5)]:
        ref = om.covariance_lagged(trajs, lag, remove_data_mean=True, reversible=True, bessel=False,
                                   stride=stride, skip=skip)
        w, sx, _, Mxx, Mxy, _ = _gpu(trajs, lag, prec, stride=stride, skip=skip, reversible=True,
                                     remove_mean=True, c0t=True)
        assert w[0] == ref["w"]
        _close(Mxx / w[0], ref["C00"], prec=prec)
        _close(Mxy / w[0], ref["C0t"], prec=prec)


def test_koopman_fixture_moments(koopman_data):
    """Inputs of the reference's Koopman tests (test_tica.py:443-452), lag 10."""
    ref = om.covariance_lagged(koopman_data, 10, remove_data_mean=True, reversible=True, bessel=False)
    w, sx, _, Mxx, Mxy, _ = _gpu(koopman_data, 10, _native.EMMA_PREC_BF16X4, reversible=True, remove_mean=True,
                                 c0t=True)
    assert w[0] == ref["w"]
    _close(Mxx / w[0], ref["C00"], rtol=1e-6, atol_rel=1e-7)
    _close(Mxy / w[0], ref["C0t"], rtol=1e-6, atol_rel=1e-7)


def test_no_pairs_gives_zero_weight():
    w, sx, _, Mxx, Mxy, _ = _gpu([np.ones((5, 3), np.float32)], 10, _native.EMMA_PREC_BF16X4, reversible=True,
                                 remove_mean=True, c0t=True)
    assert w[0] == 0.0 and np.all(Mxx == 0)


def test_tica_eigenvalues_vs_fp64():
    """HMM data (the 100k x 30 4-state chain): every eigenvalue vs the FP64 oracle.

    FP64 tiles: every eigenvalue within 1e-6 relative.  The FP32-class modes
    (bf16x4, fp32, bf16x4-acc): the three slow eigenvalues (|l| >= 0.1) within
    1e-6 relative, and every eigenvalue within an absolute bar: 1e-6 for
    bf16x4-acc, 1e-5 for the single-level modes.  The 26 noise eigenvalues of
    this chain are of order 1e-3, so the absolute error of FP32-class
    accumulation is a relative error of up to ~2e-4 on them (bf16x4-acc,
    measured 1.7e-4); only fp64 mode holds 1e-6 relative on every eigenvalue
    of such data.
    """
    from pyemma_amd.synthetic import hmm_trajectory
    X = hmm_trajectory(100000, 30, seed=1)
    lag = 10
    ref = om.covariance_lagged([X], lag, remove_data_mean=True, reversible=True, bessel=False)
    ev_ref = oeig.generalized_eigenvalues(ref["C00"], ref["C0t"])
    big = np.abs(ev_ref) >= 0.1
    assert big.sum() == 3
    for prec in PRECS:
        w, sx, _, Mxx, Mxy, _ = _gpu([X], lag, prec, reversible=True, remove_mean=True, c0t=True)
        ev = oeig.generalized_eigenvalues(Mxx / w[0], Mxy / w[0])
        rel = np.abs(ev - ev_ref) / np.abs(ev_ref)
        err = np.abs(ev - ev_ref)
        print("prec %d: rel. err all %.3g, |l| >= 0.1 %.3g, abs. err all %.3g"
              % (prec, rel.max(), rel[big].max(), err.max()))
        if prec == _native.EMMA_PREC_FP64:
            assert rel.max() < 1e-6, (prec, rel.max())
        else:
            assert rel[big].max() < 1e-6, (prec, rel[big].max())
            assert err.max() < (1e-6 if prec == _native.EMMA_PREC_BF16X4_ACC else 1e-5), (prec, err.max())


def test_accurate_mode_on_distance_features():
    """Pairwise distances of 25 Brownian particles (the 300 collinear features of the
    200 x 50,000 workload; 2 trajectories here), lag 20: the generalized eigenproblem
    amplifies the FP32 accumulation error of long MFMA chains.  BF16X4_ACC (two-level
    FP32 accumulation) keeps every one of the 300 eigenvalues within 1e-6 relative of
    the FP64 oracle (on the whole 200-trajectory workload: 3.8e-9); plain BF16X4 is
    held to 1e-4."""
    from pyemma_amd.synthetic import brownian_distances
    t = _native.torch()
    X, off = brownian_distances(2, 50000, seed=2, device=_native.device())
    Xh = X.cpu().numpy()
    lag = 20
    trajs = [Xh[off[i]:off[i + 1]] for i in range(2)]
    ref = om.covariance_lagged(trajs, lag, remove_data_mean=True, reversible=True, bessel=False)
    ev_ref = oeig.generalized_eigenvalues(ref["C00"], ref["C0t"])
    assert len(ev_ref) == 300
    errs = {}
    for prec in (_native.EMMA_PREC_BF16X4, _native.EMMA_PREC_BF16X4_ACC):
        w, sx, _, Mxx, Mxy, _ = [None if o is None else o.cpu().numpy() for o in _native.moments_lagged(
            X, off, lag, precision=prec, reversible=True, remove_mean=True, c0t=True)]
        t.cuda.synchronize()
        ev = oeig.generalized_eigenvalues(Mxx / w[0], Mxy / w[0])
        errs[prec] = (np.abs(ev - ev_ref) / np.abs(ev_ref)).max()
    print("eigenvalue rel. err (all 300): bf16x4 %.3g, bf16x4-acc %.3g"
          % (errs[_native.EMMA_PREC_BF16X4], errs[_native.EMMA_PREC_BF16X4_ACC]))
    assert errs[_native.EMMA_PREC_BF16X4_ACC] < 1e-6
    assert errs[_native.EMMA_PREC_BF16X4] < 1e-4


def test_deterministic_repeat():
    rng = np.random.default_rng(2)
    trajs = [rng.standard_normal((5000, 64)).astype(np.float32)]
    a = _gpu(trajs, 5, _native.EMMA_PREC_BF16X4, reversible=True, remove_mean=True, c0t=True)
    b = _gpu(trajs, 5, _native.EMMA_PREC_BF16X4, reversible=True, remove_mean=True, c0t=True)
    for x, y in zip(a, b):
        if x is not None:
            np.testing.assert_array_equal(x, y)
