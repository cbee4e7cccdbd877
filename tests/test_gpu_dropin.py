"""The drop-in objects on the GPU: cluster_kmeans / KmeansClustering,
covariance_lagged / LaggedCovariance and TICA, against the oracle and the
reference's own known-answer cases, plus the error paths the reference forces.

Reference cases restated here:
  * k-means: clustering/tests/test_kmeans.py:154-166 (equilibrium: k = 1 on
    the cube vertices gives exactly the origin), :168-179 (an outlier
    converges back inside the cube), :181-233 (the k = 1 center lies in the
    convex hull), :345-364 (resume from given centers moves toward the true
    means), :411-426 (four constant blocks give exactly 30/60/90/120);
  * covariance: coordinates/tests/test_covar_estimator.py (bessel, column
    selection, diagonal only, constant mean removed) through the FP64 oracle;
  * TICA: transform/tica.py:218-245 (kinetic / commute map scaling, cumvar),
    _tica_base.py:98-133 (dimension by var_cutoff, the projection),
    tica.py:165-195 (partial_fit);
  * errors: non-finite input raises InvalidDataInStreamException
    (coordinates/data/_base/datasource.py:1067-1075); no trajectory longer
    than lag + skip raises ValueError, or only warns under partial_fit
    (coordinates/estimation/covariance.py:181-190).
Tolerances: covariances in the default bf16x4-acc mode agree with the FP64
oracle to rtol 2e-5 with atol 1e-5 * max|C| (FP32-class accumulation over
spans of 1024 pairs); FP64 mode to rtol 1e-10.  k-means results (picks,
centers, inertias, labels) are bit-identical to the oracle.
"""
import logging

import numpy as np
import pytest

from oracle import eig as oeig
from oracle import moments as om
from oracle import native as on
from pyemma_amd.coordinates import assign_to_centers, cluster_kmeans, covariance_lagged, tica
from pyemma_amd.exceptions import InvalidDataInStreamException

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=2e-5, atol_rel=1e-5):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol_rel * scale)


CUBE = np.array([[1, 1, 1], [1, 1, -1], [1, -1, -1], [-1, -1, -1],
                 [-1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, 1]], dtype=np.float32)


# ---------------------------------------------------------------- k-means

def test_kmeans_equilibrium_state():
    km = cluster_kmeans(CUBE, k=1)
    assert len(km.clustercenters) == 1
    np.testing.assert_array_equal(km.clustercenters.squeeze(), np.zeros(3))


def test_kmeans_outlier_to_equilibrium():
    X = CUBE.copy()
    X[0, 1] = 1.5
    km = cluster_kmeans(X, k=2, clustercenters=np.array([[2, 0, 0], [-2, 0, 0]]), max_iter=500, n_jobs=1)
    assert np.all(np.abs(km.clustercenters) <= 1)


def test_kmeans_convex_hull():
    pts = np.array([
        [-212129 / 100000, -20411 / 50000, 2887 / 5000], [-212129 / 100000, 40827 / 100000, -5773 / 10000],
        [-141419 / 100000, -5103 / 3125, 2887 / 5000], [-141419 / 100000, 1 / 50000, -433 / 250],
        [-70709 / 50000, 3 / 100000, 17321 / 10000], [-70709 / 50000, 163301 / 100000, -5773 / 10000],
        [-70709 / 100000, -204121 / 100000, -5773 / 10000], [-70709 / 100000, -15309 / 12500, -433 / 250],
        [-17677 / 25000, -122471 / 100000, 17321 / 10000], [-70707 / 100000, 122477 / 100000, 17321 / 10000],
        [-70707 / 100000, 102063 / 50000, 2887 / 5000], [-17677 / 25000, 30619 / 25000, -433 / 250],
        [8839 / 12500, -15309 / 12500, -433 / 250], [35357 / 50000, 102063 / 50000, 2887 / 5000],
        [8839 / 12500, -204121 / 100000, -5773 / 10000], [70713 / 100000, -122471 / 100000, 17321 / 10000],
        [70713 / 100000, 30619 / 25000, -433 / 250], [35357 / 50000, 122477 / 100000, 17321 / 10000],
        [106067 / 50000, -20411 / 50000, 2887 / 5000], [141423 / 100000, -5103 / 3125, 2887 / 5000],
        [141423 / 100000, 1 / 50000, -433 / 250], [8839 / 6250, 3 / 100000, 17321 / 10000],
        [8839 / 6250, 163301 / 100000, -5773 / 10000], [106067 / 50000, 40827 / 100000, -5773 / 10000]],
        dtype=np.float32)
    res = cluster_kmeans(pts, k=1).clustercenters.astype(float).ravel()
    planes = [([-11785060650000, -6804069750000, -4811167325000], 25000531219381),
              ([-1767759097500, 1020624896250, 721685304875], 3749956484003),
              ([-70710363900000, -40824418500000, 57734973820000], 199998509082907),
              ([70710363900000, 40824418500000, -57734973820000], 199998705841169),
              ([70710363900000, -40824995850000, -28867412195000], 149999651832937),
              ([-35355181950000, 20412497925000, -28867282787500], 100001120662259),
              ([23570121300000, 13608139500000, 9622334650000], 49998241292257),
              ([0, 577350000, -204125000], 1060651231),
              ([35355181950000, -20412497925000, 28867282787500], 99997486799779),
              ([0, 72168750, 51030625], 176771554), ([0, -288675000, 102062500], 530329843),
              ([0, 0, 250], 433), ([0, -144337500, -102061250], 353560531), ([0, 0, -10000], 17321)]
    for normal, offset in planes:
        assert np.inner(np.array(normal, dtype=float), res) + offset >= 0


def test_kmeans_four_blocks():
    X = np.zeros((40000, 4))
    for i, v in enumerate([30.0, 60.0, 90.0, 120.0]):
        X[i * 10000:(i + 1) * 10000] = v
    cl = cluster_kmeans(X, k=4)
    found = {float(c[0]) for c in cl.clustercenters if np.all(c == c[0])}
    assert found == {30.0, 60.0, 90.0, 120.0}


def test_kmeans_resume_moves_toward_truth():
    rng = np.random.RandomState(32)
    X = np.hstack([rng.randn(1000) - 2.0, rng.randn(1000), rng.randn(1000) + 2.0])
    cl = cluster_kmeans(X, clustercenters=np.array([[20, 42, -29]]).T, max_iter=1, k=3, keep_data=True)
    c1 = cl.clustercenters.copy()
    cl.estimate(X, clustercenters=c1, max_iter=50)
    true = np.array([-2.0, 0.0, 2.0])
    assert np.linalg.norm(true - np.sort(cl.clustercenters.ravel())) < np.linalg.norm(true - np.sort(c1.ravel()))


@pytest.mark.parametrize("n,d,k,max_iter,tol", [(20000, 30, 64, 10, 1e-5), (6000, 300, 40, 4, 0.0),
                                                 (9000, 5, 100, 30, 1e-3)])
def test_cluster_kmeans_matches_oracle(n, d, k, max_iter, tol):
    """Seed (fixed_seed=True -> 42, kmeans.py:142-164), kmeans++ picks, Lloyd
    centers, inertias, the converged flag and the dtrajs, all bit-identical
    to the oracle's documented rule."""
    rng = np.random.default_rng(n + d)
    means = rng.standard_normal((8, d)) * 4
    X = (means[rng.integers(0, 8, n)] + rng.standard_normal((n, d))).astype(np.float32)
    trajs = [X[: n // 3], X[n // 3:]]
    cl = cluster_kmeans(trajs, k=k, max_iter=max_iter, tolerance=tol, fixed_seed=True)
    C0, _ = on.kmeanspp(X, k, 42, nthreads=8)
    np.testing.assert_array_equal(cl.initial_centers_, C0)
    C_ref, _, in_ref, conv_ref = on.lloyd(X, C0, max_iter, tol, nthreads=8)
    np.testing.assert_array_equal(cl.clustercenters, C_ref)
    np.testing.assert_array_equal(cl.inertias, in_ref)
    assert cl.converged == conv_ref
    lab_ref = on.assign(X, C_ref, nthreads=8)
    np.testing.assert_array_equal(np.concatenate(cl.dtrajs), lab_ref)
    assert [len(t) for t in cl.dtrajs] == [len(t) for t in trajs]


def test_cluster_kmeans_default_k_and_stride():
    """k default min(sqrt(N), 5000) (kmeans.py:209-216) over the strided frames;
    the dtrajs cover every (unstrided) frame."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((4000, 6)).astype(np.float32)
    cl = cluster_kmeans(X, stride=4, max_iter=3, fixed_seed=7)
    assert cl.n_clusters == int(np.sqrt(1000))
    C0, _ = on.kmeanspp(X[::4], cl.n_clusters, 7)
    np.testing.assert_array_equal(cl.initial_centers_, C0)
    assert len(cl.dtrajs[0]) == 4000
    np.testing.assert_array_equal(cl.dtrajs[0], on.assign(X, cl.clustercenters))


def test_kmeans_uniform_init_rows_of_data():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3000, 4)).astype(np.float32)
    cl = cluster_kmeans(X, k=20, init_strategy="uniform", max_iter=0, fixed_seed=5)
    rows = {tuple(r) for r in X}
    assert all(tuple(c) in rows for c in cl.initial_centers_)
    cl2 = cluster_kmeans(X, k=20, init_strategy="uniform", max_iter=0, fixed_seed=5)
    np.testing.assert_array_equal(cl.initial_centers_, cl2.initial_centers_)


# ---------------------------------------------------------------- covariance_lagged

def _trajs(seed, d, lens=(3000, 700, 15, 1200)):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((T, d)).cumsum(0) * 0.1 + rng.uniform(-2, 2, d)).astype(np.float32) for T in lens]


@pytest.mark.parametrize("precision,rtol,atol", [("bf16x4-acc", 2e-5, 1e-5), ("fp64", 1e-10, 1e-12)])
@pytest.mark.parametrize("bessel", [True, False])
@pytest.mark.parametrize("reversible", [True, False])
def test_covariance_lagged_vs_oracle(precision, rtol, atol, bessel, reversible):
    trajs = _trajs(1, 37)
    lag = 7
    cov = covariance_lagged(trajs, c00=True, c0t=True, ctt=not reversible, remove_data_mean=True,
                            reversible=reversible, bessel=bessel, lag=lag, precision=precision)
    ref = om.covariance_lagged(trajs, lag, c00=True, c0t=True, ctt=not reversible, remove_data_mean=True,
                               reversible=reversible, bessel=bessel)
    _close(cov.mean, ref["mean"], rtol=1e-6 if precision != "fp64" else 1e-12)
    _close(cov.C00_, ref["C00"], rtol, atol)
    _close(cov.C0t_, ref["C0t"], rtol, atol)
    if not reversible:
        _close(cov.mean_tau, ref["mean_tau"], rtol=1e-6 if precision != "fp64" else 1e-12)
        _close(cov.Ctt_, ref["Ctt"], rtol, atol)


def test_covariance_lagged_default_is_bessel():
    trajs = _trajs(2, 9)
    cov = covariance_lagged(trajs, lag=3, c0t=True, precision="fp64")
    ref = om.covariance_lagged(trajs, 3, c0t=True, bessel=True)
    _close(cov.C00_, ref["C00"], 1e-10, 1e-12)
    assert cov.bessel
    np.testing.assert_allclose(cov.C00_ * (ref["w"] - 1), ref["Mxx"], rtol=1e-10, atol=1e-10 * np.abs(ref["Mxx"]).max())


def test_covariance_lagged_column_selection_and_diag():
    trajs = _trajs(3, 20)
    sel = np.array([0, 5, 19, 7])
    ref = om.covariance_lagged(trajs, 4, c0t=True, remove_data_mean=True)
    cov = covariance_lagged(trajs, lag=4, c0t=True, remove_data_mean=True, column_selection=sel, precision="fp64")
    assert cov.C00_.shape == (20, 4) and cov.C0t_.shape == (20, 4)
    _close(cov.C00_, ref["C00"][:, sel], 1e-10, 1e-12)
    _close(cov.C0t_, ref["C0t"][:, sel], 1e-10, 1e-12)
    cov = covariance_lagged(trajs, lag=4, c0t=True, remove_data_mean=True, diag_only=True, precision="fp64")
    assert cov.C00_.shape == (20,)
    _close(cov.C00_, np.diag(ref["C00"]), 1e-10, 1e-12)
    _close(cov.C0t_, np.diag(ref["C0t"]), 1e-10, 1e-12)
    with pytest.raises(ValueError):
        covariance_lagged(trajs, lag=4, column_selection=sel, diag_only=True)


def test_covariance_lagged_constant_mean():
    """remove_constant_mean=c: moments of x - c (covariance.py:96-125): the
    oracle on the shifted data, no data-mean removal."""
    trajs = _trajs(4, 11)
    c = np.linspace(-1, 1, 11)
    cov = covariance_lagged(trajs, lag=2, c0t=True, remove_constant_mean=c, bessel=False, precision="fp64")
    ref = om.covariance_lagged([t.astype(np.float64) - c for t in trajs], 2, c0t=True, bessel=False)
    _close(cov.C00_, ref["C00"], 1e-9, 1e-12)
    _close(cov.C0t_, ref["C0t"], 1e-9, 1e-12)
    _close(cov.mean, ref["mean"], 1e-9, 1e-12)


def test_covariance_lagged_too_short():
    trajs = [np.ones((5, 3), np.float32), np.ones((8, 3), np.float32)]
    with pytest.raises(ValueError):
        covariance_lagged(trajs, lag=8, c0t=True)
    with pytest.raises(ValueError):
        tica(trajs, lag=10)
    # skip counts too (iterable.py:205-207): 8 frames, lag 6 + skip 2 -> no pair
    with pytest.raises(ValueError):
        covariance_lagged(trajs, lag=6, skip=2, c0t=True)


def test_partial_fit_too_short_warns(caplog):
    from pyemma_amd.coordinates.covariance import LaggedCovariance
    lc = LaggedCovariance(c0t=True, lag=10)
    with caplog.at_level(logging.WARNING):
        lc.partial_fit(np.ones((5, 3), np.float32))
    assert any("longer than lag" in r.getMessage() for r in caplog.records)
    assert not lc._estimated


# ---------------------------------------------------------------- TICA

@pytest.fixture(scope="module")
def hmm_data():
    """A 4-state birth-death chain with distinct relaxation rates (so the three
    slow TICA eigenvectors are well separated), Gaussian emissions in 12 dims,
    60,000 frames in two trajectories."""
    rng = np.random.default_rng(4)
    P = np.array([[0.99, 0.01, 0.0, 0.0], [0.004, 0.99, 0.006, 0.0],
                  [0.0, 0.015, 0.97, 0.015], [0.0, 0.0, 0.05, 0.95]])
    cum = np.cumsum(P, axis=1)
    u = rng.random(60000)
    s = np.empty(60000, dtype=np.int64)
    s[0] = 0
    for t in range(1, 60000):
        s[t] = min(int(np.searchsorted(cum[s[t - 1]], u[t], side="right")), 3)
    mu = rng.normal(0.0, 5.0, size=(4, 12))
    X = (mu[s] + rng.standard_normal((60000, 12))).astype(np.float32)
    return [X[:25000], X[25000:]]


def _ref_tica(trajs, lag):
    ref = om.covariance_lagged(trajs, lag, remove_data_mean=True, reversible=True, bessel=False)
    l, R = oeig.generalized_eig(ref["C00"], ref["C0t"])
    return ref, l, R


@pytest.mark.parametrize("precision", ["bf16x4-acc", "fp64"])
def test_tica_kinetic_map_and_projection(hmm_data, precision):
    lag = 10
    ref, l_ref, R_ref = _ref_tica(hmm_data, lag)
    t = tica(hmm_data, lag=lag, precision=precision)
    # the three slow eigenvalues of the 4-state chain (|l| >= 0.1) to 1e-6 in
    # either mode; every eigenvalue to 1e-6 in fp64 mode
    big = np.abs(l_ref) >= 0.1
    assert big.sum() == 3
    np.testing.assert_allclose(t.eigenvalues[big], l_ref[big], rtol=1e-6)
    if precision == "fp64":
        np.testing.assert_allclose(t.eigenvalues, l_ref, rtol=1e-6)
    # eigenvectors: C00-orthonormal, equal to the oracle's up to sign on the
    # slow components, kinetic map scaled by the eigenvalues
    V = t.eigenvectors / t.eigenvalues[None, :]
    for i in np.flatnonzero(big):
        r = R_ref[:, i] * np.sign(R_ref[np.argmax(np.abs(R_ref[:, i])), i])
        np.testing.assert_allclose(V[:, i], r, rtol=0, atol=1e-4 * np.abs(r).max())
        assert V[np.argmax(np.abs(V[:, i])), i] > 0  # sign convention: max-abs element positive
    # cumvar and the var_cutoff dimension
    cv = np.cumsum(t.eigenvalues ** 2) / np.sum(t.eigenvalues ** 2)
    np.testing.assert_allclose(t.cumvar, cv, rtol=1e-12)
    assert t.dimension() == int(np.searchsorted(t.cumvar, 0.95) + 1)
    np.testing.assert_allclose(t.timescales, -lag / np.log(np.abs(t.eigenvalues)), rtol=1e-12)
    # the projection: float32 of (x - mean) R in FP64, and the kinetic map
    # property cov(Y) = diag(l^2) of the projected data
    Y = t.get_output()
    X = np.concatenate(hmm_data).astype(np.float64)
    m = t.dimension()
    Yref = (X - t.mean) @ t.eigenvectors[:, :m]
    np.testing.assert_allclose(np.concatenate(Y), Yref, rtol=2e-6, atol=1e-6 * np.abs(Yref).max())
    assert all(y.dtype == np.float32 for y in Y) and [len(y) for y in Y] == [len(a) for a in hmm_data]
    cy = np.cov(np.concatenate(Y)[:, :3].astype(np.float64).T, bias=True)
    np.testing.assert_allclose(cy, np.diag(t.eigenvalues[:3] ** 2), atol=2e-3)
    np.testing.assert_array_equal(t.transform(hmm_data[0]), Y[0])


def test_tica_commute_map_and_dim(hmm_data):
    lag = 10
    t = tica(hmm_data, lag=lag, kinetic_map=False, commute_map=True, dim=4, precision="fp64")
    assert t.dimension() == 4
    _, l_ref, R_ref = _ref_tica(hmm_data, lag)
    ts = 1 - lag / np.log(np.abs(l_ref))
    reg = 0.5 * ts * np.maximum(np.tanh(np.pi * ((ts - lag) / lag) + 1), 0)
    scale = np.sqrt(reg / 2)
    for i in range(3):
        r = R_ref[:, i] * np.sign(R_ref[np.argmax(np.abs(R_ref[:, i])), i]) * scale[i]
        np.testing.assert_allclose(t.eigenvectors[:, i], r, rtol=0, atol=1e-6 * np.abs(r).max())
    Y = np.concatenate(t.get_output())
    assert Y.shape[1] == 4
    with pytest.raises(ValueError):
        tica(hmm_data, lag=lag, kinetic_map=True, commute_map=True)


def test_tica_partial_fit_equals_estimate(hmm_data):
    lag = 10
    full = tica(hmm_data, lag=lag, precision="fp64")
    from pyemma_amd.coordinates.tica import TICA
    inc = TICA(lag, precision="fp64")
    for x in hmm_data:
        inc.partial_fit(x)
    _close(inc.cov, full.cov, 1e-10, 1e-12)
    _close(inc.cov_tau, full.cov_tau, 1e-10, 1e-12)
    np.testing.assert_allclose(inc.eigenvalues, full.eigenvalues, rtol=1e-9)
    ref = om.covariance_lagged(hmm_data, lag, remove_data_mean=True, reversible=True, bessel=False)
    _close(inc.cov, ref["C00"], 1e-10, 1e-12)
    _close(inc.mean, ref["mean"], 1e-12, 1e-14)


# ---------------------------------------------------------------- non-finite input

@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_input_raises(bad):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((500, 6)).astype(np.float32)
    X[123, 4] = bad
    with pytest.raises(InvalidDataInStreamException):
        tica(X, lag=5)
    with pytest.raises(InvalidDataInStreamException):
        cluster_kmeans(X, k=5)
    with pytest.raises(InvalidDataInStreamException):
        assign_to_centers(X, centers=rng.standard_normal((3, 6)).astype(np.float32))
    with pytest.raises(InvalidDataInStreamException):
        covariance_lagged([X[:100], X], lag=3, c0t=True)
