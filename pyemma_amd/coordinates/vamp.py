"""VAMP on the GPU moments: covariances from one device pass, FP64 host algebra.

Drop-in for pyemma.coordinates.vamp / VAMP (pyemma/coordinates/transform/
vamp.py:370-1006, API api.py:1294-1443): same arguments, attributes, scores
and projections.  The three covariances C00, C0t, Ctt with separate means
come from one non-reversible pass of `emma_moments_lagged` (flags C0T | CTT
| REMOVE_MEAN: the seam LaggedCovariance(c00, c0t, ctt,
remove_data_mean=True, reversible=False) of vamp.py:578-596); projections
run on the GPU (`emma_project`).  The Chapman-Kolmogorov validator
(vamp.py:839-952, 1009-1101) builds on the MSM package and is out of scope.

The host algebra follows the published method (H. Wu and F. Noe,
"Variational approach for learning Markov processes from time series data",
J. Nonlinear Sci. 30, 23-66 (2020), sections 3-4):

* whitening: Lx with Lx^T C00 Lx = I and Ly with Ly^T Ctt Ly = I on the
  eigenspaces of C00 and Ctt above epsilon (ranks rx, ry);
* the half-weighted Koopman matrix Kw = Lx^T C0t Ly = A diag(s) B^T (SVD);
* left singular functions psi(x) = U^T (x - mean_0) with U = Lx A, right
  ones phi(y) = V^T (y - mean_t) with V = Ly B, truncated to the output
  dimension m (optionally scaled by s, the kinetic map);
* VAMP-r score of the m singular functions on covariances (C00', C0t',
  Ctt') (the model's own, or a test set's):
      sum_i t_i^r + 1,  t = singular values of
      (U^T C00' U)^-1/2 (U^T C0t' V) (V^T Ctt' V)^-1/2   (r = 1, 2);
  VAMP-E: 2 tr(S U^T C0t' V) - tr(S U^T C00' U S V^T Ctt' V) + 1, S = diag(s);
  the 1 is the constant singular function (singular value 1);
* propagation: on the bases [1, psi] of the present and [1, phi] of the
  future the model is diag(1, s); n lag times chain it through the overlap
  of the future basis with the present one.
"""
import warnings

import numpy as np
import scipy.linalg as sla

from .. import _native
from .._solvers import _bounded_threads, spd_inv_split, spd_inv_sqrt
from ._data import DeviceTrajectories, to_device
from .covariance import LaggedCovariance

__all__ = ["VAMPModel", "VAMP", "output_dimension"]

_SCALINGS = (None, "km", "kinetic map")
_SCORES = ("VAMP1", "VAMP2", "VAMPE")


def output_dimension(rank_left, rank_right, dim, singular_values):
    """Number of singular functions kept (the rule of vamp.py:128-136): all
    (min of the two whitening ranks) for dim None or 1.0, the smallest count
    whose cumulative kinetic variance reaches a float dim, else min(ranks, dim)."""
    if dim is None or (isinstance(dim, float) and dim == 1.0):
        return min(rank_left, rank_right)
    if isinstance(dim, float):
        return int(np.searchsorted(kinetic_variance(singular_values), dim)) + 1
    return int(min(rank_left, rank_right, dim))


def kinetic_variance(singular_values):
    """Cumulative share of sum s_i^2 carried by the leading singular values."""
    c = np.cumsum(np.asarray(singular_values) ** 2)
    return c / c[-1]


def _whitener(C, epsilon):
    """L with L^T C L = I on the eigenspace of C above epsilon, and its rank."""
    L = np.atleast_2d(spd_inv_split(C, epsilon=epsilon))
    return L, L.shape[1]


class _Svd(object):
    """Result of one decomposition of the covariances."""
    __slots__ = ("s", "U", "V", "rank_left", "rank_right", "m")

    def __init__(self, s, U, V, rank_left, rank_right, m):
        self.s, self.U, self.V = s, U, V
        self.rank_left, self.rank_right, self.m = rank_left, rank_right, m


class VAMPModel(object):
    """Covariances and means of a VAMP estimate, decomposed on first use.

    Every field the decomposition depends on (covariances, dim, epsilon,
    scaling) drops the cached decomposition when it is assigned.
    """

    _INPUTS = frozenset(("C00", "C0t", "Ctt", "dim", "epsilon", "scaling"))

    def __init__(self, mean_0=None, mean_t=None, C00=None, Ctt=None, C0t=None, dim=None, epsilon=1e-6,
                 scaling=None):
        self._svd = None
        self.set_model_params(mean_0=mean_0, mean_t=mean_t, C00=C00, Ctt=Ctt, C0t=C0t, dim=dim,
                              epsilon=epsilon, scaling=scaling)

    def __setattr__(self, name, value):
        if name == "scaling" and value not in _SCALINGS:
            raise ValueError('unexpected value (%s) of "scaling". Must be one of ("km", "kinetic map", None).'
                             % value)
        if name in self._INPUTS:
            object.__setattr__(self, "_svd", None)
        object.__setattr__(self, name, value)

    def set_model_params(self, mean_0, mean_t, C00, Ctt, C0t, dim, epsilon, scaling=None):
        for name, value in (("mean_0", mean_0), ("mean_t", mean_t), ("C00", C00), ("Ctt", Ctt), ("C0t", C0t),
                            ("dim", dim), ("epsilon", epsilon), ("scaling", scaling)):
            setattr(self, name, value)

    def update_model_params(self, **params):
        for name, value in params.items():
            setattr(self, name, value)

    # ---- decomposition ----
    def _decompose(self):
        with _bounded_threads():
            Lx, rx = _whitener(self.C00, self.epsilon)
            Ly, ry = _whitener(self.Ctt, self.epsilon)
            A, s, Bt = sla.svd(Lx.T.dot(self.C0t).dot(Ly), compute_uv=True, lapack_driver="gesvd")
        m = output_dimension(rx, ry, self.dim, s)
        U = Lx.dot(A[:, :m])
        V = Ly.dot(Bt[:m].T)
        if self.scaling is not None:  # kinetic map: singular functions scaled by their singular values
            U = U * s[:m]
            V = V * s[:m]
        return _Svd(s, U, V, rx, ry, m)

    def _current(self):
        if self._svd is None:
            object.__setattr__(self, "_svd", self._decompose())
        return self._svd

    @property
    def U(self):
        return self._current().U

    @property
    def V(self):
        return self._current().V

    @property
    def singular_values(self):
        return self._current().s

    @property
    def cumvar(self):
        return kinetic_variance(self.singular_values)

    def dimension(self):
        """Output dimension (vamp.py:138-148): the user's integer dim before any estimate."""
        if self.C00 is None:
            if isinstance(self.dim, int):
                warnings.warn("Returning user-input for dimension, since this model has not yet been estimated.")
                return self.dim
            raise RuntimeError("Please call set_model_params prior using this method.")
        return self._current().m

    # ---- propagation ----
    def _future_coefficients(self, observables, m, mean_free):
        """Coefficients of linear observables g(y) = G^T y on [1, phi_1..m]."""
        G = np.asarray(observables, dtype=np.float64)
        coef = np.empty((G.shape[1], m + 1))
        coef[:, 0] = 0.0 if mean_free else G.T.dot(self.mean_t)
        coef[:, 1:] = G.T.dot(self.Ctt).dot(self.V[:, :m])
        return coef

    def _present_coefficients(self, statistics, m, mean_free):
        """Coefficients of linear statistics f(x) = F^T x on [1, psi_1..m]."""
        F = np.asarray(statistics, dtype=np.float64)
        coef = np.empty((F.shape[1], m + 1))
        coef[:, 0] = 0.0 if mean_free else F.T.dot(self.mean_0)
        coef[:, 1:] = F.T.dot(self.C00).dot(self.U[:, :m])
        return coef

    def _propagator(self, m, steps):
        """(m+1) x (m+1) model of `steps` lag times between the future basis [1, phi]
        and the present basis [1, psi]: diag(1, s) once, and between steps the
        overlap of the future basis with the present one, [1, psi](y) =
        O [1, phi](y) with O[1:, 0] = U^T (mean_t - mean_0), O[1:, 1:] = U^T Ctt V."""
        sig = np.diag(np.concatenate(([1.0], self.singular_values[:m])))
        if steps == 1:
            return sig
        U, V = self.U[:, :m], self.V[:, :m]
        overlap = np.eye(m + 1)
        overlap[1:, 0] = U.T.dot(self.mean_t - self.mean_0)
        overlap[1:, 1:] = U.T.dot(self.Ctt).dot(V)
        return np.linalg.matrix_power(sig.dot(overlap), steps - 1).dot(sig)

    def expectation(self, observables, statistics, lag_multiple=1, observables_mean_free=False,
                    statistics_mean_free=False):
        """Model prediction of E[g(x_{t + n tau})] (statistics None) or of the
        lagged covariances E[g(x_{t + n tau}) f(x_t)^T] for linear observables
        g and statistics f (the quantities of vamp.py:150-262)."""
        if lag_multiple < 1:
            raise AssertionError("lag_multiple = 0 not implemented")
        m = self.dimension()
        out = self._future_coefficients(observables, m, observables_mean_free).dot(
            self._propagator(m, int(lag_multiple)))
        if statistics is None:
            return out[:, 0]
        return out.dot(self._present_coefficients(statistics, m, statistics_mean_free).T)

    # ---- scores ----
    def score(self, test_model=None, score_method="VAMP2"):
        """VAMP-1, VAMP-2 or VAMP-E score of this model's singular functions on the
        covariances of test_model (default: this model), +1 for the constant
        function (vamp.py:303-367)."""
        if score_method not in _SCORES:
            raise ValueError('"score" should be one of VAMP1, VAMP2 or VAMPE')
        data = self if test_model is None else test_model
        m = self.dimension()
        U, V = self.U[:, :m], self.V[:, :m]
        with _bounded_threads():
            cross = U.T.dot(data.C0t).dot(V)
            if score_method == "VAMPE":
                s = self.singular_values[:m]
                gram_x = U.T.dot(data.C00).dot(U)
                gram_y = V.T.dot(data.Ctt).dot(V)
                value = 2.0 * np.dot(s, np.diag(cross)) - np.trace((s[:, None] * gram_x).dot(s[:, None] * gram_y))
            else:
                wx = np.atleast_2d(spd_inv_sqrt(U.T.dot(data.C00).dot(U)))
                wy = np.atleast_2d(spd_inv_sqrt(V.T.dot(data.Ctt).dot(V)))
                t = np.linalg.svd(wx.dot(cross).dot(wy), compute_uv=False)
                value = np.sum(t) if score_method == "VAMP1" else np.sum(t ** 2)
        return float(value) + 1.0


def _model_param(name):
    """Estimator attribute stored on its model (vamp.py:531-576)."""
    return property(lambda self: getattr(self._model, name), lambda self, v: setattr(self._model, name, v))


class VAMP(object):
    """Variational approach for Markov processes (vamp.py:370-1006)."""

    dim = _model_param("dim")
    epsilon = _model_param("epsilon")
    scaling = _model_param("scaling")

    def __init__(self, lag, dim=None, scaling=None, right=False, epsilon=1e-6, stride=1, skip=0,
                 ncov_max=float("inf"), precision="fp64"):
        self._model = VAMPModel(dim=dim, epsilon=epsilon, scaling=scaling)
        self._covar = None
        self._data = None
        self._estimated = False
        self.lag = int(lag)
        self.right = right
        self.stride = stride
        self.skip = skip
        self.ncov_max = ncov_max
        self.precision = precision
        self.chunksize = None

    def describe(self):
        return "[VAMP, lag = %i; max. output dim. = %s]" % (self.lag, str(self.dim))

    @property
    def model(self):
        return self._model

    def _estimator(self):
        return LaggedCovariance(c00=True, c0t=True, ctt=True, remove_data_mean=True, reversible=False, lag=self.lag,
                                bessel=False, stride=self.stride, skip=self.skip, weights=None,
                                ncov_max=self.ncov_max, precision=self.precision)

    def _take(self, covar):
        self._model.update_model_params(mean_0=covar.mean, mean_t=covar.mean_tau, C00=covar.C00_,
                                        C0t=covar.C0t_, Ctt=covar.Ctt_)

    def _check_dim(self, dt):
        if isinstance(self.dim, int) and self.dim > dt.d:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, dt.d))

    def estimate(self, X, **kwargs):
        """vamp.py:604-623: one moments pass, then the decomposition."""
        for name, value in kwargs.items():
            if name == "chunksize":
                self.chunksize = value
            elif hasattr(self, name):
                setattr(self, name, value)
        dt = to_device(X)
        self._check_dim(dt)
        covar = self._estimator()
        covar.estimate(dt)
        self._data = dt
        self._take(covar)
        self._model.singular_values  # decompose now, as the reference's estimate does
        self._estimated = True
        return self

    fit = estimate

    def partial_fit(self, X):
        """vamp.py:625-655: moments accumulate; the decomposition waits for the next access."""
        dt = to_device(X)
        self._check_dim(dt)
        if self._covar is None:
            self._covar = self._estimator()
        self._covar.partial_fit(dt)
        if self._covar._estimated:
            self._take(self._covar)
        self._estimated = False
        return self

    def dimension(self):
        return self._model.dimension()

    def output_type(self):
        return np.float32()

    @property
    def singular_values(self):
        return self._model.singular_values

    @property
    def singular_vectors_right(self):
        return self._model.V

    @property
    def singular_vectors_left(self):
        return self._model.U

    @property
    def cumvar(self):
        return self._model.cumvar

    def expectation(self, observables, statistics, lag_multiple=1, observables_mean_free=False,
                    statistics_mean_free=False):
        return self._model.expectation(observables, statistics, lag_multiple=lag_multiple,
                                       observables_mean_free=observables_mean_free,
                                       statistics_mean_free=statistics_mean_free)

    def score(self, test_data=None, score_method="VAMP2"):
        """Self score, or the score of this model on a VAMP estimate of test_data
        with the same hyperparameters (vamp.py:953-1006)."""
        if test_data is None:
            return self._model.score(None, score_method=score_method)
        other = VAMP(self.lag, dim=self.dim, scaling=self.scaling, right=self.right, epsilon=self.epsilon,
                     stride=self.stride, skip=self.skip, ncov_max=self.ncov_max, precision=self.precision)
        other.estimate(test_data)
        return self._model.score(other.model, score_method=score_method)

    # ---- projection on the GPU (vamp.py:661-687) ----
    def _project_device(self, dt):
        m = int(self.dimension())
        mean, M = ((self._model.mean_t, self._model.V) if self.right else (self._model.mean_0, self._model.U))
        t = _native.torch()
        mean_d = t.as_tensor(np.ascontiguousarray(mean, dtype=np.float64), device=dt.X.device)
        M_d = t.as_tensor(np.ascontiguousarray(M[:, :m], dtype=np.float64), device=dt.X.device)
        return _native.project(dt.X, mean_d, M_d, m)

    def transform(self, X):
        single = not isinstance(X, (list, tuple, DeviceTrajectories))
        dt = to_device(X)
        out = dt.split(self._project_device(dt))
        return out[0] if single and dt.ntraj == 1 else out

    def get_output(self, dimensions=slice(0, None), stride=1, skip=0, chunk=None):
        if self._data is None:
            raise RuntimeError("no data: estimate() with data first")
        out = self._data.split(self._project_device(self._data))
        return [y[skip::stride][:, dimensions] for y in out]

    def fit_transform(self, X, **kwargs):
        self.estimate(X, **kwargs)
        return self.transform(X)
