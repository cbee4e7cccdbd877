"""VAMP: GPU lagged moments (C00, C0t, Ctt, separate means), host FP64 SVD, GPU projection.

Mirrors pyemma/coordinates/transform/vamp.py:39-1006.  The three covariance
matrices come from one non-reversible pass of the moments kernels
(`emma_moments_lagged` with C0T | CTT and REMOVE_MEAN, the seam
`LaggedCovariance(c00, c0t, ctt, remove_data_mean=True, reversible=False)` of
vamp.py:578-596); the Koopman SVD, scores and expectations are d x d host
FP64 algebra, as in the reference; the projections run on the GPU
(`emma_project`).  The Chapman-Kolmogorov validator (vamp.py:839-952,
1009-1101) builds on the MSM package and is not part of this path.
"""
import warnings

import numpy as np
import scipy.linalg as sla

from .. import _native
from .._solvers import _bounded_threads, spd_inv_split, spd_inv_sqrt
from ._data import DeviceTrajectories, to_device
from .covariance import LaggedCovariance

__all__ = ["VAMPModel", "VAMP"]


def _mdot(*args):
    """pyemma/util/linalg.py:30-40: chained matrix product."""
    out = args[0]
    for a in args[1:]:
        out = np.dot(out, a)
    return out


class VAMPModel(object):
    """vamp.py:39-367: covariances -> half-weighted Koopman matrix -> SVD."""

    def __init__(self, mean_0=None, mean_t=None, C00=None, Ctt=None, C0t=None, dim=None, epsilon=1e-6, scaling=None):
        self.set_model_params(mean_0=mean_0, mean_t=mean_t, C00=C00, Ctt=Ctt, C0t=C0t, dim=dim, epsilon=epsilon,
                              scaling=scaling)

    def set_model_params(self, mean_0, mean_t, C00, Ctt, C0t, dim, epsilon, scaling=None):
        self.mean_0 = mean_0
        self.mean_t = mean_t
        self.C00 = C00
        self.Ctt = Ctt
        self.C0t = C0t
        self._svd_performed = False
        self.dim = dim
        self.epsilon = epsilon
        self.scaling = scaling

    def update_model_params(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)

    @property
    def scaling(self):
        return self._scaling

    @scaling.setter
    def scaling(self, value):
        if value not in (None, "km", "kinetic map"):
            raise ValueError('unexpected value (%s) of "scaling". Must be one of ("km", "kinetic map", None).'
                             % value)
        self._scaling = value

    # the covariance setters invalidate the decomposition (vamp.py:90-116)
    @property
    def C00(self):
        return self._C00

    @C00.setter
    def C00(self, val):
        self._svd_performed = False
        self._C00 = val

    @property
    def C0t(self):
        return self._C0t

    @C0t.setter
    def C0t(self, val):
        self._svd_performed = False
        self._C0t = val

    @property
    def Ctt(self):
        return self._Ctt

    @Ctt.setter
    def Ctt(self, val):
        self._svd_performed = False
        self._Ctt = val

    @property
    def U(self):
        if not self._svd_performed:
            self._diagonalize()
        return self._U

    @property
    def V(self):
        if not self._svd_performed:
            self._diagonalize()
        return self._V

    @property
    def singular_values(self):
        if not self._svd_performed:
            self._diagonalize()
        return self._singular_values

    @staticmethod
    def _cumvar(singular_values):
        cumvar = np.cumsum(singular_values ** 2)
        cumvar /= cumvar[-1]
        return cumvar

    @property
    def cumvar(self):
        return VAMPModel._cumvar(self.singular_values)

    @staticmethod
    def _dimension(rank0, rankt, dim, singular_values):
        """vamp.py:128-136."""
        if dim is None or (isinstance(dim, float) and dim == 1.0):
            return min(rank0, rankt)
        if isinstance(dim, float):
            return np.searchsorted(VAMPModel._cumvar(singular_values), dim) + 1
        return np.min([rank0, rankt, dim])

    def dimension(self):
        """vamp.py:138-148."""
        if self.C00 is None:
            if isinstance(self.dim, int):
                warnings.warn("Returning user-input for dimension, since this model has not yet been estimated.")
                return self.dim
            raise RuntimeError("Please call set_model_params prior using this method.")
        if not self._svd_performed:
            self._diagonalize()
        return self._dimension(self._rank0, self._rankt, self.dim, self.singular_values)

    def _diagonalize(self):
        """vamp.py:264-301: W = L0^T C0t Lt with C00^-1 = L0 L0^T, Ctt^-1 = Lt Lt^T; SVD (gesvd)."""
        with _bounded_threads():
            L0 = spd_inv_split(self.C00, epsilon=self.epsilon)
            self._rank0 = L0.shape[1] if L0.ndim == 2 else 1
            Lt = spd_inv_split(self.Ctt, epsilon=self.epsilon)
            self._rankt = Lt.shape[1] if Lt.ndim == 2 else 1
            W = np.dot(L0.T, self.C0t).dot(Lt)
            A, s, BT = sla.svd(W, compute_uv=True, lapack_driver="gesvd")
            self._singular_values = s
            m = VAMPModel._dimension(self._rank0, self._rankt, self.dim, self._singular_values)
            U = np.dot(L0, A[:, :m])
            V = np.dot(Lt, BT[:m, :].T)
        if self.scaling is not None:
            U *= s[np.newaxis, 0:m]
            V *= s[np.newaxis, 0:m]
        self._U = U
        self._V = V
        self._svd_performed = True

    def expectation(self, observables, statistics, lag_multiple=1, observables_mean_free=False,
                    statistics_mean_free=False):
        """vamp.py:150-262: future expectation / lagged covariance of linear observables."""
        dim = self.dimension()
        S = np.diag(np.concatenate(([1.0], self.singular_values[0:dim])))
        V = self.V[:, 0:dim]
        U = self.U[:, 0:dim]
        m_0 = self.mean_0
        m_t = self.mean_t
        assert lag_multiple >= 1, "lag_multiple = 0 not implemented"
        if lag_multiple == 1:
            P = S
        else:
            p = np.zeros((dim + 1, dim + 1))
            p[0, 0] = 1.0
            p[1:, 0] = U.T.dot(m_t - m_0)
            p[1:, 1:] = U.T.dot(self.Ctt).dot(V)
            P = np.linalg.matrix_power(S.dot(p), lag_multiple - 1).dot(S)
        Q = np.zeros((observables.shape[1], dim + 1))
        if not observables_mean_free:
            Q[:, 0] = observables.T.dot(m_t)
        Q[:, 1:] = observables.T.dot(self.Ctt).dot(V)
        if statistics is not None:
            R = np.zeros((statistics.shape[1], dim + 1))
            if not statistics_mean_free:
                R[:, 0] = statistics.T.dot(m_0)
            R[:, 1:] = statistics.T.dot(self.C00).dot(U)
            return Q.dot(P).dot(R.T)
        return Q.dot(P)[:, 0]

    def score(self, test_model=None, score_method="VAMP2"):
        """vamp.py:303-367: VAMP1 / VAMP2 / VAMPE (+1 for the constant singular function)."""
        if test_model is None:
            test_model = self
        Uk = self.U[:, 0:self.dimension()]
        Vk = self.V[:, 0:self.dimension()]
        res = None
        with _bounded_threads():
            if score_method == "VAMP1" or score_method == "VAMP2":
                A = spd_inv_sqrt(Uk.T.dot(test_model.C00).dot(Uk))
                B = Uk.T.dot(test_model.C0t).dot(Vk)
                C = spd_inv_sqrt(Vk.T.dot(test_model.Ctt).dot(Vk))
                ABC = _mdot(A, B, C)
                if score_method == "VAMP1":
                    res = np.linalg.norm(ABC, ord="nuc")
                else:
                    res = np.linalg.norm(ABC, ord="fro") ** 2
            elif score_method == "VAMPE":
                Sk = np.diag(self.singular_values[0:self.dimension()])
                res = np.trace(2.0 * _mdot(Vk, Sk, Uk.T, test_model.C0t)
                               - _mdot(Vk, Sk, Uk.T, test_model.C00, Uk, Sk, Vk.T, test_model.Ctt))
            else:
                raise ValueError('"score" should be one of VAMP1, VAMP2 or VAMPE')
        assert res
        return res + 1


class VAMP(object):
    """Variational approach for Markov processes (vamp.py:370-1006)."""

    def __init__(self, lag, dim=None, scaling=None, right=False, epsilon=1e-6, stride=1, skip=0,
                 ncov_max=float("inf"), precision="fp64"):
        self._model = VAMPModel()
        self._covar = None
        self._data = None
        self._estimated = False
        self.lag = int(lag)
        self.dim = dim
        self.scaling = scaling
        self.right = right
        self.epsilon = epsilon
        self.stride = stride
        self.skip = skip
        self.ncov_max = ncov_max
        self.precision = precision
        self.chunksize = None

    def describe(self):
        return "[VAMP, lag = %i; max. output dim. = %s]" % (self.lag, str(self.dim))

    # parameters live on the model (vamp.py:531-576)
    @property
    def model(self):
        return self._model

    @property
    def dim(self):
        return self._model.dim

    @dim.setter
    def dim(self, value):
        self._model.dim = value

    @property
    def epsilon(self):
        return self._model.epsilon

    @epsilon.setter
    def epsilon(self, value):
        self._model.epsilon = value

    @property
    def scaling(self):
        return self._model.scaling

    @scaling.setter
    def scaling(self, value):
        self._model.scaling = value

    def _new_covar(self):
        """vamp.py:578-596: C00, C0t, Ctt of the non-reversible, data-mean-free estimator."""
        return LaggedCovariance(c00=True, c0t=True, ctt=True, remove_data_mean=True, reversible=False, lag=self.lag,
                                bessel=False, stride=self.stride, skip=self.skip, weights=None,
                                ncov_max=self.ncov_max, precision=self.precision)

    def _update_model(self, covar):
        self._model.update_model_params(mean_0=covar.mean, mean_t=covar.mean_tau, C00=covar.C00_, C0t=covar.C0t_,
                                        Ctt=covar.Ctt_)

    def estimate(self, X, **kwargs):
        """vamp.py:604-623."""
        for k, v in kwargs.items():
            if k == "chunksize":
                self.chunksize = v
            elif hasattr(self, k):
                setattr(self, k, v)
        dt = to_device(X)
        if isinstance(self.dim, int) and not self.dim <= dt.d:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, dt.d))
        covar = self._new_covar()
        covar.estimate(dt)
        self._data = dt
        self._update_model(covar)
        self._model._diagonalize()
        self._estimated = True
        return self

    def fit(self, X, **kwargs):
        return self.estimate(X, **kwargs)

    def partial_fit(self, X):
        """vamp.py:625-655: the decomposition is deferred to the first access."""
        dt = to_device(X)
        if isinstance(self.dim, int) and not self.dim <= dt.d:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, dt.d))
        if self._covar is None:
            self._covar = self._new_covar()
        self._covar.partial_fit(dt)
        if self._covar._estimated:
            self._update_model(self._covar)
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
                                       statistics_mean_free=statistics_mean_free,
                                       observables_mean_free=observables_mean_free)

    def score(self, test_data=None, score_method="VAMP2"):
        """vamp.py:953-1006: self score, or cross-validation against a model of test_data with the same
        hyperparameters."""
        if test_data is None:
            return self._model.score(None, score_method=score_method)
        est = VAMP(self.lag, dim=self.dim, scaling=self.scaling, right=self.right, epsilon=self.epsilon,
                   stride=self.stride, skip=self.skip, ncov_max=self.ncov_max, precision=self.precision)
        est.estimate(test_data)
        return self._model.score(est.model, score_method=score_method)

    # ---- projection (vamp.py:661-687) on the GPU ----
    def _project_device(self, dt):
        m = int(self.dimension())
        if self.right:
            mean, M = self._model.mean_t, self._model.V[:, 0:m]
        else:
            mean, M = self._model.mean_0, self._model.U[:, 0:m]
        t = _native.torch()
        mean_d = t.as_tensor(np.ascontiguousarray(mean, dtype=np.float64), device=dt.X.device)
        M_d = t.as_tensor(np.ascontiguousarray(M, dtype=np.float64), device=dt.X.device)
        return _native.project(dt.X, mean_d, M_d, m)

    def transform(self, X):
        single = isinstance(X, np.ndarray) or (not isinstance(X, (list, tuple, DeviceTrajectories)))
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
