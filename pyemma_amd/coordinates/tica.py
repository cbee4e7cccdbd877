"""TICA: GPU lagged moments, host FP64 eigensolve, GPU projection.

Mirrors pyemma/coordinates/transform/tica.py:37-245 and _tica_base.py:22-239.
"""
import logging
import warnings

import numpy as np

from .. import _native
from .._solvers import eig_corr
from ..exceptions import ZeroRankError
from ._data import DeviceTrajectories, to_device
from .covariance import LaggedCovariance
from .source import FeatureSource

log = logging.getLogger("pyemma_amd.coordinates.transform.tica.TICA")


class TICAModel(object):
    """Model parameters (TICAModelBase, _tica_base.py:22-30)."""

    def __init__(self):
        self.mean = None
        self.cov = None
        self.cov_tau = None
        self.cumvar = None
        self.eigenvalues = None
        self.eigenvectors = None

    def update_model_params(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)


class TICA(object):
    r"""Time-lagged independent component analysis (TICA), tica.py:42-141."""

    _DEFAULT_VARIANCE_CUTOFF = 0.95

    def __init__(self, lag, dim=-1, var_cutoff=0.95, kinetic_map=True, commute_map=False, epsilon=1e-6,
                 stride=1, skip=0, reversible=True, weights=None, ncov_max=float("inf"), precision="bf16x4-acc"):
        if kinetic_map and commute_map:
            raise ValueError("Trying to use both kinetic_map and commute_map. Use either or.")
        if (kinetic_map or commute_map) and not reversible:
            kinetic_map = False
            commute_map = False
            warnings.warn("Cannot use kinetic_map or commute_map for non-reversible processes, both will be set to"
                          "False.")
        self._covar = None
        self._model = TICAModel()
        self._estimated = False
        self._data = None
        self.lag = int(lag)
        self.dim = dim
        self.var_cutoff = var_cutoff
        self.kinetic_map = kinetic_map
        self.commute_map = commute_map
        self.epsilon = epsilon
        self.stride = stride
        self.skip = skip
        self.reversible = reversible
        self.weights = weights
        self.ncov_max = ncov_max
        self.precision = precision
        self.chunksize = None

    # ---- parameters with the reference's validation (_tica_base.py:46-83) ----
    @property
    def dim(self):
        return self._dim

    @dim.setter
    def dim(self, value):
        if int(value) > -1:
            self._var_cutoff = 1.0
        self._dim = int(value)

    @property
    def var_cutoff(self):
        return self._var_cutoff

    @var_cutoff.setter
    def var_cutoff(self, value):
        v = float(value)
        if not hasattr(self, "_dim"):
            raise RuntimeError("need to set dim before var_cutoff")
        if not (0 < v <= 1.0):
            raise ValueError("variance cutoff has to be in interval (0, 1.0]")
        if v != TICA._DEFAULT_VARIANCE_CUTOFF and self.dim != -1 and v != 1.0:
            raise ValueError("Trying to set both the number of dimension and the subspace variance. "
                             "Use either one or the other.")
        self._var_cutoff = v

    @property
    def model(self):
        return self._model

    def describe(self):
        try:
            dim = self.dimension()
        except RuntimeError:
            dim = self.dim
        return "[TICA, lag = %i; max. output dim. = %i]" % (self.lag, dim)

    # ---- estimation ----
    def _new_covar(self):
        return LaggedCovariance(c00=True, c0t=True, ctt=False, remove_data_mean=True, reversible=self.reversible,
                                lag=self.lag, bessel=False, stride=self.stride, skip=self.skip,
                                weights=self.weights, ncov_max=self.ncov_max, precision=self.precision)

    def estimate(self, X, **kwargs):
        """Estimate mean, C00, C0t on the GPU, then diagonalize (tica.py:156-216)."""
        for k, v in kwargs.items():
            if k == "chunksize":
                self.chunksize = v
            elif hasattr(self, k):
                setattr(self, k, v)
        # a FeatureSource is streamed chunk by chunk (its features never all exist at once)
        dt = X if isinstance(X, FeatureSource) else to_device(X)
        if not self.dim <= dt.d:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, dt.d))
        covar = self._new_covar()
        covar.estimate(dt)
        self._covar = covar
        self._data = dt
        self.model.update_model_params(mean=covar.mean, cov=covar.C00_, cov_tau=covar.C0t_)
        self._diagonalize()
        return self

    def fit(self, X, **kwargs):
        return self.estimate(X, **kwargs)

    def partial_fit(self, X):
        """tica.py:165-195: eigendecomposition deferred to first property access."""
        dt = X if isinstance(X, FeatureSource) else to_device(X)
        if not self.dim <= dt.d:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, dt.d))
        if self._covar is None:
            self._covar = self._new_covar()
        self._covar.partial_fit(dt)
        if self._covar._estimated:
            self.model.update_model_params(mean=self._covar.mean, cov=self._covar.C00_, cov_tau=self._covar.C0t_)
        self._estimated = False
        return self

    def _diagonalize(self):
        """tica.py:218-245."""
        try:
            eigenvalues, eigenvectors = eig_corr(self.cov, self.cov_tau, self.epsilon, sign_maxelement=True)
        except ZeroRankError:
            raise ZeroRankError("All input features are constant in all time steps. No dimension would be left "
                                "after dimension reduction.")
        if self.kinetic_map and self.commute_map:
            raise ValueError("Trying to use both kinetic_map and commute_map. Use either or.")
        if self.kinetic_map:
            eigenvectors *= eigenvalues[None, :]
        if self.commute_map:
            timescales = 1 - self.lag / np.log(np.abs(eigenvalues))
            regularized_timescales = 0.5 * timescales * np.maximum(
                np.tanh(np.pi * ((timescales - self.lag) / self.lag) + 1), 0)
            eigenvectors *= np.sqrt(regularized_timescales / 2)
        cumvar = np.cumsum(np.abs(eigenvalues) ** 2)
        cumvar /= cumvar[-1]
        self.model.update_model_params(cumvar=cumvar, eigenvalues=eigenvalues, eigenvectors=eigenvectors)
        self._estimated = True

    def _lazy(self):
        if not self._estimated:
            if self.model.cov is None:
                raise RuntimeError("TICA has not been estimated")
            self._diagonalize()

    # ---- model properties (_tica_base.py:89-239) ----
    @property
    def mean(self):
        return self.model.mean

    @property
    def cov(self):
        return self.model.cov

    @property
    def cov_tau(self):
        return self.model.cov_tau

    @property
    def eigenvalues(self):
        self._lazy()
        return self.model.eigenvalues

    @property
    def eigenvectors(self):
        self._lazy()
        return self.model.eigenvectors

    @property
    def cumvar(self):
        self._lazy()
        return self.model.cumvar

    @property
    def timescales(self):
        self._lazy()
        return -self.lag / np.log(np.abs(self.eigenvalues))

    @property
    def feature_TIC_correlation(self):
        feature_sigma = np.sqrt(np.diag(self.cov))
        return np.dot(self.cov, self.eigenvectors[:, :self.dimension()]) / feature_sigma[:, np.newaxis]

    def dimension(self):
        """_tica_base.py:98-115."""
        if self.dim > -1:
            return self.dim
        if self._estimated or self.model.eigenvalues is not None:
            dim = len(self.eigenvalues)
            if self.var_cutoff < 1.0:
                dim = min(dim, np.searchsorted(self.cumvar, self.var_cutoff) + 1)
            return int(dim)
        if self.var_cutoff == 1.0 and self._data is not None:
            return self._data.d
        raise RuntimeError("Requested dimension, but the dimension depends on the cumulative variance and the "
                           "transformer has not yet been estimated. Call estimate() before.")

    def output_type(self):
        V = self.eigenvectors[:, 0:self.dimension()]
        if np.all(np.isreal(V)) or np.allclose(np.imag(V), 0):
            return np.float32()
        return np.complex64

    # ---- projection ----
    def _project_device(self, dt):
        m = self.dimension()
        V = self.eigenvectors[:, :m]
        if np.iscomplexobj(V):
            if not np.allclose(np.imag(V), 0):
                raise NotImplementedError("complex TICA eigenvectors (non-reversible, complex pair) are not "
                                          "supported by the GPU projection")
            V = np.real(V)
        t = _native.torch()
        mean = t.as_tensor(np.ascontiguousarray(self.mean, dtype=np.float64), device=dt.X.device)
        Vd = t.as_tensor(np.ascontiguousarray(V, dtype=np.float64), device=dt.X.device)
        return _native.project(dt.X, mean, Vd, m)

    def transform(self, X):
        """Project X (ndarray or list) onto the leading components (_tica_base.py:117-133)."""
        if isinstance(X, FeatureSource):
            return [y.cpu().numpy() for y in X.per_trajectory(self._project_device)]
        single = isinstance(X, np.ndarray) or (not isinstance(X, (list, tuple, DeviceTrajectories)))
        dt = to_device(X)
        Y = self._project_device(dt)
        out = dt.split(Y)
        return out[0] if single and dt.ntraj == 1 else out

    def transform_device(self, X):
        """Projection kept on the device: (DeviceTrajectories in) -> torch tensor (N, m)."""
        dt = to_device(X)
        return self._project_device(dt)

    def get_output(self, dimensions=slice(0, None), stride=1, skip=0, chunk=None):
        """Projected estimation data, one array per trajectory."""
        if self._data is None:
            raise RuntimeError("no data: estimate() with data first")
        if isinstance(self._data, FeatureSource):
            out = [y.cpu().numpy() for y in self._data.per_trajectory(self._project_device)]
        else:
            Y = self._project_device(self._data)
            out = self._data.split(Y)
        return [y[skip::stride][:, dimensions] for y in out]

    def fit_transform(self, X, **kwargs):
        self.estimate(X, **kwargs)
        return self.transform(X)
