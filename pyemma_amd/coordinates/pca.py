"""PCA: GPU instantaneous moments (lag 0, mean-free), host FP64 eigh, GPU projection.

Mirrors pyemma/coordinates/transform/pca.py:47-265.  The covariance is the
lag-0 branch of the moments kernels (`emma_moments_lagged` with lag 0 and
REMOVE_MEAN: `RunningCovar(xx=True, xy=False, yy=False, remove_mean=True)`
of pca.py:202-216), normalized with Bessel's correction as `cov_XX(bessel=True)`
(pca.py:247); the eigendecomposition is numpy's eigh on the host, as in
pca.py:218-231.
"""
import numpy as np

from .. import _native
from .._solvers import _bounded_threads
from ._data import DeviceTrajectories, to_device
from .covariance import LaggedCovariance

__all__ = ["PCAModel", "PCA"]


class PCAModel(object):
    """pca.py:47-53."""

    def __init__(self):
        self.mean = None
        self.eigenvectors = None
        self.eigenvalues = None
        self.cumvar = None

    def update_model_params(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)


class PCA(object):
    """Principal component analysis (pca.py:56-265)."""

    _DEFAULT_VAR_CUTOFF = 0.95

    def __init__(self, dim=-1, var_cutoff=0.95, mean=None, stride=1, skip=0, precision="fp64"):
        if dim != -1 and var_cutoff != PCA._DEFAULT_VAR_CUTOFF:
            raise ValueError("Trying to set both the number of dimension and the subspace variance. Use either or.")
        self._model = PCAModel()
        self._covar = None
        self._data = None
        self._estimated = False
        self.dim = dim
        self.var_cutoff = var_cutoff
        self.stride = stride
        self.skip = skip
        self.precision = precision
        self.cov = None
        self.mu = None
        self.chunksize = None
        self._model.mean = mean

    def describe(self):
        return "[PCA, output dimension = %i]" % self.dim

    def dimension(self):
        """pca.py:117-133."""
        if self.dim != -1:
            return self.dim
        if self._estimated:
            dim = len(self._model.eigenvalues)
            if self.var_cutoff < 1.0:
                dim = min(dim, np.searchsorted(self.cumvar, self.var_cutoff) + 1)
            return int(dim)
        if self.var_cutoff == 1.0 and self._data is not None:
            return self._data.d
        raise RuntimeError("Requested dimension, but the dimension depends on the cumulative variance and the "
                           "transformer has not yet been parametrized. Call parametrize() before.")

    def _lazy(self):
        if not self._estimated:
            if self.cov is None:
                raise RuntimeError("PCA has not been estimated")
            self._diagonalize()
            self._estimated = True

    @property
    def cumvar(self):
        self._lazy()
        return self._model.cumvar

    @property
    def eigenvalues(self):
        self._lazy()
        return self._model.eigenvalues

    @property
    def eigenvectors(self):
        self._lazy()
        return self._model.eigenvectors

    @property
    def mean(self):
        return self._model.mean

    @property
    def model(self):
        return self._model

    def _new_covar(self):
        return LaggedCovariance(c00=True, c0t=False, ctt=False, remove_data_mean=True, reversible=False, lag=0,
                                bessel=True, stride=self.stride, skip=self.skip, precision=self.precision)

    def _diagonalize(self):
        """pca.py:218-231: eigh, descending order, cumulative variance."""
        with _bounded_threads():
            v, R = np.linalg.eigh(self.cov)
        I = np.argsort(v)[::-1]
        eigenvalues = v[I]
        eigenvectors = R[:, I]
        cumvar = np.cumsum(eigenvalues)
        cumvar /= cumvar[-1]
        self._model.update_model_params(eigenvalues=eigenvalues, eigenvectors=eigenvectors, cumvar=cumvar)

    def _estimate(self, dt, partial):
        if self._covar is None or not partial:
            self._covar = self._new_covar()
        if partial:
            self._covar.partial_fit(dt)
        else:
            self._covar.estimate(dt)
        self.cov = self._covar.C00_
        self.mu = self._covar.mean
        self._model.update_model_params(mean=self._covar.mean)

    def estimate(self, X, **kwargs):
        """pca.py:233-255."""
        for k, v in kwargs.items():
            if k == "chunksize":
                self.chunksize = v
            elif hasattr(self, k):
                setattr(self, k, v)
        dt = to_device(X)
        self._data = dt
        self._estimate(dt, partial=False)
        self._diagonalize()
        self._estimated = True
        return self

    def fit(self, X, **kwargs):
        return self.estimate(X, **kwargs)

    def partial_fit(self, X):
        """pca.py:168-174."""
        dt = to_device(X)
        if self._data is None:
            self._data = dt
        self._estimate(dt, partial=True)
        self._estimated = False
        return self

    @property
    def feature_PC_correlation(self):
        """pca.py:176-200."""
        feature_sigma = np.array(np.sqrt(np.diag(self.cov)), ndmin=2)
        PC_sigma = np.array(np.sqrt(self.eigenvalues[:self.dimension()]), ndmin=2)
        return np.dot(self.cov, self.eigenvectors[:, : self.dimension()]) / feature_sigma.T.dot(PC_sigma)

    # ---- projection (pca.py:257-265) on the GPU ----
    def _project_device(self, dt):
        m = int(self.dimension())
        t = _native.torch()
        mean = t.as_tensor(np.ascontiguousarray(self.mean, dtype=np.float64), device=dt.X.device)
        V = t.as_tensor(np.ascontiguousarray(self.eigenvectors[:, :m], dtype=np.float64), device=dt.X.device)
        return _native.project(dt.X, mean, V, m)

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
