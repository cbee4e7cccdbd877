"""Koopman reweighting of TICA (coordinates/estimation/koopman.py:28-140).

The first pass estimates the non-reversible, mean-free C00 and C0t and the
means of X and Y on the GPU (LaggedCovariance, reversible=False); the small
(d+1) x (d+1) eigenproblem of the Koopman matrix on the modified basis
(PC|1) is solved on the host in FP64.  The resulting linear weight model
w(x) = x . u + u0 is evaluated for every frame on the GPU (emma_row_affine)
and fed to the weighted moments of the second, reversible pass.
"""
import numpy as np
import scipy.linalg as scl

from .. import _native
from .._solvers import sort_by_norm, spd_inv_split
from ._data import DeviceTrajectories, to_device
from .covariance import LaggedCovariance

__all__ = ["KoopmanWeights", "KoopmanEstimator"]


def _compute_u(K):
    """koopman.py:28-45: eigenvector of K^T for the eigenvalue of largest modulus, last entry 1."""
    M = K.shape[0] - 1
    l, U = scl.eig(K.T)
    l, U = sort_by_norm(l, U)
    u = np.real(U[:, 0])
    v = np.zeros(M + 1)
    v[M] = 1.0
    return u / np.dot(u, v)


class KoopmanWeights(object):
    """koopman.py:48-55: w(x) = x . u + u_const."""

    def __init__(self, u, u_const):
        self._u = np.asarray(u, dtype=np.float64)
        self._u_const = float(u_const)

    def weights(self, X):
        return np.asarray(X).dot(self._u) + self._u_const

    def device_weights(self, Xt):
        """FP64 weights of every row of a device frame matrix."""
        return _native.row_affine(Xt, self._u, self._u_const)


class KoopmanEstimator(object):
    """koopman.py:58-140 (the estimator only computes u; TICA uses it)."""

    def __init__(self, lag, epsilon=1e-6, stride=1, skip=0, ncov_max=float("inf"), precision="fp64"):
        self.lag = lag
        self.epsilon = epsilon
        self.stride = stride
        self.skip = skip
        self.ncov_max = ncov_max
        self._covar = LaggedCovariance(c00=True, c0t=True, remove_data_mean=True, reversible=False, lag=lag,
                                       bessel=False, stride=stride, skip=skip, ncov_max=ncov_max,
                                       precision=precision)
        self._K = None
        self._R = None

    def estimate(self, data, **kwargs):
        self._covar.estimate(data, **kwargs)
        self._finish_estimation()
        return self

    def partial_fit(self, data):
        self._covar.partial_fit(data)
        self._K = None
        return self

    def _finish_estimation(self):
        """koopman.py:83-94."""
        R = spd_inv_split(self._covar.C00_, epsilon=self.epsilon, canonical_signs=True)
        M = R.shape[1]
        K = np.dot(R.T, np.dot(self._covar.C0t_, R))
        K = np.vstack((K, np.dot((self._covar.mean_tau - self._covar.mean), R)))
        ex1 = np.zeros((M + 1, 1))
        ex1[M, 0] = 1.0
        self._K = np.hstack((K, ex1))
        self._R = R

    @property
    def K_pc_1(self):
        if self._K is None:
            self._finish_estimation()
        return self._K

    @property
    def u_pc_1(self):
        return _compute_u(self.K_pc_1)

    @property
    def mean(self):
        return self._covar.mean

    @property
    def R(self):
        if self._K is None:
            self._finish_estimation()
        return self._R

    @property
    def u(self):
        """koopman.py:116-125: weights in the input basis."""
        u_mod = self.u_pc_1
        R = self.R
        N = R.shape[0]
        u_input = np.zeros(N + 1)
        u_input[0:N] = R.dot(u_mod[0:-1])
        u_input[N] = u_mod[-1] - self.mean.dot(R.dot(u_mod[0:-1]))
        return u_input

    @property
    def weights(self):
        u_input = self.u
        return KoopmanWeights(u_input[0:-1], u_input[-1])


def frame_weights(weights, dt):
    """Per-row FP64 device weights of DeviceTrajectories dt (None: unweighted).

    Accepts what LaggedCovariance.weights accepts (covariance.py:130-158): a
    number (the same weight for every frame), one array per trajectory (or one
    array for a single trajectory), or an object with a weights(X) method
    evaluated per trajectory.
    """
    import numbers

    t = _native.torch()
    if weights is None:
        return None
    if isinstance(weights, numbers.Number):
        return t.full((dt.n_frames,), float(weights), dtype=t.float64, device=dt.X.device)
    if isinstance(weights, KoopmanWeights):
        return weights.device_weights(dt.X)
    if hasattr(weights, "weights") and callable(getattr(weights, "weights")):
        parts = [np.asarray(weights.weights(dt.trajectory(i).cpu().numpy()), dtype=np.float64).ravel()
                 for i in range(dt.ntraj)]
    else:
        if isinstance(weights, np.ndarray) and weights.ndim == 1 and dt.ntraj == 1:
            weights = [weights]
        if not isinstance(weights, (list, tuple)):
            raise ValueError("weights must be a number, arrays, or an object with a weights(X) method")
        if len(weights) != dt.ntraj:
            raise ValueError("number of weight arrays did not match number of input data sets. {} vs. {}"
                             .format(len(weights), dt.ntraj))
        parts = [np.asarray(w, dtype=np.float64).ravel() for w in weights]
    for i, w in enumerate(parts):
        if w.size < dt.lengths[i]:
            raise ValueError("weight array %d has %d entries for a trajectory of %d frames"
                             % (i, w.size, dt.lengths[i]))
    cat = np.concatenate([w[:dt.lengths[i]] for i, w in enumerate(parts)])
    return t.as_tensor(cat, device=dt.X.device)
