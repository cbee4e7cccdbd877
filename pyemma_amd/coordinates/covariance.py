"""LaggedCovariance: lagged (co)moments on the GPU, FP64 Chan storage on the host.

Mirrors pyemma/coordinates/estimation/covariance.py:36-343.  One estimate()
call hands every trajectory to emma_moments_lagged in a single launch
sequence (the device merges its tiles exactly, in FP64); partial_fit() calls
are merged on the host with the pairwise update of
_ext/variational/estimators/running_moments.py:35-148.
"""
import logging
import math
import numbers

import numpy as np

from .. import _native
from ._data import to_device

log = logging.getLogger("pyemma_amd.coordinates.estimation.covariance.LaggedCovariance")

PRECISIONS = {"bf16x4": _native.EMMA_PREC_BF16X4, "fp32": _native.EMMA_PREC_FP32, "fp64": _native.EMMA_PREC_FP64,
              "bf16x4-acc": _native.EMMA_PREC_BF16X4_ACC}


class Moments(object):
    """running_moments.py:9-85."""

    def __init__(self, w, sx, sy, Mxy):
        self.w = float(w)
        self.sx = sx
        self.sy = sy
        self.Mxy = Mxy

    def combine(self, other, mean_free=False):
        """Chan et al. pairwise update (running_moments.py:35-61)."""
        w1, w2 = self.w, other.w
        w = w1 + w2
        q = w2 / w1
        dsx = q * self.sx - other.sx
        dsy = q * self.sy - other.sy
        self.w = w
        self.sx = self.sx + other.sx
        self.sy = self.sy + other.sy
        if mean_free:
            d = dsx * dsy if self.Mxy.ndim == 1 else np.outer(dsx, dsy)
            self.Mxy = self.Mxy + other.Mxy + (w1 / (w2 * w)) * d
        else:
            self.Mxy = self.Mxy + other.Mxy
        return self

    @property
    def mean_x(self):
        return self.sx / self.w

    @property
    def mean_y(self):
        return self.sy / self.w

    def covar(self, bessel=True):
        return self.Mxy / (self.w - 1) if bessel else self.Mxy / self.w


class MomentsStorage(object):
    """running_moments.py:88-148."""

    def __init__(self, nsave, remove_mean=False, rtol=1.5):
        self.nsave = nsave
        self.storage = []
        self.rtol = rtol
        self.remove_mean = remove_mean

    def _can_merge_tail(self):
        return len(self.storage) >= 2 and self.storage[-2].w <= self.storage[-1].w * self.rtol

    def store(self, moments):
        if len(self.storage) == self.nsave:
            self.storage[-1].combine(moments, mean_free=self.remove_mean)
        else:
            self.storage.append(moments)
        while self._can_merge_tail():
            M = self.storage.pop()
            self.storage[-1].combine(M, mean_free=self.remove_mean)

    @property
    def moments(self):
        while len(self.storage) > 1:
            M = self.storage.pop()
            self.storage[-1].combine(M, mean_free=self.remove_mean)
        return self.storage[0]


class LaggedCovariance(object):
    r"""Compute lagged covariances between time series (GPU moments).

    Parameters follow covariance.py:96-125; `precision` selects the
    covariance-tile arithmetic: 'bf16x4' (bf16 hi/lo split on MFMA) and 'fp32'
    (FP32 MFMA) accumulate tiles in FP32 over at most 1024 pairs and merge in
    FP64; 'fp64' runs FP64 MFMA with FP64 accumulation (agrees with the FP64
    reference computation to rounding, also for near-zero eigenvalues).
    'bf16x4-acc' (the default) runs 'bf16x4' with two-level FP32 accumulation
    (sub-spans of 64 pairs added into the 1024-pair span sum; TICA eigenvalues
    within 1e-6 relative of FP64 on 300 collinear distance features).
    """

    def __init__(self, c00=True, c0t=False, ctt=False, remove_constant_mean=None, remove_data_mean=False,
                 reversible=False, bessel=True, sparse_mode="auto", modify_data=False, lag=0, weights=None,
                 stride=1, skip=0, chunksize=NotImplemented, ncov_max=float("inf"), column_selection=None,
                 diag_only=False, precision="bf16x4-acc"):
        if (c0t or ctt) and lag == 0:
            raise ValueError("lag must be positive if c0t=True or ctt=True")
        if remove_constant_mean is not None and remove_data_mean:
            raise ValueError("Subtracting the data mean and a constant vector simultaneously is not supported.")
        if remove_constant_mean is not None:
            remove_constant_mean = np.asarray(remove_constant_mean, dtype=np.float64).ravel()
        if column_selection is not None and diag_only:
            raise ValueError("Computing only parts of the diagonal is not supported.")
        if ctt and reversible:
            raise ValueError("Combining compute_YY and symmetrize=True is meaningless.")
        if precision not in PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(PRECISIONS))
        self.c00, self.c0t, self.ctt = c00, c0t, ctt
        self.remove_constant_mean = remove_constant_mean
        self.remove_data_mean = remove_data_mean
        self.reversible = reversible
        self.bessel = bessel
        self.sparse_mode = sparse_mode
        self.modify_data = modify_data
        self.lag = int(lag)
        self.weights = weights
        self.stride = int(stride)
        self.skip = int(skip)
        self.ncov_max = ncov_max
        self.column_selection = None if column_selection is None else np.asarray(column_selection)
        self.diag_only = diag_only
        self.precision = precision
        self.chunksize = None
        self._rc = None
        self._estimated = False

    # ------------------------------------------------------------------
    def _init_storage(self, n_chunks):
        nsave = min(int(max(math.log(max(n_chunks, 1), 2), 2)), self.ncov_max)
        rm = self.remove_data_mean
        self._rc = {
            "xx": MomentsStorage(nsave, remove_mean=rm) if self.c00 or not self.c0t else None,
            "xy": MomentsStorage(nsave, remove_mean=rm) if self.c0t else None,
            "yy": MomentsStorage(nsave, remove_mean=rm) if self.ctt else None,
        }

    def _moments_launch(self, dt, skip=None):
        """The moments call of dt; its results stay on the device (no host sync)."""
        from .koopman import frame_weights

        wts = frame_weights(self.weights, dt)
        if wts is not None and self.ctt:
            # running_moments.py:255-256
            raise ValueError("Use of weights is not implemented for compute_YY==True")
        return _native.moments_lagged(
            dt.X, dt.offsets_dev, self.lag, stride=self.stride, skip=self.skip if skip is None else skip,
            reversible=self.reversible,
            remove_mean=self.remove_data_mean, c0t=self.c0t, ctt=self.ctt,
            precision=PRECISIONS[self.precision], weights=wts)

    def _device_moments(self, dt, skip=None):
        return self._moments_finish(self._moments_launch(dt, skip=skip))

    def _moments_finish(self, dev):
        """Host FP64 moments of a launched call, with the constant-mean shift."""
        lag = self.lag
        w, sx, sy, Mxx, Mxy, Myy = dev
        h = lambda t: None if t is None else t.cpu().numpy()
        w = float(w.item())
        sx, sy, Mxx, Mxy, Myy = h(sx), h(sy), h(Mxx), h(Mxy), h(Myy)
        if lag == 0 or (self.c00 and not self.c0t and not self.ctt):
            sy = sx
        c = self.remove_constant_mean
        if c is not None:
            # moments of (x - c): exact algebra on the unshifted FP64 sums
            if c.size != sx.size:
                raise ValueError("remove_constant_mean has wrong dimension")
            sx0, sy0 = sx, sy
            sx = sx0 - w * c
            sy = sy0 - w * c
            corr = lambda M, a, b: M - np.outer(a, c) - np.outer(c, b) + w * np.outer(c, c)
            Mxx = corr(Mxx, sx0, sx0)
            if Mxy is not None:
                Mxy = corr(Mxy, sx0, sy0)
            if Myy is not None:
                Myy = corr(Myy, sy0, sy0)
        return w, sx, sy, Mxx, Mxy, Myy

    def _estimate_source(self, src, partial_fit=False):
        """Chunks of a FeatureSource: one moments call per chunk of trajectory
        segments (the skip is applied by the segmentation, so each call reads
        its segments with skip 0), merged by the storage's Chan update like
        the reference's chunks (covariance.py:219-244)."""
        import numbers

        if self.weights is not None and not isinstance(self.weights, numbers.Number) \
                and not hasattr(self.weights, "weights") and not hasattr(self.weights, "device_weights"):
            raise ValueError("per-trajectory weight arrays are not supported with a streamed source; "
                             "use a weights object or a number")
        n_chunks = src.n_chunks(lag=self.lag, stride=self.stride, skip=self.skip)
        if n_chunks == 0:
            longest = int(max(src.trajectory_lengths()) if src.ntraj else 0)
            if partial_fit:
                log.warning("Could not use data passed to partial_fit(), because no single data set "
                            "[longest=%i] is longer than lag+skip [%i]", longest, self.lag + self.skip)
                return self
            raise ValueError("None single dataset [longest=%i] is longer than lag+skip [%i]."
                             % (longest, self.lag + self.skip))
        if self._rc is None or not partial_fit:
            self._init_storage(n_chunks)
        # every chunk's moments call is queued before any result is read back,
        # so the device runs the chunks back to back; the storage then merges
        # them in chunk order
        pending = [self._moments_launch(dt, skip=0)
                   for _, dt in src.iter_chunks(lag=self.lag, stride=self.stride, skip=self.skip)]
        for dev in pending:
            self._store(*self._moments_finish(dev))
        self._estimated = True
        return self

    def _store(self, w, sx, sy, Mxx, Mxy, Myy):
        if self._rc["xx"] is not None:
            self._rc["xx"].store(Moments(w, sx, sx, Mxx))
        if self._rc["xy"] is not None:
            self._rc["xy"].store(Moments(w, sx, sy, Mxy))
        if self._rc["yy"] is not None:
            self._rc["yy"].store(Moments(w, sy, sy, Myy))

    def _estimate(self, data, partial_fit=False):
        from .source import FeatureSource

        if isinstance(data, FeatureSource):
            return self._estimate_source(data, partial_fit=partial_fit)
        dt = to_device(data)
        usable = [n for n in dt.lengths if n > self.lag + self.skip]
        if not usable:
            longest = int(max(dt.lengths) if dt.ntraj else 0)
            if partial_fit:
                log.warning("Could not use data passed to partial_fit(), because no single data set "
                            "[longest=%i] is longer than lag+skip [%i]", longest, self.lag + self.skip)
                return self
            raise ValueError("None single dataset [longest=%i] is longer than lag+skip [%i]."
                             % (longest, self.lag + self.skip))
        if self._rc is None or not partial_fit:
            self._init_storage(len(usable))
        self._store(*self._device_moments(dt))
        self._estimated = True
        return self

    def estimate(self, X, **kwargs):
        for k, v in kwargs.items():
            if k == "chunksize":
                self.chunksize = v
            elif hasattr(self, k):
                setattr(self, k, v)
        return self._estimate(X, partial_fit=False)

    def fit(self, X, **kwargs):
        return self.estimate(X, **kwargs)

    def partial_fit(self, X):
        self._estimate(X, partial_fit=True)
        return self

    # ------------------------------------------------------------------
    def _check_estimated(self):
        if not self._estimated:
            raise RuntimeError("LaggedCovariance has not been estimated")

    def _select(self, C):
        if self.diag_only:
            return np.diag(C).copy()
        if self.column_selection is not None:
            return C[:, self.column_selection]
        return C

    @property
    def mean(self):
        """covariance.py:269-272 (RunningCovar.mean_X, running_moments.py:326-332)."""
        self._check_estimated()
        if self._rc["xx"] is not None:
            return self._rc["xx"].moments.mean_x
        return self._rc["xy"].moments.mean_y

    @property
    def mean_tau(self):
        self._check_estimated()
        if self._rc["xy"] is not None:
            return self._rc["xy"].moments.mean_y
        return self._rc["yy"].moments.mean_y

    @property
    def C00_(self):
        self._check_estimated()
        return self._select(self._rc["xx"].moments.covar(bessel=self.bessel))

    @property
    def C0t_(self):
        self._check_estimated()
        return self._select(self._rc["xy"].moments.covar(bessel=self.bessel))

    @property
    def Ctt_(self):
        self._check_estimated()
        return self._select(self._rc["yy"].moments.covar(bessel=self.bessel))

    # deprecated aliases of the reference
    @property
    def cov(self):
        return self.C00_

    @property
    def cov_tau(self):
        return self.C0t_

    @property
    def nsave(self):
        for key in ("xx", "xy", "yy"):
            if self._rc and self._rc[key] is not None:
                return self._rc[key].nsave
        return None
