"""FP64 restatement of the reference's lagged moment estimator.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Every function cites the reference code it restates.  Only the dense code
path is restated: for floating-point input with the default
``sparse_tol=0.0`` the reference's constant-column detector marks every
column variable on the first row (covar_c/covartools_cpp.h:34-42, the test
``diff >= tol`` is ``0 >= 0``), so the sparse/constant branches of
moments.py are never taken for TICA on float data.

Pinned by (tests/test_oracle_moments.py):
  * the closed forms of the reference's own tests
    (coordinates/tests/test_covar_estimator.py:19-90,
    coordinates/tests/test_tica.py:203-230), rtol 1e-10;
  * chunk-size invariance of the Chan merge (test_tica.py:143-163);
  * a brute-force loop over individual pairs in exact rational arithmetic
    (fractions.Fraction) on tiny inputs.
"""
import math

import numpy as np

__all__ = ["lagged_pairs", "moments_XX", "moments_XXXY", "moments_block", "Moments",
           "MomentsStorage", "covariance_lagged", "n_pairs"]


def n_pairs(T, lag, stride=1, skip=0):
    """Number of (x_t, x_{t+lag}) pairs a trajectory of length T yields.

    Restates the legacy lagged iterator (iterable.py:152-160): X frames are
    traj[skip::stride], Y frames traj[skip+lag::stride], truncated to the
    shorter; trajectories with T <= skip+lag yield none (iterable.py:205-207,
    covariance.py:181-190)."""
    start = skip + lag
    if T <= start:
        return 0
    return (T - start + stride - 1) // stride


def lagged_pairs(traj, lag, stride=1, skip=0):
    """(X, Y) arrays of one trajectory; same pairing as iterable.py:95-160."""
    n = n_pairs(len(traj), lag, stride, skip)
    X = traj[skip: skip + n * stride: stride]
    Y = traj[skip + lag: skip + lag + n * stride: stride]
    assert len(X) == len(Y) == n
    return X, Y


def _f64(X):
    # moments.py:181-224 (_copy_convert): float input is copied to float64.
    return np.array(X, dtype=np.float64, order="C", copy=True)


def moments_XX(X, remove_mean=False, weights=None):
    """moments.py:601-696 restricted to the dense path (weights: moments.py:267-303, 540-553)."""
    X0 = _f64(X)
    T = X0.shape[0]
    wt = None if weights is None else np.asarray(weights, dtype=np.float64).ravel()[:T]
    w = float(T) if wt is None else float(wt.sum())
    sx = X0.sum(axis=0) if wt is None else (wt[:, None] * X0).sum(axis=0)   # moments.py:270-281
    if remove_mean:
        X0 -= sx / w                          # moments.py:326-358
    Cxx = X0.T.dot(X0) if wt is None else (wt[:, None] * X0).T.dot(X0)    # moments.py:545-553
    return w, sx, Cxx


def moments_XXXY(X, Y, remove_mean=False, symmetrize=False, weights=None):
    """moments.py:699-839 restricted to the dense path.

    symmetrize: w = 2T, sx = sy = sum X + sum Y (moments.py:290-296),
    Cxx = X0'X0 + Y0'Y0, Cxy = X0'Y0 + Y0'X0 (moments.py:548-559).
    Otherwise w = T, Cxx = X0'X0, Cxy = X0'Y0 with separately centered X, Y.
    weights (one per pair, the weight of its x frame) multiply every term
    and T becomes their sum (moments.py:267-303, 540-553).
    """
    X0 = _f64(X)
    Y0 = _f64(Y)
    T = X0.shape[0]
    wt = np.ones(T) if weights is None else np.asarray(weights, dtype=np.float64).ravel()[:T]
    sx_raw = (wt[:, None] * X0).sum(axis=0)
    sy_raw = (wt[:, None] * Y0).sum(axis=0)
    if symmetrize:
        sx = sx_raw + sy_raw
        sy = sx
        w = 2.0 * wt.sum()
    else:
        sx, sy, w = sx_raw, sy_raw, float(wt.sum())
    if remove_mean:
        X0 -= sx / w
        Y0 -= sy / w
    WX, WY = wt[:, None] * X0, wt[:, None] * Y0
    if symmetrize:
        Cxx = WX.T.dot(X0) + WY.T.dot(Y0)
        Cxy = WX.T.dot(Y0) + WY.T.dot(X0)
    else:
        Cxx = WX.T.dot(X0)
        Cxy = WX.T.dot(Y0)
    return w, sx, sy, Cxx, Cxy


def moments_block(X, Y, remove_mean=False):
    """moments.py:842-973 (dense path): XX, XY, YX, YY blocks, not symmetrized."""
    X0 = _f64(X)
    Y0 = _f64(Y)
    w = float(X0.shape[0])
    sx = X0.sum(axis=0)
    sy = Y0.sum(axis=0)
    if remove_mean:
        X0 -= sx / w
        Y0 -= sy / w
    Cxx = X0.T.dot(X0)
    Cxy = X0.T.dot(Y0)
    Cyy = Y0.T.dot(Y0)
    return w, (sx, sy), ((Cxx, Cxy), (Cxy.T, Cyy))


class Moments(object):
    """running_moments.py:9-85: weight, sums and (co)moment matrix."""

    def __init__(self, w, sx, sy, Mxy):
        self.w = float(w)
        self.sx = sx
        self.sy = sy
        self.Mxy = Mxy

    def combine(self, other, mean_free=False):
        """Chan et al. pairwise update, running_moments.py:35-61."""
        w1, w2 = self.w, other.w
        w = w1 + w2
        q = w2 / w1
        dsx = q * self.sx - other.sx
        dsy = q * self.sy - other.sy
        self.w = w
        self.sx = self.sx + other.sx
        self.sy = self.sy + other.sy
        if mean_free:
            self.Mxy = self.Mxy + other.Mxy + (w1 / (w2 * w)) * np.outer(dsx, dsy)
        else:
            self.Mxy = self.Mxy + other.Mxy
        return self

    def covar(self, bessel=True):
        """running_moments.py:71-85."""
        return self.Mxy / (self.w - 1) if bessel else self.Mxy / self.w


class MomentsStorage(object):
    """running_moments.py:88-148: bounded list of partial moments."""

    def __init__(self, nsave, remove_mean=False, rtol=1.5):
        self.nsave = nsave
        self.storage = []
        self.rtol = rtol
        self.remove_mean = remove_mean

    def _can_merge_tail(self):
        if len(self.storage) < 2:
            return False
        return self.storage[-2].w <= self.storage[-1].w * self.rtol

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


def covariance_lagged(trajs, lag, c00=True, c0t=True, ctt=False, remove_data_mean=False,
                      reversible=False, bessel=True, stride=1, skip=0, chunksize=None,
                      ncov_max=float("inf"), weights=None):
    """FP64 lagged covariances of a list of trajectories.

    Restates LaggedCovariance._estimate / _init_covar (covariance.py:160-252)
    driving RunningCovar.add (running_moments.py:233-308): every trajectory
    is cut into chunks of ``chunksize`` pairs (None: whole trajectory), each
    chunk's moments are stored with the Chan merge.  Returns a dict with
    w, sx, sy, Mxx, Mxy, Myy (unnormalized, as stored) and mean, mean_tau,
    C00, C0t, Ctt as the estimator exposes them (covariance.py:269-307).
    """
    if isinstance(trajs, np.ndarray):
        trajs = [trajs]
    if (c0t or ctt) and lag == 0:
        raise ValueError("lag must be positive if c0t=True or ctt=True")
    chunks = []
    for i, traj in enumerate(trajs):
        traj = np.asarray(traj)
        if traj.ndim == 1:
            traj = traj[:, None]
        if lag > 0:
            X, Y = lagged_pairs(traj, lag, stride, skip)
        else:
            X = traj[skip::stride]
            Y = None
        n = len(X)
        if n == 0:
            continue
        # frame weights follow the frames (covariance.py:198-227): the weight
        # of a pair is the weight of its x frame
        W = None if weights is None else np.asarray(weights[i], dtype=np.float64).ravel()[skip::stride][:n]
        cs = n if not chunksize else int(chunksize)
        for lo in range(0, n, cs):
            chunks.append((X[lo:lo + cs], None if Y is None else Y[lo:lo + cs],
                           None if W is None else W[lo:lo + cs]))
    if not chunks:
        raise ValueError("None single dataset is longer than lag+skip [%i]." % (lag + skip))
    nsave = min(int(max(math.log(len(chunks), 2), 2)), ncov_max)
    sxx = MomentsStorage(nsave, remove_mean=remove_data_mean)
    sxy = MomentsStorage(nsave, remove_mean=remove_data_mean) if c0t else None
    syy = MomentsStorage(nsave, remove_mean=remove_data_mean) if ctt else None
    for X, Y, W in chunks:
        if Y is None or (c00 and not c0t and not ctt):
            w, sx, Cxx = moments_XX(X, remove_mean=remove_data_mean, weights=W)
            sxx.store(Moments(w, sx, sx, Cxx))
        elif not ctt:
            w, sx, sy, Cxx, Cxy = moments_XXXY(X, Y, remove_mean=remove_data_mean,
                                               symmetrize=reversible, weights=W)
            sxx.store(Moments(w, sx, sx, Cxx))
            sxy.store(Moments(w, sx, sy, Cxy))
        else:
            if reversible:
                raise ValueError("Combining compute_YY and symmetrize=True is meaningless.")
            if W is not None:
                raise NotImplementedError("weighted Ctt is not restated")
            w, s, C = moments_block(X, Y, remove_mean=remove_data_mean)
            sxx.store(Moments(w, s[0], s[0], C[0][0]))
            if c0t:
                sxy.store(Moments(w, s[0], s[1], C[0][1]))
            syy.store(Moments(w, s[1], s[1], C[1][1]))
    out = {}
    mxx = sxx.moments
    out.update(w=mxx.w, sx=mxx.sx, Mxx=mxx.Mxy, mean=mxx.sx / mxx.w, C00=mxx.covar(bessel))
    if sxy is not None:
        mxy = sxy.moments
        out.update(sy=mxy.sy, Mxy=mxy.Mxy, mean_tau=mxy.sy / mxy.w, C0t=mxy.covar(bessel))
    if syy is not None:
        myy = syy.moments
        out.update(Myy=myy.Mxy, Ctt=myy.covar(bessel))
        out.setdefault("mean_tau", myy.sy / myy.w)
        out.setdefault("sy", myy.sy)
    return out
