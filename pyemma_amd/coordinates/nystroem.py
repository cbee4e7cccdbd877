"""Nystroem TICA: column-sampled TICA on one GPU moments pass.

Drop-in for pyemma.coordinates.tica_nystroem / NystroemTICA and the oASIS
column sampler (pyemma/coordinates/transform/nystroem_tica.py:40-786, API
api.py:1446-1551).  The reference streams the data once per selection
round to get the covariance columns of the newly selected features
(LaggedCovariance with column_selection).  Here one pass of the moments
kernels yields every column of C00 and C0t at the cost of one such round;
the selection rounds then read host copies of those columns, so the
selected columns, the error vector and the eigenproblem are the same
numbers, without re-reading the frames.

Column sampling (oASIS: R. Patel, T. Goldstein, E. Dyer, A. Mirhoseini,
R. Baraniuk, "oASIS: Adaptive Column Sampling for Kernel Matrix
Approximation", arXiv:1505.05208, Algorithm 1; spectral oASIS: F. Litzinger,
L. Boninsegna, H. Wu, F. Nueske, R. Patel, R. Baraniuk, F. Noe, C. Clementi,
"Rapid calculation of molecular kinetics using compressed sensing",
J. Chem. Theory Comput. 14, 2771 (2018)).  For an SPD matrix A (here C00)
known through its diagonal and a growing set S of columns:

* W = A[S, S], C = A[:, S], R = W^-1 C^T; the Nystroem approximation is
  A ~ C R and its diagonal error e_i = (C R)_ii - A_ii (<= 0);
* adding column j borders W by b = A[S, j] and A_jj; with q = W^-1 b and
  the Schur complement sigma = A_jj - b^T q (zero: j adds nothing),
      W'^-1 = [[W^-1 + q q^T / sigma, -q / sigma], [-q^T / sigma, 1 / sigma]],
      R'    = [[R + q (b^T R - a_j^T) / sigma], [(a_j^T - b^T R) / sigma]],
  a_j = A[:, j]; no inverse is recomputed;
* selection: oASIS takes the largest |e_i|; spectral oASIS spreads a batch
  over the approximate eigenvectors weighted by the error (greedy farthest
  points); 'random' draws uniformly among the unselected columns.
"""
import logging
import warnings
from types import FunctionType

import numpy as np

from .._solvers import eig_corr, sort_by_norm, spd_inv_split
from ..exceptions import ZeroRankError
from ._data import to_device
from .covariance import LaggedCovariance
from .tica import TICA

log = logging.getLogger("pyemma_amd.coordinates.transform.nystroem_tica.NystroemTICA")

__all__ = ["oASIS_Nystroem", "NystroemTICA", "selection_strategy"]


class oASIS_Nystroem(object):
    """Nystroem approximation of an SPD matrix from a growing set of its columns.

    d        the matrix diagonal (n,)
    C_k      the known columns (n, k)
    columns  their indices (k,)
    """

    def __init__(self, d, C_k, columns):
        self._diag = np.array(d, dtype=np.float64)
        self._cols = np.array(columns, dtype=int).ravel()
        self._C = np.array(C_k, dtype=np.float64).reshape(self._diag.shape[0], self._cols.shape[0])
        self._recompute()
        self._picker = selection_strategy(self, "spectral-oasis", nsel=len(self._cols))

    # ---- state ----
    @property
    def n(self):
        return self._diag.shape[0]

    @property
    def k(self):
        return self._cols.shape[0]

    @property
    def Ck(self):
        return self._C

    @property
    def Wk(self):
        return self._C[self._cols, :]

    @property
    def Wk_inv(self):
        return self._Winv

    @property
    def Rk(self):
        return self._R

    @property
    def diag(self):
        return self._diag

    @property
    def column_indices(self):
        return np.array(self._cols)

    @property
    def error(self):
        return self._err

    def _refresh_error(self):
        # diagonal of C R minus the true diagonal
        self._err = np.sum(self._C * self._R.T, axis=1) - self._diag

    def _recompute(self):
        """W^-1 (pseudo-inverse) and R from scratch."""
        self._Winv = np.linalg.pinv(self.Wk)
        self._R = self._Winv.dot(self._C.T)
        self._refresh_error()

    update_inverse = _recompute

    # ---- selection ----
    def set_selection_strategy(self, strategy="spectral-oasis", nsel=1, neig=None):
        self._picker = selection_strategy(self, strategy, nsel, neig)

    def select_columns(self):
        return self._picker.select()

    # ---- growth ----
    def add_column(self, col, icol, update_error=True):
        """Border W^-1 and R with column icol (values col = A[:, icol]); False when
        its Schur complement vanishes (the column is already in the span)."""
        col = np.asarray(col, dtype=np.float64).ravel()
        b = col[self._cols]
        q = self._R[:, icol]
        sigma = float(self._diag[icol] - b.dot(q))
        if np.isclose(sigma, 0):
            return False
        k = self.k
        bR = b.dot(self._R)                     # b^T R, (n,)
        Winv = np.empty((k + 1, k + 1))
        Winv[:k, :k] = self._Winv + np.outer(q, q) / sigma
        Winv[:k, k] = Winv[k, :k] = -q / sigma
        Winv[k, k] = 1.0 / sigma
        R = np.empty((k + 1, self.n))
        R[:k] = self._R + np.outer(q, bR - col) / sigma
        R[k] = (col - bR) / sigma
        self._cols = np.append(self._cols, int(icol))
        R[:, self._cols] = np.eye(k + 1)        # exact on the selected columns
        self._Winv, self._R = Winv, R
        self._C = np.column_stack((self._C, col))
        if update_error:
            self._refresh_error()
        return True

    def add_columns(self, C_k_new, columns_new):
        """Add columns in order; returns the indices actually added."""
        C_k_new = np.asarray(C_k_new).reshape(self.n, -1)
        taken = [c for j, c in enumerate(np.asarray(columns_new).ravel())
                 if self.add_column(C_k_new[:, j], c, update_error=False)]
        self._refresh_error()
        return np.array(taken)

    # ---- approximations ----
    def approximate_matrix(self):
        return self._C.dot(self._R)

    def approximate_column(self, i):
        return self._C.dot(self._R[:, i])

    def approximate_cholesky(self, epsilon=1e-6):
        """F with F F^T = C W^-1 C^T (rank-truncated square root of W^-1)."""
        return self._C.dot(spd_inv_split(self.Wk, epsilon=epsilon))

    def approximate_eig(self, epsilon=1e-6):
        """Eigenpairs of the Nystroem approximation C W^-1 C^T via the small
        factor F: F^T F = Q diag(s) Q^T, eigenvectors F^+T Q, unit length."""
        F = self.approximate_cholesky(epsilon=epsilon)
        s, Q = np.linalg.eigh(F.T.dot(F))
        s, Q = sort_by_norm(s, Q)
        vecs = np.linalg.pinv(F.T).dot(Q)
        norms = np.sqrt(np.sum(vecs * vecs, axis=0))
        live = ~np.all(np.isclose(vecs, 0), axis=0)
        vecs[:, live] /= norms[live]
        return s, vecs


class _Picker(object):
    """Common part of the selection rules: how many columns may still be taken,
    and no selection once the approximation is exact."""

    def __init__(self, oasis_obj, strategy, nsel=1, neig=None):
        self._oasis = oasis_obj
        self._strategy = strategy
        self._nsel = nsel
        self._neig = neig

    @property
    def strategy(self):
        return self._strategy

    @property
    def nsel(self):
        return self._nsel

    def _budget(self):
        left = self._oasis.n - self._oasis.k
        if left == 0:
            warnings.warn("Requested more columns but there are none left. Returning None.")
            return 0
        if self._nsel > left:
            warnings.warn("Requested more columns than are left to select. Returning only %d columns." % left)
            return left
        return self._nsel

    def select(self):
        err = self._oasis.error
        if np.allclose(err, 0):
            return None
        count = self._budget()
        if count == 0:
            return None
        return self._pick(count, err)

    def _pick(self, count, err):
        raise NotImplementedError


class _RandomPicker(_Picker):
    """Uniformly among the columns not selected yet (global numpy RNG, as the reference)."""

    def _pick(self, count, err):
        free = np.setdiff1d(np.arange(self._oasis.n), self._oasis.column_indices)
        return np.sort(np.random.choice(free, count, replace=False)).astype(int)


class _OasisPicker(_Picker):
    """The columns with the largest diagonal error (oASIS)."""

    def _pick(self, count, err):
        return np.argsort(np.abs(err))[::-1][:count]


class _SpectralOasisPicker(_Picker):
    """Spectral oASIS: columns spread over the error-weighted approximate
    eigenvectors.  Row i of Z = e_i * [1, v_1(i), .., v_r(i)] (the constant
    vector first) is column i's point; points are taken greedily, each the
    farthest from the ones already covered.  The covered set starts with the
    origin and the point of column 0, whose distance the reference's search
    counts in every round after the first (nystroem_tica.py:744-756), so that
    the selections are the same."""

    def _pick(self, count, err):
        if count == 1:
            return np.array([np.argmax(np.abs(err))])
        o = self._oasis
        if np.allclose(o.Wk, 0):
            vecs = np.ones((o.k, o.k))
        else:
            vecs = self._with_constant(o.approximate_eig()[1])
        r = vecs.shape[1] if self._neig is None else min(self._neig, vecs.shape[1])
        Z = err[:, None] * vecs[:, :r]
        gap = np.linalg.norm(Z, axis=1)                              # distance to the origin
        chosen = [int(np.argmax(gap))]
        gap = np.minimum(gap, np.linalg.norm(Z - Z[0], axis=1))      # column 0 counts as covered
        for _ in range(1, count):
            gap = np.minimum(gap, np.linalg.norm(Z - Z[chosen[-1]], axis=1))
            chosen.append(int(np.argmax(gap)))
        return np.unique(chosen)

    @staticmethod
    def _with_constant(vecs):
        """Make the first eigenvector exactly constant when it nearly is; otherwise
        prepend a constant one."""
        if np.isclose(np.min(vecs[:, 0]), np.max(vecs[:, 0])):
            vecs[:, 0] = 1
            return vecs
        return np.column_stack((np.ones(vecs.shape[0]), vecs))

    _fix_constant_evec = _with_constant


# the reference's class name of the spectral rule (exported by nystroem_tica)
SelectionStrategySpectralOasis = _SpectralOasisPicker

_PICKERS = {"random": _RandomPicker, "oasis": _OasisPicker, "spectral-oasis": _SpectralOasisPicker}


def selection_strategy(oasis_obj, strategy="spectral-oasis", nsel=1, neig=None):
    """Selection rule by name: 'random', 'oasis' or 'spectral-oasis'."""
    name = strategy.lower()
    if name not in _PICKERS:
        raise ValueError("Selected strategy is unknown: " + str(strategy))
    return _PICKERS[name](oasis_obj, name, nsel=nsel, neig=neig)


class NystroemTICA(TICA):
    """Sparse-sampling TICA (nystroem_tica.py:51-271)."""

    def __init__(self, lag, max_columns, dim=-1, var_cutoff=0.95, epsilon=1e-6, stride=1, skip=0, reversible=True,
                 ncov_max=float("inf"), initial_columns=None, nsel=1, selection_strategy="spectral-oasis",
                 neig=None, precision="fp64"):
        super(NystroemTICA, self).__init__(lag, dim=dim, var_cutoff=var_cutoff, kinetic_map=False,
                                           commute_map=False, epsilon=epsilon, stride=stride, skip=skip,
                                           reversible=reversible, ncov_max=ncov_max, precision=precision)
        self.max_columns = max_columns
        self.initial_columns = initial_columns
        self.nsel = nsel
        self.selection_strategy = selection_strategy
        self.neig = neig
        self._oasis = None
        self.model.diag = None
        self.model.column_indices = None

    @property
    def initial_columns(self):
        return self._initial_columns

    @initial_columns.setter
    def initial_columns(self, value):
        """None, a count, a function of the input dimension, or the indices
        themselves (nystroem_tica.py:169-182)."""
        if value is None:
            value = 1
        if isinstance(value, int):
            count = value

            def value(n_features):
                return np.random.choice(n_features, count, replace=False)
        elif isinstance(value, np.ndarray):
            value = np.asarray(value, dtype=int).ravel()
        elif not isinstance(value, FunctionType):
            raise ValueError("initial_columns has to be one of these types (None, int, function, ndarray),"
                             "but was {}".format(type(value)))
        self._initial_columns = value

    def describe(self):
        try:
            dim = self.dimension()
        except RuntimeError:
            dim = self.dim
        return "[NystroemTICA, lag = %i; max. columns = %i; max. output dim. = %i]" % (self.lag, self.max_columns,
                                                                                       dim)

    def estimate(self, X, **kwargs):
        """One moments pass for all covariance columns, then the selection rounds
        on host copies (nystroem_tica.py:201-245)."""
        for name, value in kwargs.items():
            if name == "chunksize":
                self.chunksize = value
            elif hasattr(self, name):
                setattr(self, name, value)
        dt = to_device(X)
        if self.dim > dt.d:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, dt.d))
        if callable(self.initial_columns):
            self.initial_columns = self.initial_columns(dt.d)
        first = np.asarray(self.initial_columns)
        if first.ndim != 1:
            raise ValueError("initial_columns must be either None, an integer, a list, or a 1-d numpy array.")
        first = first.astype(int)
        covar = LaggedCovariance(c00=True, c0t=True, ctt=False, remove_data_mean=True, reversible=self.reversible,
                                 lag=self.lag, bessel=False, stride=self.stride, skip=self.skip,
                                 ncov_max=self.ncov_max, precision=self.precision)
        covar.estimate(dt)
        C00, C0t = covar.C00_, covar.C0t_
        sampler = oASIS_Nystroem(np.diag(C00).copy(), C00[:, first], first)
        sampler.set_selection_strategy(strategy=self.selection_strategy, nsel=self.nsel, neig=self.neig)
        lagged_cols = [C0t[:, first]]
        while sampler.k < min(self.max_columns, sampler.n):
            pick = sampler.select_columns()
            if pick is None or len(pick) == 0 or np.all(np.isin(pick, sampler.column_indices)):
                log.warning("Iteration ended prematurely: No more columns to select.")
                break
            taken = sampler.add_columns(C00[:, pick], pick)
            if len(taken):
                lagged_cols.append(C0t[:, pick[np.isin(pick, taken)]])
        self._oasis = sampler
        self._covar = covar
        self._data = dt
        self.model.update_model_params(mean=covar.mean, diag=np.diag(C00).copy(), cov=sampler.Ck,
                                       cov_tau=np.column_stack(lagged_cols), column_indices=sampler.column_indices)
        self._diagonalize()
        return self

    def partial_fit(self, X):
        raise NotImplementedError("NystroemTICA has no partial_fit (as in the reference)")

    def _diagonalize(self):
        """TICA on the selected rows and columns only (nystroem_tica.py:247-266)."""
        cols = self.model.column_indices
        try:
            eigenvalues, eigenvectors = eig_corr(self._oasis.Wk, self.model.cov_tau[cols, :], self.epsilon,
                                                 sign_maxelement=True)
        except ZeroRankError:
            raise ZeroRankError("All input features are constant in all time steps. No dimension would be left "
                                "after dimension reduction.")
        power = np.cumsum(np.abs(eigenvalues) ** 2)
        self.model.update_model_params(cumvar=power / power[-1], eigenvalues=eigenvalues, eigenvectors=eigenvectors)
        self._estimated = True

    @property
    def column_indices(self):
        return self.model.column_indices

    def _project_device(self, dt):
        """The eigenvectors are coefficients of the selected features: Y = (X - mean)[:, columns] V."""
        from .. import _native

        t = _native.torch()
        m = self.dimension()
        V = np.zeros((dt.d, m))
        V[np.asarray(self.model.column_indices, dtype=int), :] = np.real(self.eigenvectors[:, :m])
        mean = t.as_tensor(np.ascontiguousarray(self.mean, dtype=np.float64), device=dt.X.device)
        Vd = t.as_tensor(np.ascontiguousarray(V, dtype=np.float64), device=dt.X.device)
        return _native.project(dt.X, mean, Vd, m)
