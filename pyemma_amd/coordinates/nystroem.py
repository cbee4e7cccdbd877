"""Nystroem TICA: column-sampled TICA on one GPU moments pass.

Mirrors pyemma/coordinates/transform/nystroem_tica.py:40-786 (NystroemTICA,
oASIS_Nystroem and its selection strategies) and api.tica_nystroem
(api.py:1446-1551).  The reference streams the data once per selection
round to get the covariance columns of the newly selected features
(LaggedCovariance with column_selection).  Here one pass of the moments
kernels yields every column of C00 and C0t at the cost of one selection
round; the oASIS rounds then slice host copies of those matrices, so the
selected columns, the error vectors and the eigenproblem are the same
numbers the reference builds, without re-reading the frames.
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
    """Nystroem approximation of an spd matrix A from k of its columns (nystroem_tica.py:274-643).

    Keeps C_k = A[:, columns], W_k^-1 with W_k = A[columns, columns] and
    R_k = W_k^-1 C_k^T, so that A ~ C_k R_k; the per-column error of the
    approximated diagonal drives the column selection.
    """

    def __init__(self, d, C_k, columns):
        self._d = np.array(d)
        self._C_k = np.array(C_k)
        self._columns = np.array(columns)
        self._k = self._columns.shape[0]
        self._n = self._C_k.shape[0]
        self.update_inverse()
        self._compute_error()
        self._selection_strategy = selection_strategy(self, "spectral-oasis", nsel=len(columns))

    @property
    def n(self):
        return self._n

    @property
    def k(self):
        return self._k

    @property
    def Ck(self):
        return self._C_k

    @property
    def Wk(self):
        return self._C_k[self._columns, :]

    @property
    def Wk_inv(self):
        return self._W_k_inv

    @property
    def diag(self):
        return self._d

    @property
    def column_indices(self):
        return np.array(self._columns)

    @property
    def error(self):
        return self._err

    def _compute_error(self):
        # err_i = (C_k R_k)_ii - A_ii
        self._err = np.einsum("ij,ji->i", self._C_k, self._R_k) - self._d

    def set_selection_strategy(self, strategy="spectral-oasis", nsel=1, neig=None):
        self._selection_strategy = selection_strategy(self, strategy, nsel, neig)

    def select_columns(self):
        return self._selection_strategy.select()

    def update_inverse(self):
        self._W_k_inv = np.linalg.pinv(self.Wk)
        self._R_k = np.dot(self._W_k_inv, self._C_k.T)

    def add_column(self, col, icol, update_error=True):
        """Bordered-inverse update with the Schur complement of the new column (nystroem_tica.py:466-533);
        returns False (nothing added) when the complement is numerically zero."""
        k = self._k
        b = col[self._columns][:, None]
        q = self._R_k[:, icol][:, None]
        schur = float((self._d[icol] - np.dot(b.T, q)).item())
        if np.isclose(schur, 0):
            return False
        s = 1.0 / schur
        qC = np.dot(b.T, self._R_k)
        Winv = np.zeros((k + 1, k + 1))
        Winv[0:k, 0:k] = self._W_k_inv + s * np.dot(q, q.T)
        Winv[0:k, k] = -s * q[0:k, 0]
        Winv[k, 0:k] = -s * q[0:k, 0].T
        Winv[k, k] = s
        R = np.vstack((self._R_k + s * np.dot(q, (qC - col.T)), s * (-qC + col.T)))
        cols = np.append(self._columns, icol)
        R[:, cols] = np.eye(k + 1)
        self._W_k_inv = Winv
        self._R_k = R
        self._C_k = np.hstack((self._C_k, col[:, None]))
        self._k += 1
        self._columns = cols
        if update_error:
            self._compute_error()
        return True

    def add_columns(self, C_k_new, columns_new):
        added = [c for i, c in enumerate(columns_new) if self.add_column(C_k_new[:, i], c, update_error=False)]
        self._compute_error()
        return np.array(added)

    def approximate_matrix(self):
        return np.dot(self._C_k, self._R_k)

    def approximate_column(self, i):
        return np.dot(self._C_k, self._R_k[:, i])

    def approximate_cholesky(self, epsilon=1e-6):
        L0 = spd_inv_split(self.Wk, epsilon=epsilon)
        return np.dot(self._C_k, L0)

    def approximate_eig(self, epsilon=1e-6):
        L = self.approximate_cholesky(epsilon=epsilon)
        s, V = np.linalg.eigh(np.dot(L.T, L))
        s, V = sort_by_norm(s, V)
        V = np.dot(np.linalg.pinv(L.T), V)
        for i in range(V.shape[1]):
            if not np.allclose(V[:, i], 0):
                V[:, i] /= np.sqrt(np.dot(V[:, i], V[:, i]))
        return s, V


class SelectionStrategy(object):
    """nystroem_tica.py:646-707."""

    def __init__(self, oasis_obj, strategy="spectral-oasis", nsel=1, neig=None):
        self._oasis_obj = oasis_obj
        self._strategy = strategy
        self._nsel = nsel
        self._neig = neig

    @property
    def strategy(self):
        return self._strategy

    @property
    def nsel(self):
        return self._nsel

    def _check_nsel(self):
        o = self._oasis_obj
        if o._n == o._k:
            warnings.warn("Requested more columns but there are none left. Returning None.")
            return None
        if self._nsel > o._n - o._k:
            ncols = o._n - o._k
            warnings.warn("Requested more columns than are left to select. Returning only " + str(ncols)
                          + " columns.")
            return ncols
        return self._nsel

    def select(self):
        err = self._oasis_obj.error
        if np.allclose(err, 0):
            return None
        nsel = self._check_nsel()
        if nsel is None:
            return None
        return self._select(nsel, err)

    def _select(self, nsel, err):
        raise NotImplementedError("Classes derived from SelectionStrategy must override the _select() method.")


class SelectionStrategyRandom(SelectionStrategy):
    def _select(self, nsel, err):
        sel = []
        while len(sel) < nsel:
            i = np.random.choice(self._oasis_obj._n)
            if not (i in sel or i in self._oasis_obj._columns):
                sel.append(i)
        return np.array(sel, dtype=int)


class SelectionStrategyOasis(SelectionStrategy):
    def _select(self, nsel, err):
        return np.argsort(np.abs(err))[::-1][:nsel]


class SelectionStrategySpectralOasis(SelectionStrategy):
    """nystroem_tica.py:727-767: farthest points of the error-weighted approximate eigenvectors."""

    def _select(self, nsel, err):
        if nsel == 1:
            return np.array([np.argmax(np.abs(err))])
        o = self._oasis_obj
        if np.allclose(o.Wk, 0):
            evec = np.ones((o.k, o.k))
        else:
            _, evec = o.approximate_eig()
            evec = self._fix_constant_evec(evec)
        neig = evec.shape[1] if self._neig is None else min(self._neig, evec.shape[1])
        W = err[:, None] * evec[:, :neig]
        sel = np.zeros(nsel, dtype=int)
        d_to_0 = np.linalg.norm(W, axis=1)
        sel[0] = np.argmax(d_to_0)
        for i in range(1, nsel):
            # distance to every entry of sel, the not yet filled ones (index 0) included, as in the reference
            d_to_i = d_to_0
            for k in range(nsel):
                d_to_i = np.minimum(d_to_i, np.linalg.norm(W - W[sel[k], :], axis=1))
            sel[i] = np.argmax(d_to_i)
        return np.unique(sel)

    @staticmethod
    def _fix_constant_evec(evecs):
        evec0 = evecs[:, 0]
        if np.isclose(np.min(evec0), np.max(evec0)):
            evecs[:, 0] = 1
        else:
            evecs = np.hstack([np.ones((evecs.shape[0], 1)), evecs])
        return evecs


def selection_strategy(oasis_obj, strategy="spectral-oasis", nsel=1, neig=None):
    strategy = strategy.lower()
    if strategy == "random":
        return SelectionStrategyRandom(oasis_obj, strategy, nsel=nsel, neig=neig)
    if strategy == "oasis":
        return SelectionStrategyOasis(oasis_obj, strategy, nsel=nsel, neig=neig)
    if strategy == "spectral-oasis":
        return SelectionStrategySpectralOasis(oasis_obj, strategy, nsel=nsel, neig=neig)
    raise ValueError("Selected strategy is unknown: " + str(strategy))


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
    def initial_columns(self, initial_columns):
        """nystroem_tica.py:169-182."""
        if not (initial_columns is None or isinstance(initial_columns, (int, FunctionType, np.ndarray))):
            raise ValueError("initial_columns has to be one of these types (None, int, function, ndarray),"
                             "but was {}".format(type(initial_columns)))
        if initial_columns is None:
            initial_columns = 1
        if isinstance(initial_columns, int):
            i = initial_columns
            initial_columns = lambda N: np.random.choice(N, i, replace=False)  # noqa: E731
        if isinstance(initial_columns, np.ndarray):
            initial_columns = np.asarray(initial_columns, dtype=int).ravel()
        self._initial_columns = initial_columns

    def describe(self):
        try:
            dim = self.dimension()
        except RuntimeError:
            dim = self.dim
        return "[NystroemTICA, lag = %i; max. columns = %i; max. output dim. = %i]" % (self.lag, self.max_columns,
                                                                                       dim)

    def estimate(self, X, **kwargs):
        """nystroem_tica.py:201-245, with all covariance columns from one moments pass."""
        for k, v in kwargs.items():
            if k == "chunksize":
                self.chunksize = v
            elif hasattr(self, k):
                setattr(self, k, v)
        dt = to_device(X)
        indim = dt.d
        if not self.dim <= indim:
            raise RuntimeError("requested more output dimensions (%i) than dimension of input data (%i)"
                               % (self.dim, indim))
        if callable(self.initial_columns):
            self.initial_columns = self.initial_columns(indim)
        if not len(np.array(self.initial_columns).shape) == 1:
            raise ValueError("initial_columns must be either None, an integer, a list, or a 1-d numpy array.")
        covar = LaggedCovariance(c00=True, c0t=True, ctt=False, remove_data_mean=True, reversible=self.reversible,
                                 lag=self.lag, bessel=False, stride=self.stride, skip=self.skip,
                                 ncov_max=self.ncov_max, precision=self.precision)
        covar.estimate(dt)
        C00, C0t = covar.C00_, covar.C0t_
        cols0 = np.asarray(self.initial_columns, dtype=int)
        cov_tau = C0t[:, cols0]
        self._oasis = oASIS_Nystroem(np.diag(C00).copy(), C00[:, cols0], cols0)
        self._oasis.set_selection_strategy(strategy=self.selection_strategy, nsel=self.nsel, neig=self.neig)
        while self._oasis.k < np.min((self.max_columns, self._oasis.n)):
            cols = self._oasis.select_columns()
            if cols is None or len(cols) == 0 or np.all(np.isin(cols, self._oasis.column_indices)):
                log.warning("Iteration ended prematurely: No more columns to select.")
                break
            ix = self._oasis.add_columns(C00[:, cols], cols)
            ix = np.isin(cols, ix)
            if np.any(ix):
                cov_tau = np.concatenate((cov_tau, C0t[:, cols][:, ix]), axis=1)
        self._covar = covar
        self._data = dt
        self.model.update_model_params(mean=covar.mean, diag=np.diag(C00).copy(), cov=self._oasis.Ck,
                                       cov_tau=cov_tau, column_indices=self._oasis.column_indices)
        self._diagonalize()
        return self

    def partial_fit(self, X):
        raise NotImplementedError("NystroemTICA has no partial_fit (as in the reference)")

    def _diagonalize(self):
        """nystroem_tica.py:247-266: the k x k problem on the selected rows and columns."""
        Wktau = self.model.cov_tau[self.model.column_indices, :]
        try:
            eigenvalues, eigenvectors = eig_corr(self._oasis.Wk, Wktau, self.epsilon, sign_maxelement=True)
        except ZeroRankError:
            raise ZeroRankError("All input features are constant in all time steps. No dimension would be left "
                                "after dimension reduction.")
        cumvar = np.cumsum(np.abs(eigenvalues) ** 2)
        cumvar /= cumvar[-1]
        self.model.update_model_params(cumvar=cumvar, eigenvalues=eigenvalues, eigenvectors=eigenvectors)
        self._estimated = True

    @property
    def column_indices(self):
        return self.model.column_indices

    def _project_device(self, dt):
        """The eigenvectors are coefficients of the selected features: Y = (X - mean)[:, columns] V."""
        from .. import _native

        t = _native.torch()
        m = self.dimension()
        cols = np.asarray(self.model.column_indices, dtype=int)
        V = np.zeros((dt.d, m))
        V[cols, :] = np.real(self.eigenvectors[:, :m])
        mean = t.as_tensor(np.ascontiguousarray(self.mean, dtype=np.float64), device=dt.X.device)
        Vd = t.as_tensor(np.ascontiguousarray(V, dtype=np.float64), device=dt.X.device)
        return _native.project(dt.X, mean, Vd, m)
