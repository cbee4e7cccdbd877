"""User API with the reference's signatures (pyemma/coordinates/api.py).

tica             api.py:1046-1290
covariance_lagged api.py:1554-1661
cluster_kmeans   api.py:1727-1866
assign_to_centers api.py:2060-2156
vamp             api.py:1294-1443
tica_nystroem    api.py:1446-1551
pca              api.py:900-1042
"""
import types
import warnings

import numpy as np

from .clustering import AssignCenters, KmeansClustering
from .covariance import LaggedCovariance
from .nystroem import NystroemTICA
from .pca import PCA
from .tica import TICA
from .vamp import VAMP

__all__ = ["tica", "covariance_lagged", "cluster_kmeans", "assign_to_centers", "vamp", "pca", "tica_nystroem"]


def _weights_arg(weights, data, lag, stride, skip, ncov_max, reversible=True, for_tica=True, precision="fp64"):
    """api.py:1252-1272 (tica) / 1633-1650 (covariance_lagged): resolve the weights argument.

    'koopman' runs the Koopman estimator (a non-reversible moments pass on the
    GPU and a small host eigenproblem) and returns its weight model.
    """
    from .koopman import KoopmanEstimator

    if isinstance(weights, str):
        if weights == "koopman":
            if data is None:
                raise ValueError("Data must be supplied for reweighting='koopman'")
            if for_tica and not reversible:
                raise ValueError("Koopman re-weighting is designed for reversible processes, set reversible=True")
            koop = KoopmanEstimator(lag=lag, stride=stride, skip=skip, ncov_max=ncov_max, precision=precision)
            # the Koopman estimator needs every frame at once: a streamed source is featurized whole
            from .source import FeatureSource
            koop.estimate(data.materialize() if isinstance(data, FeatureSource) else data)
            return koop.weights
        if weights == "empirical":
            return None
        raise ValueError("reweighting must be either 'empirical', 'koopman' or an object with a weights(data) "
                         "method.")
    if hasattr(weights, "weights") and isinstance(getattr(weights, "weights"), types.MethodType):
        return weights
    if isinstance(weights, (list, tuple)) and (not for_tica or all(isinstance(w, np.ndarray) for w in weights)):
        if for_tica and data is not None and len(data) != len(weights):
            raise ValueError("len of weights({}) must match len of data({}).".format(len(weights), len(data)))
        return weights
    if isinstance(weights, np.ndarray) and not for_tica:
        return weights
    raise ValueError("reweighting must be either 'empirical', 'koopman' or an object with a weights(data) method.")


def tica(data=None, lag=10, dim=-1, var_cutoff=0.95, kinetic_map=True, commute_map=False, weights="empirical",
         stride=1, remove_mean=True, skip=0, reversible=True, ncov_max=float("inf"), chunksize=None, **kwargs):
    r"""Time-lagged independent component analysis (api.py:1046-1290)."""
    precision = kwargs.pop("precision", "bf16x4-acc")
    weights = _weights_arg(weights, data, lag, stride, skip, ncov_max, reversible=reversible, for_tica=True,
                           precision=precision)
    if not remove_mean:
        warnings.warn("remove_mean option is deprecated. The mean is removed from the data by default, otherwise "
                      "it cannot be guaranteed that all eigenvalues will be smaller than one.", DeprecationWarning)
    res = TICA(lag, dim=dim, var_cutoff=var_cutoff, kinetic_map=kinetic_map, commute_map=commute_map, skip=skip,
               stride=stride, weights=weights, reversible=reversible, ncov_max=ncov_max, precision=precision)
    if data is not None:
        res.estimate(data, chunksize=chunksize)
    else:
        res.chunksize = chunksize
    return res


def vamp(data=None, lag=10, dim=None, scaling=None, right=False, ncov_max=float("inf"), stride=1, skip=0,
         chunksize=None, **kwargs):
    r"""Variational approach for Markov processes (api.py:1294-1443)."""
    res = VAMP(lag, dim=dim, scaling=scaling, right=right, skip=skip, ncov_max=ncov_max,
               precision=kwargs.pop("precision", "fp64"))
    if data is not None:
        res.estimate(data, stride=stride, chunksize=chunksize)
    else:
        res.chunksize = chunksize
    return res


def tica_nystroem(max_columns, data=None, lag=10, dim=-1, var_cutoff=0.95, epsilon=1e-6, stride=1, skip=0,
                  reversible=True, ncov_max=float("inf"), chunksize=None, initial_columns=None, nsel=1, neig=None,
                  **kwargs):
    r"""Sparse-sampling (Nystroem) TICA (api.py:1446-1551)."""
    res = NystroemTICA(lag, max_columns, dim=dim, var_cutoff=var_cutoff, epsilon=epsilon, skip=skip,
                       reversible=reversible, ncov_max=ncov_max, initial_columns=initial_columns, nsel=nsel,
                       neig=neig, precision=kwargs.pop("precision", "fp64"))
    if data is not None:
        res.estimate(data, stride=stride, chunksize=chunksize)
    else:
        res.chunksize = chunksize
    return res


def pca(data=None, dim=-1, var_cutoff=0.95, stride=1, mean=None, skip=0, chunksize=None, **kwargs):
    r"""Principal component analysis (api.py:900-1042)."""
    if mean is not None:
        warnings.warn("provided mean ignored", DeprecationWarning)
    res = PCA(dim=dim, var_cutoff=var_cutoff, mean=None, skip=skip, stride=stride,
              precision=kwargs.pop("precision", "fp64"))
    if data is not None:
        res.estimate(data, chunksize=chunksize)
    else:
        res.chunksize = chunksize
    return res


def covariance_lagged(data=None, c00=True, c0t=True, ctt=False, remove_constant_mean=None, remove_data_mean=False,
                      reversible=False, bessel=True, lag=0, weights="empirical", stride=1, skip=0, chunksize=None,
                      ncov_max=float("inf"), column_selection=None, diag_only=False, **kwargs):
    r"""Lagged covariances between time series (api.py:1554-1661)."""
    precision = kwargs.pop("precision", "bf16x4-acc")
    weights = _weights_arg(weights, data, lag, stride, skip, ncov_max, for_tica=False, precision=precision)
    lc = LaggedCovariance(c00=c00, c0t=c0t, ctt=ctt, remove_constant_mean=remove_constant_mean,
                          remove_data_mean=remove_data_mean, reversible=reversible, bessel=bessel, lag=lag,
                          weights=weights, stride=stride, skip=skip, ncov_max=ncov_max,
                          column_selection=column_selection, diag_only=diag_only, precision=precision)
    if data is not None:
        lc.estimate(data, chunksize=chunksize)
    else:
        lc.chunksize = chunksize
    return lc


def cluster_kmeans(data=None, k=None, max_iter=10, tolerance=1e-5, stride=1, metric="euclidean",
                   init_strategy="kmeans++", fixed_seed=False, n_jobs=None, chunksize=None, skip=0, keep_data=False,
                   clustercenters=None, **kwargs):
    r"""k-means clustering (api.py:1727-1866)."""
    res = KmeansClustering(n_clusters=k, max_iter=max_iter, metric=metric, tolerance=tolerance,
                           init_strategy=init_strategy, fixed_seed=fixed_seed, n_jobs=n_jobs, skip=skip,
                           keep_data=keep_data, clustercenters=clustercenters, stride=stride,
                           assign_mode=kwargs.pop("assign_mode", "bf16x3"))
    if data is not None:
        res.estimate(data, chunksize=chunksize)
    else:
        res.chunksize = chunksize
    return res


def assign_to_centers(data=None, centers=None, stride=1, return_dtrajs=True, metric="euclidean", n_jobs=None,
                      chunksize=None, skip=0, **kwargs):
    r"""Assign data to the nearest cluster centers (api.py:2060-2156)."""
    if centers is None:
        raise ValueError("You have to provide centers in form of a filename or NumPy array or a reader created by "
                         "source function")
    res = AssignCenters(centers, metric=metric, n_jobs=n_jobs, skip=skip, stride=stride,
                        assign_mode=kwargs.pop("assign_mode", "bf16x3"))
    if data is not None:
        res.estimate(data, chunksize=chunksize)
        if return_dtrajs:
            return res.dtrajs
    else:
        res.chunksize = chunksize
    return res
