"""ctypes binding of the C oracle (oracle/emma_oracle.c).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by (tests/test_oracle_kmeans.py):
  * assign: brute force in FP64 where the margin between the two nearest
    centers exceeds the FP32 rounding bound, the perfect-assignment
    known-answer test of clustering/tests/test_assign.py:48-109, and
    thread-count invariance (test_assign.py:232-239);
  * kmeans++: centers are rows of X, a fixed seed reproduces itself
    (test_kmeans.py:64-102), every pick has positive D^2, and an
    independent pure-Python restatement of the documented pick rule on
    small inputs;
  * Lloyd: the exact-center known answers of test_kmeans.py:154-166 and
    :411-426, non-increasing cost, and an independent NumPy restatement of
    the documented center update.
The D^2 sampling order and RNG, and the Lloyd accumulation order, have no
counterpart in the reference tree (deeptime is not vendored): those rules
are the documented definition used here, parity unpinned against deeptime.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        f64p = ctypes.POINTER(ctypes.c_double)
        L.oracle_assign.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, f32p, ctypes.c_int, i32p,
                                    f32p, ctypes.c_int]
        L.oracle_assign.restype = None
        L.oracle_dist.argtypes = [f32p, f32p, ctypes.c_int]
        L.oracle_dist.restype = ctypes.c_float
        L.oracle_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_uniform.restype = ctypes.c_double
        L.oracle_block_sum.argtypes = [f32p, ctypes.c_int64, ctypes.c_int64]
        L.oracle_block_sum.restype = ctypes.c_double
        L.oracle_blocked_sum.argtypes = [f32p, ctypes.c_int64]
        L.oracle_blocked_sum.restype = ctypes.c_double
        L.oracle_kmpp_pick.argtypes = [f32p, ctypes.c_int64, f64p, ctypes.c_double]
        L.oracle_kmpp_pick.restype = ctypes.c_int64
        L.oracle_kmeanspp.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_uint64, f32p, i64p, ctypes.c_int]
        L.oracle_kmeanspp.restype = ctypes.c_int
        L.oracle_update_centers.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, i32p, f32p,
                                            ctypes.c_int, i64p]
        L.oracle_update_centers.restype = ctypes.c_int
        L.oracle_lloyd.argtypes = [f32p, ctypes.c_int64, ctypes.c_int, f32p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, i32p, f64p,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.oracle_lloyd.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _f32(X):
    X = np.require(X, dtype=np.float32, requirements="C")
    if X.ndim == 1:
        X = X[:, None]
    return X


def assign(X, C, nthreads=1, return_dists=False):
    """Labels of the sequential FP32 loop (first minimum wins)."""
    X = _f32(X)
    C = _f32(C)
    n, d = X.shape
    labels = np.empty(n, dtype=np.int32)
    dists = np.empty(n, dtype=np.float32)
    lib().oracle_assign(_p(X, ctypes.c_float), n, d, _p(C, ctypes.c_float), C.shape[0],
                        _p(labels, ctypes.c_int32), _p(dists, ctypes.c_float), int(nthreads))
    return (labels, dists) if return_dists else labels


def dist(x, c):
    x = np.require(x, dtype=np.float32, requirements="C").ravel()
    c = np.require(c, dtype=np.float32, requirements="C").ravel()
    return float(lib().oracle_dist(_p(x, ctypes.c_float), _p(c, ctypes.c_float), x.size))


def uniform(seed, counter):
    return float(lib().oracle_uniform(int(seed) & (2**64 - 1), int(counter)))


def block_sum(v, lo, hi):
    v = np.require(v, dtype=np.float32, requirements="C")
    return float(lib().oracle_block_sum(_p(v, ctypes.c_float), int(lo), int(hi)))


def blocked_sum(v):
    v = np.require(v, dtype=np.float32, requirements="C")
    return float(lib().oracle_blocked_sum(_p(v, ctypes.c_float), v.size))


def kmpp_pick(d2, bsum, u):
    d2 = np.require(d2, dtype=np.float32, requirements="C")
    bsum = np.require(bsum, dtype=np.float64, requirements="C")
    return int(lib().oracle_kmpp_pick(_p(d2, ctypes.c_float), d2.size, _p(bsum, ctypes.c_double),
                                      float(u)))


def kmeanspp(X, k, seed, nthreads=1):
    """(centers (k,d) float32, picked row indices (k,) int64)."""
    X = _f32(X)
    n, d = X.shape
    C = np.empty((k, d), dtype=np.float32)
    idx = np.empty(k, dtype=np.int64)
    rc = lib().oracle_kmeanspp(_p(X, ctypes.c_float), n, d, int(k), int(seed) & (2**64 - 1),
                               _p(C, ctypes.c_float), _p(idx, ctypes.c_int64), int(nthreads))
    if rc != 0:
        raise RuntimeError("oracle_kmeanspp failed")
    return C, idx


def update_centers(X, labels, C):
    X = _f32(X)
    C = np.array(_f32(C), copy=True)
    labels = np.require(labels, dtype=np.int32, requirements="C")
    counts = np.empty(C.shape[0], dtype=np.int64)
    rc = lib().oracle_update_centers(_p(X, ctypes.c_float), X.shape[0], X.shape[1],
                                     _p(labels, ctypes.c_int32), _p(C, ctypes.c_float), C.shape[0],
                                     _p(counts, ctypes.c_int64))
    if rc != 0:
        raise RuntimeError("oracle_update_centers failed")
    return C, counts


def lloyd(X, C0, max_iter, tol, nthreads=1):
    """(centers, labels of last assignment, inertias, converged)."""
    X = _f32(X)
    C = np.array(_f32(C0), copy=True)
    n, d = X.shape
    labels = np.zeros(n, dtype=np.int32)
    inertias = np.zeros(max(int(max_iter), 1), dtype=np.float64)
    conv = ctypes.c_int(0)
    it = lib().oracle_lloyd(_p(X, ctypes.c_float), n, d, _p(C, ctypes.c_float), C.shape[0],
                            int(max_iter), float(tol), _p(labels, ctypes.c_int32),
                            _p(inertias, ctypes.c_double), ctypes.byref(conv), int(nthreads))
    if it < 0:
        raise RuntimeError("oracle_lloyd failed")
    return C, labels, inertias[:it], bool(conv.value)
