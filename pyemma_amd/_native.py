"""ctypes binding of libemma_amd.so (declared in include/emma_amd.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every entry point raises instead of falling back to anything else.
Device buffers are torch tensors (plumbing only); pointers and the current
HIP stream are handed to the C ABI.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemma_amd.so")

EMMA_OK = 0
EMMA_ERR_INVALID_ARG = -1
EMMA_MOM_SYMMETRIZE = 1
EMMA_MOM_REMOVE_MEAN = 2
EMMA_MOM_C0T = 4
EMMA_MOM_CTT = 8
EMMA_MOM_WEIGHTED = 16
EMMA_PREC_FP32 = 0
EMMA_PREC_BF16X4 = 1
EMMA_PREC_FP64 = 2
EMMA_PREC_BF16X4_ACC = 3
EMMA_ASSIGN_FP32_PLAIN = 0
EMMA_ASSIGN_BF16X3_RECHECK = 1
EMMA_DIST_INVERSE = 1

_lib = None


class EmmaError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EmmaError("libemma_amd.so is not built (%s); run `python -m pyemma_amd._build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, i64, u32, u64, dbl = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double)
    L.emma_version.restype = ctypes.c_char_p
    L.emma_version.argtypes = []
    L.emma_last_error.restype = ctypes.c_char_p
    L.emma_last_error.argtypes = []
    L.emma_moments_workspace_size.restype = sz
    L.emma_moments_workspace_size.argtypes = [i64, i32, i32, u32, i32]
    L.emma_moments_lagged.restype = ctypes.c_int
    L.emma_moments_lagged.argtypes = [vp, i64, i32, vp, i32, i32, i32, i32, u32, i32, vp, vp, vp, vp, vp, vp,
                                      vp, sz, vp]
    L.emma_moments_lagged_weighted.restype = ctypes.c_int
    L.emma_moments_lagged_weighted.argtypes = [vp, i64, i32, vp, i32, i32, i32, i32, u32, i32, vp, vp, vp, vp, vp,
                                               vp, vp, vp, sz, vp]
    L.emma_row_affine.restype = ctypes.c_int
    L.emma_row_affine.argtypes = [vp, i64, i32, vp, dbl, vp, vp]
    L.emma_assign_workspace_size.restype = sz
    L.emma_assign_workspace_size.argtypes = [i64, i32, i32, i32]
    L.emma_assign.restype = ctypes.c_int
    L.emma_assign.argtypes = [vp, i64, i32, vp, i32, vp, i32, vp, vp, sz, vp]
    L.emma_assign_variants.restype = i32
    L.emma_assign_variants.argtypes = [i32]
    L.emma_assign_select.restype = ctypes.c_int
    L.emma_assign_select.argtypes = [i32, i32, i32]
    L.emma_kmeanspp_workspace_size.restype = sz
    L.emma_kmeanspp_workspace_size.argtypes = [i64, i32, i32]
    L.emma_kmeanspp.restype = ctypes.c_int
    L.emma_kmeanspp.argtypes = [vp, i64, i32, i32, u64, vp, vp, vp, sz, vp]
    L.emma_lloyd_workspace_size.restype = sz
    L.emma_lloyd_workspace_size.argtypes = [i64, i32, i32, i32]
    L.emma_lloyd.restype = ctypes.c_int
    L.emma_lloyd.argtypes = [vp, i64, i32, vp, i32, i32, dbl, i32, vp, vp, vp, vp, vp, sz, vp]
    L.emma_project.restype = ctypes.c_int
    L.emma_project.argtypes = [vp, i64, i32, vp, vp, i32, i32, vp, vp]
    L.emma_count_nonfinite.restype = ctypes.c_int
    L.emma_count_nonfinite.argtypes = [vp, i64, vp, vp]
    L.emma_count_transitions.restype = ctypes.c_int
    L.emma_count_transitions.argtypes = [vp, i64, vp, i32, i32, i32, i32, vp, vp, vp, vp]
    L.emma_pair_distances.restype = ctypes.c_int
    L.emma_pair_distances.argtypes = [vp, i64, i32, vp, i32, u32, vp, vp]
    L.emma_max_state.restype = ctypes.c_int
    L.emma_max_state.argtypes = [vp, i64, vp, vp]
    L.emma_assign_noalloc.restype = ctypes.c_int
    L.emma_assign_noalloc.argtypes = [vp, i64, i32, vp, i32, vp, i32, vp, vp]
    L.emma_probe_mfma_bf16.restype = ctypes.c_int
    L.emma_probe_mfma_bf16.argtypes = [vp, vp, vp, vp, vp]
    _lib = L
    return L


def version():
    return lib().emma_version().decode()


def check(rc, what):
    if rc != EMMA_OK:
        msg = lib().emma_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError("%s: %s" % (what, msg))
        raise EmmaError("%s failed (%d): %s" % (what, rc, msg))


# ----------------------------------------------------------------------
# device plumbing
# ----------------------------------------------------------------------
def torch():
    import torch as _t
    return _t


def device():
    t = torch()
    if not t.cuda.is_available():
        raise EmmaError("no GPU visible: the MI355X path needs a HIP device")
    return t.device("cuda", t.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


_ws_cache = {}


def workspace(nbytes, tag="default"):
    """Reusable device scratch buffer (grown on demand, per tag)."""
    t = torch()
    nbytes = max(int(nbytes), 256)
    buf = _ws_cache.get(tag)
    if buf is None or buf.numel() < nbytes:
        buf = t.empty(nbytes, dtype=t.uint8, device=device())
        _ws_cache[tag] = buf
    return buf


def as_device_f32(X):
    """2-D C-contiguous float32 device tensor from ndarray/tensor."""
    t = torch()
    if isinstance(X, t.Tensor):
        Xt = X.to(device=device(), dtype=t.float32)
    else:
        Xn = np.require(X, dtype=np.float32, requirements="C")
        Xt = t.from_numpy(Xn).to(device())
    if Xt.dim() == 1:
        Xt = Xt.reshape(-1, 1)
    return Xt.contiguous()


# ----------------------------------------------------------------------
# typed wrappers of the C ABI (device tensors in, device tensors out)
# ----------------------------------------------------------------------
def count_nonfinite(Xt):
    t = torch()
    out = t.empty(1, dtype=t.int64, device=Xt.device)
    check(lib().emma_count_nonfinite(ptr(Xt), Xt.numel(), ptr(out), stream_ptr()), "emma_count_nonfinite")
    return out


def moments_lagged(Xt, traj_offsets, lag, stride=1, skip=0, reversible=True, remove_mean=True, c0t=True,
                   ctt=False, precision=EMMA_PREC_BF16X4, weights=None):
    """Unnormalized FP64 moments (w, sx, sy, Mxx, Mxy, Myy) as device tensors.

    weights: optional FP64 device tensor with one frame weight per row of Xt.
    """
    t = torch()
    n_rows, d = Xt.shape
    off = traj_offsets
    if not isinstance(off, t.Tensor):
        off = t.as_tensor(np.asarray(off, dtype=np.int64), device=Xt.device)
    ntraj = off.numel() - 1
    flags = 0
    if reversible and lag > 0:
        flags |= EMMA_MOM_SYMMETRIZE
    if remove_mean:
        flags |= EMMA_MOM_REMOVE_MEAN
    if c0t and lag > 0:
        flags |= EMMA_MOM_C0T
    if ctt and lag > 0:
        flags |= EMMA_MOM_CTT
    if weights is not None:
        flags |= EMMA_MOM_WEIGHTED
        if weights.dtype != t.float64 or weights.numel() != n_rows or weights.device != Xt.device:
            raise ValueError("weights must be a float64 device tensor with one entry per frame")
        weights = weights.contiguous()
    L = lib()
    ws_bytes = L.emma_moments_workspace_size(n_rows, d, ntraj, flags, precision)
    ws = workspace(ws_bytes, "moments")
    f64 = dict(dtype=t.float64, device=Xt.device)
    w = t.empty(1, **f64)
    sx = t.empty(d, **f64)
    sy = t.empty(d, **f64)
    Mxx = t.empty((d, d), **f64)
    Mxy = t.empty((d, d), **f64) if flags & EMMA_MOM_C0T else None
    Myy = t.empty((d, d), **f64) if flags & EMMA_MOM_CTT else None
    check(L.emma_moments_lagged_weighted(ptr(Xt), n_rows, d, ptr(off), ntraj, int(lag), int(stride), int(skip),
                                         flags, int(precision), ptr(weights), ptr(w), ptr(sx), ptr(sy), ptr(Mxx),
                                         ptr(Mxy), ptr(Myy), ptr(ws), ws.numel(), stream_ptr()),
          "emma_moments_lagged")
    return w, sx, sy, Mxx, Mxy, Myy


def assign(Xt, Ct, mode=EMMA_ASSIGN_BF16X3_RECHECK, labels=None, n_rechecked=None):
    t = torch()
    n, d = Xt.shape
    k = Ct.shape[0]
    if Ct.shape[1] != d:
        raise ValueError("cluster centers have wrong dimension. Have dim=%i, but input has %i" % (Ct.shape[1], d))
    if labels is None:
        labels = t.empty(n, dtype=t.int32, device=Xt.device)
    L = lib()
    ws = workspace(L.emma_assign_workspace_size(n, d, k, mode), "assign")
    check(L.emma_assign(ptr(Xt), n, d, ptr(Ct), k, ptr(labels), int(mode), ptr(n_rechecked), ptr(ws), ws.numel(),
                        stream_ptr()), "emma_assign")
    return labels


def kmeanspp(Xt, k, seed):
    t = torch()
    n, d = Xt.shape
    C = t.empty((k, d), dtype=t.float32, device=Xt.device)
    idx = t.empty(k, dtype=t.int64, device=Xt.device)
    L = lib()
    ws = workspace(L.emma_kmeanspp_workspace_size(n, d, k), "kmeanspp")
    check(L.emma_kmeanspp(ptr(Xt), n, d, int(k), int(seed) & (2 ** 64 - 1), ptr(C), ptr(idx), ptr(ws), ws.numel(),
                          stream_ptr()), "emma_kmeanspp")
    return C, idx


def lloyd(Xt, Ct, max_iter, tol, mode=EMMA_ASSIGN_BF16X3_RECHECK):
    """Runs Lloyd in place on Ct; returns (labels, inertias[max_iter], n_iter[1], converged[1])."""
    t = torch()
    n, d = Xt.shape
    k = Ct.shape[0]
    labels = t.empty(n, dtype=t.int32, device=Xt.device)
    inertias = t.zeros(max(int(max_iter), 1), dtype=t.float64, device=Xt.device)
    n_iter = t.zeros(1, dtype=t.int32, device=Xt.device)
    conv = t.zeros(1, dtype=t.int32, device=Xt.device)
    L = lib()
    ws = workspace(L.emma_lloyd_workspace_size(n, d, k, mode), "lloyd")
    check(L.emma_lloyd(ptr(Xt), n, d, ptr(Ct), k, int(max_iter), float(tol), int(mode), ptr(labels), ptr(inertias),
                       ptr(n_iter), ptr(conv), ptr(ws), ws.numel(), stream_ptr()), "emma_lloyd")
    return labels, inertias, n_iter, conv


def project(Xt, mean_dev, V_dev, m):
    """Y = (X - mean) V[:, :m] (FP64 math, float32 out)."""
    t = torch()
    n, d = Xt.shape
    Y = t.empty((n, m), dtype=t.float32, device=Xt.device)
    ldv = V_dev.shape[1]
    check(lib().emma_project(ptr(Xt), n, d, ptr(mean_dev), ptr(V_dev), ldv, int(m), ptr(Y), stream_ptr()),
          "emma_project")
    return Y


def row_affine(Xt, u, u0):
    """out[i] = X[i] . u + u0 in FP64 (device tensor)."""
    t = torch()
    n, d = Xt.shape
    u_dev = t.as_tensor(np.asarray(u, dtype=np.float64).ravel(), device=Xt.device)
    if u_dev.numel() != d:
        raise ValueError("weight vector has %d entries, data has %d columns" % (u_dev.numel(), d))
    out = t.empty(n, dtype=t.float64, device=Xt.device)
    check(lib().emma_row_affine(ptr(Xt), n, d, ptr(u_dev), float(u0), ptr(out), stream_ptr()), "emma_row_affine")
    return out


def max_state(S_dev):
    """Largest entry of an int32 device tensor (-1 when empty), as a device int32 tensor."""
    t = torch()
    out = t.empty(1, dtype=t.int32, device=S_dev.device)
    check(lib().emma_max_state(ptr(S_dev), S_dev.numel(), ptr(out), stream_ptr()), "emma_max_state")
    return out


def count_transitions(S_dev, traj_offsets, lag, sliding, nstates, hist=True):
    """Dense transition counts (nstates x nstates int64), state histogram and the
    number of out-of-range entries, all device tensors."""
    t = torch()
    off = traj_offsets
    if not isinstance(off, t.Tensor):
        off = t.as_tensor(np.asarray(off, dtype=np.int64), device=S_dev.device)
    C = t.empty((nstates, nstates), dtype=t.int64, device=S_dev.device)
    H = t.empty(nstates, dtype=t.int64, device=S_dev.device) if hist else None
    bad = t.empty(1, dtype=t.int64, device=S_dev.device)
    check(lib().emma_count_transitions(ptr(S_dev), S_dev.numel(), ptr(off), off.numel() - 1, int(lag),
                                       1 if sliding else 0, int(nstates), ptr(C), ptr(H), ptr(bad), stream_ptr()),
          "emma_count_transitions")
    return C, H, bad
