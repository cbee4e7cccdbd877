"""emma_assign_noalloc through the C ABI against emma_assign (the workspace
entry point): identical labels and recheck counts, on two calls in a row
with the second one smaller (the shared scratch keeps the first call's
contents), for sizes that fit one chunk and for sizes that run in several
chunks, plus the host-side rejection of a center set beyond the scratch."""
import ctypes

import numpy as np
import pytest

from pyemma_amd import _native

pytestmark = pytest.mark.gpu


def _with_workspace(X, C, mode):
    t = _native.torch()
    nre = t.zeros(1, dtype=t.int64, device=X.device)
    lab = _native.assign(X, C, mode=mode, n_rechecked=nre)
    t.cuda.synchronize()
    return lab.cpu().numpy(), int(nre.item())


def _noalloc(X, C, mode):
    t = _native.torch()
    n, d = X.shape
    lab = t.full((n,), -7, dtype=t.int32, device=X.device)
    nre = t.full((1,), -7, dtype=t.int64, device=X.device)
    rc = _native.lib().emma_assign_noalloc(_native.ptr(X), n, d, _native.ptr(C), C.shape[0], _native.ptr(lab),
                                           int(mode), _native.ptr(nre), _native.stream_ptr())
    _native.check(rc, "emma_assign_noalloc")
    t.cuda.synchronize()
    return lab.cpu().numpy(), int(nre.item())


def _data(n, d, k, seed):
    t = _native.torch()
    g = t.Generator(device=_native.device())
    g.manual_seed(seed)
    means = t.randn((16, d), generator=g, device=_native.device()) * 3
    comp = t.randint(0, 16, (n,), generator=g, device=_native.device())
    X = (means[comp] + t.randn((n, d), generator=g, device=_native.device())).contiguous()
    rows = t.randperm(n, generator=g, device=_native.device())[:k]
    C = (X[rows] + 0.2 * t.randn((k, d), generator=g, device=_native.device())).contiguous()
    return X, C


@pytest.mark.parametrize("mode", [_native.EMMA_ASSIGN_BF16X3_RECHECK, _native.EMMA_ASSIGN_FP32_PLAIN])
@pytest.mark.parametrize("d,k,n1,n2", [(30, 100, 50_000, 20_000),       # one chunk each
                                       (300, 1000, 70_000, 9_000),      # one chunk each
                                       (512, 8_000, 250_000, 140_000)])  # chunks of ~130k frames
def test_noalloc_matches_workspace_entry(mode, d, k, n1, n2):
    if mode == _native.EMMA_ASSIGN_FP32_PLAIN and k > 1000:
        pytest.skip("the plain loop at k = 8,000 takes long and has no scratch")
    X, C = _data(n1, d, k, seed=d + k)
    for n in (n1, n2):  # second call smaller: the scratch holds the first call's contents
        Xs = X[:n].contiguous()
        ref_lab, ref_nre = _with_workspace(Xs, C, mode)
        lab, nre = _noalloc(Xs, C, mode)
        np.testing.assert_array_equal(lab, ref_lab)
        assert nre == ref_nre, (n, nre, ref_nre)


def test_noalloc_rejects_centers_beyond_scratch():
    t = _native.torch()
    d, k = 512, 20_000
    X = t.zeros((10, d), dtype=t.float32, device=_native.device())
    C = t.zeros((k, d), dtype=t.float32, device=_native.device())
    lab = t.empty(10, dtype=t.int32, device=X.device)
    rc = _native.lib().emma_assign_noalloc(_native.ptr(X), 10, d, _native.ptr(C), k, _native.ptr(lab),
                                           _native.EMMA_ASSIGN_BF16X3_RECHECK, ctypes.c_void_p(0),
                                           _native.stream_ptr())
    assert rc == _native.EMMA_ERR_INVALID_ARG
    assert b"scratch" in _native.lib().emma_last_error()
