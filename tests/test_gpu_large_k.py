"""k-means and assignment at the largest configured shape: d = 512 features
(mixture of 64 Gaussians, means N(0, 4 I), diagonal variances log-uniform in
[0.1, 2]) and k = 4096 centers picked by kmeans++ from the data
(coordinates/clustering/kmeans.py:209-267, interface.py:157-167).

Against the oracle, bit for bit: the labels of 8,192 frames, the 4,096
kmeans++ picks on 20,000 frames, and two Lloyd iterations (centers,
inertias, labels) on those frames.
"""
import os

import numpy as np
import pytest

from oracle import native as on
from pyemma_amd import _native
from pyemma_amd.synthetic import gaussian_mixture

pytestmark = pytest.mark.gpu

D, K = 512, 4096
CORES = len(os.sched_getaffinity(0))


@pytest.fixture(scope="module")
def mixture():
    X = gaussian_mixture(20000, D, seed=3, device=_native.device())
    C, idx = _native.kmeanspp(X, K, 17)
    _native.torch().cuda.synchronize()
    return X, C, idx


def test_kmeanspp_picks_k4096_d512(mixture):
    X, C, idx = mixture
    C_ref, idx_ref = on.kmeanspp(X.cpu().numpy(), K, 17, nthreads=CORES)
    np.testing.assert_array_equal(idx.cpu().numpy(), idx_ref)
    np.testing.assert_array_equal(C.cpu().numpy(), C_ref)


def test_assign_k4096_d512_bit_exact(mixture):
    t = _native.torch()
    X, C, _ = mixture
    Xs = X[:8192].contiguous()
    nre = t.zeros(1, dtype=t.int64, device=X.device)
    lab = _native.assign(Xs, C, n_rechecked=nre)
    t.cuda.synchronize()
    ref = on.assign(Xs.cpu().numpy(), C.cpu().numpy(), nthreads=CORES)
    np.testing.assert_array_equal(lab.cpu().numpy(), ref)
    # the centers are frames: each of those frames is its own center (distance 0)
    print("rechecked %d of 8192 frames" % int(nre.item()))


def test_lloyd_two_iterations_k4096_d512(mixture):
    t = _native.torch()
    X, C0, _ = mixture
    C = C0.clone()
    labels, inert, n_iter, conv = _native.lloyd(X, C, 2, 0.0)
    t.cuda.synchronize()
    C_ref, lab_ref, in_ref, conv_ref = on.lloyd(X.cpu().numpy(), C0.cpu().numpy(), 2, 0.0, nthreads=CORES)
    assert int(n_iter.item()) == len(in_ref) == 2
    np.testing.assert_array_equal(inert.cpu().numpy()[:2], in_ref)
    np.testing.assert_array_equal(C.cpu().numpy(), C_ref)
    np.testing.assert_array_equal(labels.cpu().numpy(), lab_ref)
    assert bool(conv.item()) == conv_ref
