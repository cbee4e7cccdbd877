"""GPU kmeans++ and Lloyd through the C ABI vs the C oracle (identical picks,
bit-identical centers, inertias and labels)."""
import numpy as np
import pytest

from oracle import native as on
from pyemma_amd import _native

pytestmark = pytest.mark.gpu


def _kmpp(X, k, seed):
    t = _native.torch()
    C, idx = _native.kmeanspp(_native.as_device_f32(X), k, seed)
    t.cuda.synchronize()
    return C.cpu().numpy(), idx.cpu().numpy()


def _lloyd(X, C0, max_iter, tol, mode):
    t = _native.torch()
    Xt = _native.as_device_f32(X)
    C = _native.as_device_f32(C0).clone()
    labels, inert, n_iter, conv = _native.lloyd(Xt, C, max_iter, tol, mode=mode)
    t.cuda.synchronize()
    it = int(n_iter.item())
    return C.cpu().numpy(), labels.cpu().numpy(), inert.cpu().numpy()[:it], bool(conv.item())


@pytest.mark.parametrize("n,d,k,seed", [(9000, 3, 6, 1), (20000, 30, 50, 42), (5000, 300, 20, 7),
                                        (4097, 2, 300, 3), (100, 4, 100, 9)])
def test_kmeanspp_identical_picks(n, d, k, seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, d)) * rng.uniform(0.5, 2, d)).astype(np.float32)
    C_ref, idx_ref = on.kmeanspp(X, k, seed, nthreads=4)
    C, idx = _kmpp(X, k, seed)
    np.testing.assert_array_equal(idx, idx_ref)
    np.testing.assert_array_equal(C, C_ref)


def test_kmeanspp_duplicate_points_total_zero():
    X = np.repeat(np.arange(3, dtype=np.float32)[:, None], 50, axis=0)
    C_ref, idx_ref = on.kmeanspp(X, 5, 11)
    C, idx = _kmpp(X, 5, 11)
    np.testing.assert_array_equal(idx, idx_ref)


@pytest.mark.parametrize("mode", [_native.EMMA_ASSIGN_FP32_PLAIN, _native.EMMA_ASSIGN_BF16X3_RECHECK])
@pytest.mark.parametrize("n,d,k", [(6000, 5, 12), (9000, 30, 100), (3000, 300, 40)])
def test_lloyd_bit_identical(mode, n, d, k):
    rng = np.random.default_rng(n + k)
    X = (rng.standard_normal((n, d)) + rng.integers(0, 4, (n, 1)) * 3.0).astype(np.float32)
    C0, _ = on.kmeanspp(X, k, 5)
    C_ref, lab_ref, in_ref, conv_ref = on.lloyd(X, C0, 8, 0.0, nthreads=4)
    C, lab, inert, conv = _lloyd(X, C0, 8, 0.0, mode)
    np.testing.assert_array_equal(inert, in_ref)
    np.testing.assert_array_equal(C, C_ref)
    np.testing.assert_array_equal(lab, lab_ref)
    assert conv == conv_ref


def test_lloyd_converges_and_stops(golden_hmm):
    g = golden_hmm
    C, lab, inert, conv = _lloyd(g["X"], g["kmpp_centers"], 5, 0.0, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    np.testing.assert_array_equal(C, g["lloyd_centers"])
    np.testing.assert_array_equal(inert, g["lloyd_inertias"])
    X = np.zeros((40000, 4), dtype=np.float32)
    for i, v in enumerate([30.0, 60.0, 90.0, 120.0]):
        X[i * 10000:(i + 1) * 10000] = v
    C0, _ = on.kmeanspp(X, 4, 42)
    C, lab, inert, conv = _lloyd(X, C0, 10, 1e-5, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    assert sorted(C[:, 0].tolist()) == [30.0, 60.0, 90.0, 120.0]
    assert conv and len(inert) < 10
