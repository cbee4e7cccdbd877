"""TICA projection on the GPU (emma_project, the transform of _tica_base.py:98-133):
Y = (X - mean) V[:, :m] in FP64, rounded to float32, against numpy's FP64 product
(rounding of the FP64 results may differ by the summation order: within 2 float32
ulp plus 1e-12 of the row scale), with ragged sizes, more outputs than one pass
(m > 32), ldv > m and a partial last workgroup."""
import numpy as np
import pytest

from pyemma_amd import _native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,m,ldv", [(1000, 300, 32, 40), (777, 30, 5, 5), (4097, 3, 2, 3),
                                       (50, 512, 70, 70), (1, 1, 1, 1), (300, 33, 33, 64)])
def test_projection_vs_fp64(n, d, m, ldv):
    t = _native.torch()
    rng = np.random.default_rng(n + d)
    X = (rng.standard_normal((n, d)) * 2 + 1).astype(np.float32)
    mean = rng.standard_normal(d)
    V = rng.standard_normal((d, ldv))
    dev = _native.device()
    Y = _native.project(_native.as_device_f32(X), t.as_tensor(mean, device=dev), t.as_tensor(V, device=dev), m)
    t.cuda.synchronize()
    got = Y.cpu().numpy()
    ref = (X.astype(np.float64) - mean) @ V[:, :m]
    assert got.shape == (n, m) and got.dtype == np.float32
    scale = np.abs(X.astype(np.float64) - mean) @ np.abs(V[:, :m])
    tol = 2 * np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-12 * scale
    assert np.all(np.abs(got.astype(np.float64) - ref) <= tol)


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 17, 24, 25, 31, 40, 57, 64])
def test_projection_output_blocks(m):
    """Every output tiling of the kernel: one 16-output tile (m <= 16), two
    (m <= 32), passes of 32 (m > 32), each with a partial last tile; 700
    frames (a partial last workgroup) of 45 dims (a partial last dim chunk),
    ldv = m + 3."""
    test_projection_vs_fp64(700, 45, m, m + 3)


def test_projection_empty():
    t = _native.torch()
    dev = _native.device()
    X = t.zeros((0, 4), dtype=t.float32, device=dev)
    Y = _native.project(X, t.zeros(4, dtype=t.float64, device=dev), t.zeros((4, 2), dtype=t.float64, device=dev), 2)
    assert tuple(Y.shape) == (0, 2)
