"""The rounding of v_mfma_f32_32x32x16_bf16's FP32 accumulation, on the device.

The near-tie bound of the bf16x3 assignment (assign_error_coeff in
pyemma_amd/csrc/assign.hip) charges every addition of the matrix core's FP32
accumulation with at most u = 2^-24 relative error, i.e. round to nearest.
A matrix core that truncated (or rounded toward +-inf) would double that
term.  These cases put one exact bf16 x bf16 product next to an FP32
accumulator value so that round-to-nearest-even, truncation, rounding up or
down, round-half-away and a datapath without a sticky bit all give different
results, and require the round-to-nearest-even one.  A second probe sums 16
equal products into one accumulator and accepts either a single rounding of
the exact sum or rounding after every addition (both satisfy the bound).
"""
import numpy as np
import pytest

from pyemma_amd import _native

pytestmark = pytest.mark.gpu

ULP1 = 2.0 ** -23  # spacing of float32 just above 1


def _bf16_bits(x):
    a = np.asarray(x, dtype=np.float32)
    bits = a.view(np.uint32)
    assert np.all(bits & 0xFFFF == 0), "value not exactly representable in bf16"
    return (bits >> 16).astype(np.uint16)


def _probe(A, B, C):
    t = _native.torch()
    dev = _native.device()
    Ad = t.as_tensor(_bf16_bits(A).view(np.int16), device=dev)
    Bd = t.as_tensor(_bf16_bits(B).view(np.int16), device=dev)
    Cd = t.as_tensor(np.asarray(C, dtype=np.float32), device=dev)
    Dd = t.empty((32, 32), dtype=t.float32, device=dev)
    _native.check(_native.lib().emma_probe_mfma_bf16(_native.ptr(Ad), _native.ptr(Bd), _native.ptr(Cd),
                                                     _native.ptr(Dd), _native.stream_ptr()), "emma_probe_mfma_bf16")
    t.cuda.synchronize()
    return Dd.cpu().numpy()


# (C, a, b) with the product a * b exact in FP32; expected = round-to-nearest-
# even of C + a * b (computed exactly in float64, then rounded once)
CASES = [
    (1.0, 0.75, ULP1),                        # 0.75 ulp up: RNE rounds up, truncation/down do not
    (-1.0, -0.75, ULP1),                      # 0.75 ulp down in magnitude: RNE away from zero, up/trunc not
    (1.0, 0.25, ULP1),                        # 0.25 ulp: RNE stays, rounding up does not
    (-1.0, -0.25, ULP1),                      # 0.25 ulp below -1: RNE stays, rounding down does not
    (1.0, 0.5, ULP1),                         # tie to even (stays at 1); round-half-away goes up
    (1.0 + ULP1, 0.5, ULP1),                  # tie to even (goes up to 1 + 2 ulp)
    (1.0, 1.0 + 2.0 ** -7, 0.5 * ULP1),       # just above the tie: needs the sticky bit
    (1.0, -(1.0 + 2.0 ** -7), 0.25 * ULP1),   # just below 1 - tie: sticky bit on subtraction
    (3.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -7),  # product with 15 significant bits added to 3
    (1.0, -1.0, 1.0 - 2.0 ** -8),             # cancellation: exact result 2^-8
    (2.0 ** 20, 0.75, 2.0 ** -3 * 1.0),       # 0.75 * 2^-3 vs ulp(2^20) = 2^-3
    (-(2.0 ** 20), -1.0 - 2.0 ** -7, 2.0 ** -4),  # just above half an ulp below -2^20
]


def _rne(x):
    return np.float32(x)  # numpy's float64 -> float32 conversion rounds to nearest even


def test_single_product_rounds_to_nearest_even():
    A = np.zeros((32, 16), np.float32)
    B = np.zeros((16, 32), np.float32)
    C = np.zeros((32, 32), np.float32)
    for i, (c, a, b) in enumerate(CASES):
        A[i, 0] = a
        B[0, i] = b
        C[i, i] = c
        assert np.float64(np.float32(c)) == c
    D = _probe(A, B, C)
    exact_prod = np.outer(A[:, 0].astype(np.float64), B[0, :].astype(np.float64))
    for i, (c, a, b) in enumerate(CASES):
        want = _rne(np.float64(c) + np.float64(a) * np.float64(b))
        assert D[i, i] == want, "case %d: C=%r a*b=%r: got %r, round-to-nearest-even %r" % (
            i, c, a * b, float(D[i, i]), float(want))
    # every other element: C = 0 plus one product, exact in FP32
    off = ~np.eye(32, dtype=bool)
    np.testing.assert_array_equal(D[off], exact_prod.astype(np.float32)[off])


def test_sixteen_products_one_accumulator():
    """1 + 16 * (3 * 2^-29): each product is 3/32 ulp of 1.  One rounding of the
    exact sum (0.75 ulp) gives 1 + ulp; rounding after every addition gives 1.
    Both are within the bound; a result of neither kind is not."""
    p = 3.0 * 2.0 ** -29
    A = np.full((32, 16), 3.0 * 2.0 ** -15, np.float32)
    B = np.full((16, 32), 2.0 ** -14, np.float32)
    C = np.ones((32, 32), np.float32)
    D = _probe(A, B, C)
    fused = _rne(1.0 + 16 * p)
    sequential = np.float32(1.0)
    for _ in range(16):
        sequential = np.float32(np.float64(sequential) + p)
    vals = set(np.unique(D).tolist())
    assert vals <= {float(fused), float(sequential)}, vals
    assert len(vals) == 1
    print("16 products into 1.0: %r (%s)" % (float(D[0, 0]), "one rounding of the exact sum"
                                            if D[0, 0] == fused else "rounded after each addition"))
