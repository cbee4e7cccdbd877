"""GPU assignment through the C ABI vs the sequential-FP32 oracle: bit-exact labels."""
import numpy as np
import pytest

from oracle import native as on
from pyemma_amd import _native

pytestmark = pytest.mark.gpu

MODES = [_native.EMMA_ASSIGN_FP32_PLAIN, _native.EMMA_ASSIGN_BF16X3_RECHECK]


def _run(X, C, mode):
    t = _native.torch()
    Xt = _native.as_device_f32(X)
    Ct = _native.as_device_f32(C)
    nre = t.zeros(1, dtype=t.int64, device=Xt.device)
    lab = _native.assign(Xt, Ct, mode=mode, n_rechecked=nre)
    t.cuda.synchronize()
    return lab.cpu().numpy(), int(nre.item())


def _data(n, d, k, seed, offset=0.0, scale=1.0):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, d)) * scale + offset).astype(np.float32)
    C = X[rng.choice(n, k, replace=n < k)] + (rng.standard_normal((k, d)) * 0.3 * scale).astype(np.float32)
    return X, C.astype(np.float32)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,d,k", [(1000, 30, 100), (2049, 300, 257), (777, 3, 1), (1500, 1, 5),
                                   (3000, 128, 512), (1024, 512, 64), (513, 17, 33), (300, 64, 1000)])
def test_labels_bit_exact(mode, n, d, k):
    X, C = _data(n, d, k, seed=n + d + k)
    ref = on.assign(X, C, nthreads=4)
    lab, nre = _run(X, C, mode)
    np.testing.assert_array_equal(lab, ref)
    if mode == _native.EMMA_ASSIGN_FP32_PLAIN:
        assert nre == 0


@pytest.mark.parametrize("mode", MODES)
def test_far_from_origin_and_large_spread(mode):
    """Shifted data (mean 1e3) and mixed scales: the shift and the bound must hold."""
    X, C = _data(4000, 40, 80, seed=5, offset=1000.0, scale=0.5)
    np.testing.assert_array_equal(_run(X, C, mode)[0], on.assign(X, C, nthreads=4))


@pytest.mark.parametrize("mode", MODES)
def test_exact_ties_first_index_wins(mode):
    rng = np.random.default_rng(1)
    C = rng.standard_normal((10, 8)).astype(np.float32)
    C = np.concatenate([C, C[::-1]], axis=0)  # every center twice
    X = np.concatenate([C, rng.standard_normal((500, 8)).astype(np.float32)], axis=0)
    ref = on.assign(X, C)
    lab, nre = _run(X, C, mode)
    np.testing.assert_array_equal(lab, ref)
    if mode == _native.EMMA_ASSIGN_BF16X3_RECHECK:
        assert nre >= 20  # duplicated centers always tie


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("bad_center", [None, 0, 3])
def test_non_finite_rows_and_centers(mode, bad_center):
    """NaN and inf in frames (and NaN in one center): the reference loop starts
    from center 0 and replaces it only on a smaller distance, so a NaN
    distance never wins and a NaN distance to center 0 keeps label 0."""
    X, C = _data(700, 20, 40, seed=11)
    X[5, 3] = np.nan
    X[17, :] = np.inf
    X[18, 0] = -np.inf
    X[40:44, 7] = 3e19  # squares overflow to inf for every center
    if bad_center is not None:
        C[bad_center, 2] = np.nan
    ref = on.assign(X, C)
    np.testing.assert_array_equal(_run(X, C, mode)[0], ref)


@pytest.mark.parametrize("mode", MODES)
def test_listed_candidate_after_a_tie(mode):
    """Exact ties where the later center of a tied pair is also the nearer in
    the MFMA pass's rounding: the first minimum in index order must win."""
    rng = np.random.default_rng(2)
    base = rng.standard_normal((16, 64)).astype(np.float32)
    C = np.concatenate([base, base], axis=0)  # center j and j + 16 identical
    X = (base[rng.integers(0, 16, 3000)] + rng.standard_normal((3000, 64)).astype(np.float32) * 0.05).astype(np.float32)
    ref = on.assign(X, C)
    lab, nre = _run(X, C, mode)
    np.testing.assert_array_equal(lab, ref)
    assert np.all(lab < 16)


def test_golden_hmm_labels(golden_hmm):
    g = golden_hmm
    for mode in MODES:
        np.testing.assert_array_equal(_run(g["X"], g["lloyd_centers"], mode)[0], g["labels"])


def test_recheck_rate_is_small_on_separated_data():
    from pyemma_amd.synthetic import hmm_trajectory
    X = hmm_trajectory(20000, 30, seed=3)
    C = X[::200][:100].copy()
    lab, nre = _run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    np.testing.assert_array_equal(lab, on.assign(X, C, nthreads=4))
    assert nre < 0.2 * len(X)


def test_mfma_dimension_limit_and_plain_beyond():
    X, C = _data(300, 600, 7, seed=2)
    with pytest.raises(ValueError):
        _run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    np.testing.assert_array_equal(_run(X, C, _native.EMMA_ASSIGN_FP32_PLAIN)[0], on.assign(X, C))


def test_deterministic_repeat():
    X, C = _data(5000, 300, 300, seed=9)
    a, na = _run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    b, nb = _run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    np.testing.assert_array_equal(a, b)
    assert na == nb


@pytest.mark.parametrize("d,k", [(20, 70), (64, 256), (128, 200), (128, 300), (200, 100), (300, 1000), (512, 64)])
def test_every_tile_variant_bit_exact(d, k):
    """Each tile variant (waves per workgroup, workgroups per CU, frame blocks per wave) gives the
    oracle's labels, including a ragged last workgroup."""
    L = _native.lib()
    n = 4 * 1024 + 77
    X, C = _data(n, d, k, seed=d * 7 + k)
    ref = on.assign(X, C, nthreads=4)
    try:
        for v in range(L.emma_assign_variants(d)):
            _native.check(L.emma_assign_select(d, k, v), "emma_assign_select")
            lab, _ = _run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)
            np.testing.assert_array_equal(lab, ref, err_msg="variant %d" % v)
    finally:
        _native.check(L.emma_assign_select(d, k, 0), "emma_assign_select")


@pytest.mark.parametrize("d", [64, 128, 300, 301, 510, 512])
@pytest.mark.parametrize("aligned", [True, False])
def test_frame_load_forms(d, aligned):
    """Both frame-load forms of the MFMA passes give the oracle's labels in every tile
    variant: 16-byte loads (d % 4 == 0 and 16-byte aligned frames; in groups of
    K-steps beyond 8 K-steps) and the dword loads otherwise (d % 4 != 0, or frames
    starting 4 bytes past a 16-byte boundary)."""
    t = _native.torch()
    L = _native.lib()
    n, k = 3 * 1024 + 45, 160
    X, C = _data(n, d, k, seed=d + (0 if aligned else 1))
    ref = on.assign(X, C, nthreads=4)
    if aligned:
        Xt = _native.as_device_f32(X)
    else:
        buf = t.empty(n * d + 1, dtype=t.float32, device=_native.device())
        Xt = buf[1:].view(n, d)
        Xt.copy_(t.from_numpy(X))
        assert Xt.data_ptr() % 16 == 4
    Ct = _native.as_device_f32(C)
    try:
        for v in range(L.emma_assign_variants(d)):
            _native.check(L.emma_assign_select(d, k, v), "emma_assign_select")
            lab = _native.assign(Xt, Ct, mode=_native.EMMA_ASSIGN_BF16X3_RECHECK)
            t.cuda.synchronize()
            np.testing.assert_array_equal(lab.cpu().numpy(), ref, err_msg="variant %d" % v)
    finally:
        _native.check(L.emma_assign_select(d, k, 0), "emma_assign_select")


@pytest.mark.parametrize("reps", [2, 8, 9, 12])
def test_recheck_candidate_paths(reps):
    """Every center repeated `reps` times: each frame ties between `reps`
    centers (every frame is listed).  Up to 8 candidates the tie is settled by
    the reference loop over the recorded candidates, above 8 by the full loop
    over every center.  The first minimum wins in both."""
    rng = np.random.default_rng(reps)
    base = rng.standard_normal((6, 24)).astype(np.float32)
    C = np.repeat(base, reps, axis=0)
    X = (base[rng.integers(0, 6, 3000)] + 0.3 * rng.standard_normal((3000, 24))).astype(np.float32)
    ref = on.assign(X, C, nthreads=4)
    lab, nre = _run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)
    np.testing.assert_array_equal(lab, ref)
    assert nre == 3000
    assert np.all(ref % reps == 0)


def _every_variant(X, C):
    """Labels of X under every variant (both first-pass pipelines, every tile shape)."""
    L = _native.lib()
    d, k = X.shape[1], C.shape[0]
    out = []
    try:
        for v in range(L.emma_assign_variants(d)):
            _native.check(L.emma_assign_select(d, k, v), "emma_assign_select")
            out.append(_run(X, C, _native.EMMA_ASSIGN_BF16X3_RECHECK)[0])
    finally:
        _native.check(L.emma_assign_select(d, k, -1), "emma_assign_select")
    return out


@pytest.mark.parametrize("d", [64, 300])
@pytest.mark.parametrize("case", ["outliers", "tiny", "huge", "mixed_scale"])
def test_coarse_pass_edge_cases(d, case):
    """The coarse pass scales frames and centers by 2^t (max |b| near 2^9) into
    fp16.  Outlier frames beyond 64 max|b| overflow fp16 and must reach the fine
    pass; tiny (1e-30) and huge (1e30) data clamp the exponent t to +-60; frames
    with some dimensions 1e-6 of the others underflow fp16 there.  Every
    variant gives the oracle's labels."""
    rng = np.random.default_rng({"outliers": 1, "tiny": 2, "huge": 3, "mixed_scale": 4}[case] + d)
    X, C = _data(3000, d, 40, seed=d + 1)
    if case == "outliers":
        X[::7] *= 5000.0
    elif case == "tiny":
        X, C = X * np.float32(1e-30), C * np.float32(1e-30)
    elif case == "huge":
        X, C = X * np.float32(1e30), C * np.float32(1e30)
    else:
        scale = np.where(rng.random(d) < 0.3, 1e-6, 1.0).astype(np.float32)
        X, C = X * scale, C * scale
    X, C = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(C, np.float32)
    ref = on.assign(X, C, nthreads=4)
    for v, lab in enumerate(_every_variant(X, C)):
        np.testing.assert_array_equal(lab, ref, err_msg="variant %d" % v)
