"""Pins the C oracle of the k-means half (oracle/emma_oracle.c).

Independent restatements (NumPy / pure Python) of each documented rule,
brute force in FP64 where margins are large, and the reference's known-answer
tests (clustering/tests/test_assign.py:48-109, test_kmeans.py:154-166,
:411-426, :64-102).
"""
import math

import numpy as np
import pytest

from oracle import native as on


def _labels_numpy_seq(X, C):
    """Sequential FP32 loop in NumPy: every op rounds to float32, no FMA."""
    X = X.astype(np.float32)
    C = C.astype(np.float32)
    acc = np.zeros((X.shape[0], C.shape[0]), dtype=np.float32)
    for i in range(X.shape[1]):
        t = (X[:, None, i] - C[None, :, i]).astype(np.float32)
        acc = (acc + (t * t).astype(np.float32)).astype(np.float32)
    return np.argmin(acc, axis=1).astype(np.int32), acc


@pytest.mark.parametrize("n,d,k", [(500, 3, 7), (300, 37, 50), (200, 130, 33)])
def test_assign_equals_numpy_sequential_loop(n, d, k):
    rng = np.random.default_rng(n + d + k)
    X = (rng.standard_normal((n, d)) * 3 + 10).astype(np.float32)
    C = X[rng.choice(n, k, replace=False)] + rng.standard_normal((k, d)).astype(np.float32) * 0.1
    lab, dist = on.assign(X, C, return_dists=True)
    ref, acc = _labels_numpy_seq(X, C)
    np.testing.assert_array_equal(lab, ref)
    np.testing.assert_array_equal(dist, acc[np.arange(n), ref])


def test_assign_first_minimum_wins():
    X = np.array([[1.0, 0.0], [0.0, 0.0], [1.0, 1.0]], dtype=np.float32)
    C = np.array([[2.0, 0.0], [0.0, 1.0], [1.0, 0.0]], dtype=np.float32)
    # [0,0] and [1,1] tie between centers 1 and 2 (distance 1): the lower index wins
    assert on.assign(X, C).tolist() == [2, 1, 1]


def test_assign_fp64_bruteforce_where_margin_is_large():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2000, 16)).astype(np.float32)
    C = rng.standard_normal((40, 16)).astype(np.float32)
    lab = on.assign(X, C)
    D = ((X.astype(np.float64)[:, None, :] - C.astype(np.float64)[None]) ** 2).sum(-1)
    srt = np.sort(D, axis=1)
    gamma = 18 * 2.0 ** -24 / (1 - 18 * 2.0 ** -24)
    safe = (srt[:, 1] - srt[:, 0]) > 4 * gamma * srt[:, 1]
    assert safe.mean() > 0.99
    np.testing.assert_array_equal(lab[safe], np.argmin(D, axis=1)[safe])


def test_assign_perfect_known_answer():
    """test_assign.py:48-109: well separated Gaussian blobs, centers at the means."""
    rng = np.random.default_rng(0)
    means = np.array([[-3, 0], [-1, 1], [0, 0], [1, -1], [4, 2]], dtype=float)
    ns = 1000
    X = np.zeros((5 * ns, 2))
    for i in range(5):
        X[i * ns:(i + 1) * ns] = 0.1 * rng.standard_normal() + means[i]
    lab = on.assign(X.astype(np.float32), means.astype(np.float32))
    np.testing.assert_array_equal(lab, np.repeat(np.arange(5), ns))


def test_assign_thread_invariance():
    """test_assign.py:232-239."""
    rng = np.random.default_rng(9)
    X = rng.random((5000, 2)).astype(np.float32)
    C = X[rng.choice(5000, 100)]
    np.testing.assert_array_equal(on.assign(X, C, nthreads=1), on.assign(X, C, nthreads=4))


# ---------------------------------------------------------------- kmeans++
def _uniform_py(seed, counter):
    M = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    z = mix((seed + (counter + 1) * 0x9E3779B97F4A7C15) & M)
    return (z >> 11) * 2.0 ** -53


def _block_sum_py(v, lo, hi):
    s = [0.0] * 64
    for l in range(64):
        a = 0.0
        for i in range(lo + l, hi, 64):
            a = a + float(v[i])
        s[l] = a
    for off in (32, 16, 8, 4, 2, 1):
        s = [s[l] + s[l ^ off] for l in range(64)]
    return s[0]


def _pick_py(d2, u):
    """The documented pick rule, restated in plain Python."""
    n = len(d2)
    B, S, G = 4096, 64, 64
    nb = (n + B - 1) // B
    bsum = [_block_sum_py(d2, b * B, min((b + 1) * B, n)) for b in range(nb)]
    nsb = (nb + S - 1) // S
    sbs = []
    for s in range(nsb):
        a = 0.0
        for b in range(s * S, min((s + 1) * S, nb)):
            a = a + bsum[b]
        sbs.append(a)
    total = 0.0
    for a in sbs:
        total = total + a
    if not total > 0:
        return min(int(u * n), n - 1)
    T = u * total

    def walk(items, start):
        run, last, last_before = start, None, 0.0
        for key, val in items:
            prev = run
            run = run + val
            if val > 0:
                last, last_before = key, prev
            if run > T:
                return key, prev
        return last, last_before
    s, before = walk([(s, sbs[s]) for s in range(nsb)], 0.0)
    b, before = walk([(b, bsum[b]) for b in range(s * S, min((s + 1) * S, nb))], before)
    lo, hi = b * B, min((b + 1) * B, n)
    segs = []
    for g in range(lo, hi, G):
        a = 0.0
        for i in range(g, min(g + G, hi)):
            a = a + float(d2[i])
        segs.append((g, a))
    g, before = walk(segs, before)
    return walk([(i, float(d2[i])) for i in range(g, min(g + G, hi))], before)[0]


def test_uniform_stream_matches_python():
    for seed in (0, 42, 2 ** 32 - 1, 123456789):
        for c in (0, 1, 7, 999):
            assert on.uniform(seed, c) == _uniform_py(seed, c)


def test_block_sum_matches_python():
    rng = np.random.default_rng(4)
    v = rng.random(5000).astype(np.float32) * 1e3
    assert on.block_sum(v, 0, 4096) == _block_sum_py(v, 0, 4096)
    assert on.block_sum(v, 4096, 5000) == _block_sum_py(v, 4096, 5000)


@pytest.mark.parametrize("seed", [1, 42, 7777])
def test_kmeanspp_matches_python_restatement(seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((9000, 3)) * np.array([1.0, 2.0, 0.5])).astype(np.float32)
    k = 6
    C, idx = on.kmeanspp(X, k, seed)
    # restate: d2 with the exact loop (NumPy sequential), picks by the Python rule
    n = len(X)
    picks = [min(int(_uniform_py(seed, 0) * n), n - 1)]
    d2 = None
    for r in range(1, k):
        c = X[picks[-1]]
        t = [(X[:, i] - c[i]).astype(np.float32) for i in range(3)]
        di = np.zeros(n, dtype=np.float32)
        for ti in t:
            di = (di + (ti * ti).astype(np.float32)).astype(np.float32)
        d2 = di if d2 is None else np.where(di < d2, di, d2).astype(np.float32)
        picks.append(_pick_py(d2, _uniform_py(seed, r)))
    assert idx.tolist() == picks
    np.testing.assert_array_equal(C, X[idx])


def test_kmeanspp_picks_have_positive_weight_and_distinct():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((3000, 4)).astype(np.float32)
    C, idx = on.kmeanspp(X, 50, 5)
    assert len(set(idx.tolist())) == 50
    np.testing.assert_array_equal(C, X[idx])
    C2, idx2 = on.kmeanspp(X, 50, 5)
    np.testing.assert_array_equal(idx, idx2)


# ---------------------------------------------------------------- Lloyd
def _update_numpy(X, labels, C):
    C = C.copy()
    for j in range(C.shape[0]):
        mem = np.nonzero(labels == j)[0]
        if len(mem) == 0:
            continue
        acc = np.zeros(X.shape[1])
        for q in range(0, len(mem), 256):
            ch = np.zeros(X.shape[1])
            for m in mem[q:q + 256]:
                ch = ch + X[m].astype(np.float64)
            acc = acc + ch
        C[j] = (acc / len(mem)).astype(np.float32)
    return C


def test_update_centers_matches_numpy():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((1500, 5)).astype(np.float32)
    C = X[:7].copy()
    lab = on.assign(X, C)
    Cn, cnt = on.update_centers(X, lab, C)
    np.testing.assert_array_equal(Cn, _update_numpy(X, lab, C))
    assert cnt.sum() == len(X)


def test_lloyd_cube_equilibrium():
    """test_kmeans.py:154-166: k=1 on the cube vertices -> exactly the origin."""
    X = np.array([[1, 1, 1], [1, 1, -1], [1, -1, -1], [-1, -1, -1], [-1, 1, 1], [-1, -1, 1], [-1, 1, -1],
                  [1, -1, 1]], dtype=np.float32)
    C0, _ = on.kmeanspp(X, 1, 42)
    C, lab, inert, conv = on.lloyd(X, C0, 10, 1e-5)
    np.testing.assert_array_equal(C.squeeze(), np.zeros(3))


def test_lloyd_four_blocks():
    """test_kmeans.py:411-426: four constant blocks -> centers exactly 30/60/90/120."""
    X = np.zeros((40000, 4), dtype=np.float32)
    X[0:10000] = 30.0
    X[10000:20000] = 60.0
    X[20000:30000] = 90.0
    X[30000:] = 120.0
    C0, _ = on.kmeanspp(X, 4, 42)
    C, lab, inert, conv = on.lloyd(X, C0, 10, 1e-5)
    got = sorted(C[:, 0].tolist())
    assert got == [30.0, 60.0, 90.0, 120.0]
    assert all(np.all(c == c[0]) for c in C)
    assert conv


def _dist_to_own_center(X, C, lab):
    """Sequential FP32 distance of each frame to the center of its label (NumPy, every op rounded)."""
    X = X.astype(np.float32)
    Cl = C.astype(np.float32)[lab]
    acc = np.zeros(X.shape[0], dtype=np.float32)
    for i in range(X.shape[1]):
        t = (X[:, i] - Cl[:, i]).astype(np.float32)
        acc = (acc + (t * t).astype(np.float32)).astype(np.float32)
    return acc


def test_lloyd_cost_nonincreasing_and_stopping():
    """The cost of iteration t is the k-means objective of kmeans.py:67-75,
    C(S) = sum_i sum_{x in S_i} ||x - mu_i||^2 with S_i the clusters formed by
    that iteration's assignment and mu_i their means (the updated centers)."""
    rng = np.random.default_rng(12)
    X = rng.standard_normal((4000, 6)).astype(np.float32)
    C0, _ = on.kmeanspp(X, 20, 3)
    C, lab, inert, conv = on.lloyd(X, C0, 30, 0.0)
    assert np.all(np.diff(inert) <= 1e-9 * inert[:-1])
    # the last cost: the last labels against the centers they produced
    assert math.isclose(inert[-1], on.blocked_sum(_dist_to_own_center(X, C, lab)), rel_tol=0, abs_tol=0)
    # iteration by iteration: labels of the centers before the update, cost against the
    # member means (the means minimize the cost of fixed clusters: never above the cost
    # against the centers before the update)
    Cp = C0
    for it in range(4):
        Cn, lab_it, inert_it, _ = on.lloyd(X, C0, it + 1, 0.0)
        np.testing.assert_array_equal(lab_it, on.assign(X, Cp))
        Cm, _ = on.update_centers(X, lab_it, Cp)
        np.testing.assert_array_equal(Cn, Cm)
        cost = on.blocked_sum(_dist_to_own_center(X, Cn, lab_it))
        assert inert_it[-1] == cost
        assert cost <= on.blocked_sum(_dist_to_own_center(X, Cp, lab_it))
        Cp = Cn
