"""Host-only: the chunk plan of a streamed FeatureSource covers every lagged
pair of every trajectory exactly once (the pairs of iterable.py:182-288:
t = skip + j * stride, t + lag < T), and with lag 0 partitions the frames."""
import numpy as np
import pytest

from pyemma_amd.coordinates.source import segment_plan


def _pairs_full(lengths, lag, stride, skip):
    return sorted((i, t) for i, T in enumerate(lengths) for t in range(skip, T - lag, stride))


def _pairs_plan(plan, lag, stride):
    out = []
    for chunk in plan:
        for i, a, b in chunk:
            # the segment read as its own trajectory with skip 0
            out.extend((i, a + tl) for tl in range(0, (b - a) - lag, stride))
    return sorted(out)


@pytest.mark.parametrize("seed", range(12))
def test_every_pair_once(seed):
    rng = np.random.default_rng(seed)
    lengths = list(rng.integers(0, 400, size=int(rng.integers(1, 8))))
    lag = int(rng.integers(0, 30))
    stride = int(rng.integers(1, 5))
    skip = int(rng.integers(0, 20))
    chunk = int(rng.integers(1, 300))
    plan = segment_plan(lengths, chunk, lag=lag, stride=stride, skip=skip)
    assert _pairs_plan(plan, lag, stride) == _pairs_full(lengths, lag, stride, skip)
    L = max(stride, (max(chunk - lag, 1) // stride) * stride)
    for c in plan:
        assert sum(b - a for _, a, b in c) < chunk + L + lag  # a chunk closes once it holds chunk frames
        for i, a, b in c:
            assert (a - skip) % stride == 0 and 0 <= a < b <= lengths[i]


def test_lag0_partitions_frames():
    lengths = [5, 0, 17, 3, 40]
    plan = segment_plan(lengths, 8)
    got = sorted((i, f) for c in plan for i, a, b in c for f in range(a, b))
    assert got == [(i, f) for i, T in enumerate(lengths) for f in range(T)]


def test_short_trajectories_give_no_segment():
    assert segment_plan([5, 10], 100, lag=10) == []
    assert segment_plan([5, 11], 100, lag=10) == [[(1, 0, 11)]]
