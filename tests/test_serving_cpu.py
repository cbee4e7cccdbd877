"""Host logic of the discretization service that needs no GPU: the arena allocator."""
import numpy as np

from pyemma_amd.serving import FirstFit


def test_firstfit_alloc_free_coalesce():
    a = FirstFit(4096)
    o1 = a.alloc(100)
    o2 = a.alloc(300)
    o3 = a.alloc(256)
    assert (o1, o2, o3) == (0, 256, 768)
    assert a.in_use == 1024
    assert a.alloc(4096) is None
    a.free(o2)
    assert a.alloc(512) == 256  # first fit reuses the hole
    a.free(o1)
    a.free(o3)
    a.free(256)
    assert a.in_use == 0 and a._free == [(0, 4096)]
    assert a.peak == 1024


def test_firstfit_random_sequence_never_overlaps():
    rng = np.random.default_rng(0)
    a = FirstFit(1 << 20)
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            off = list(live)[rng.integers(len(live))]
            a.free(off)
            del live[off]
        else:
            n = int(rng.integers(1, 50000))
            off = a.alloc(n)
            if off is not None:
                live[off] = FirstFit.aligned(n)
        spans = sorted(live.items())
        for (o, n), (o2, _) in zip(spans, spans[1:]):
            assert o + n <= o2
        assert all(o + n <= a.size for o, n in spans)
        assert a.in_use == sum(live.values())
        assert a.in_use + sum(ln for _, ln in a._free) == a.size
