"""Tile-shape autotuner of the assignment kernel.

The MFMA assignment kernel exists in a few tile variants (waves per
workgroup x workgroups per CU, emma_assign_variants).  Which one is fastest
depends on the dimension (register footprint of the frame fragments, size of
a center tile) and on the number of centers (tiles per frame block).  The
first call for a (dimension, center count) bucket times every variant on a
sample of the caller's own frames (up to 2^21 frames: smaller samples fill
the device for only a few rounds, where the variants tie and the full-size
winner is not visible) with HIP events on the current stream,
selects the fastest through emma_assign_select, and remembers the choice for
the rest of the process (a lock guards the table; the library's own table is
mutex-guarded).  Variants differ only in speed: labels are identical.
"""
import threading

from . import _native

__all__ = ["tune_assign", "tuned"]

_lock = threading.Lock()
_tuned = {}  # (ks, k bucket) -> (variant, [ms per variant])


def _bucket(d, k):
    ks = -(-d // 16)
    for s in (2, 4, 8, 12, 16, 19, 24, 32):
        if ks <= s:
            ks = s
            break
    kp = -(-k // 32) * 32
    b = 0
    while (32 << b) < kp and b < 15:
        b += 1
    return ks, b


def tuned():
    with _lock:
        return dict(_tuned)


def tune_assign(Xt, Ct, mode=_native.EMMA_ASSIGN_BF16X3_RECHECK, sample=2097152, reps=2):
    """Select the fastest assignment tile variant for these sizes (once per bucket)."""
    if mode != _native.EMMA_ASSIGN_BF16X3_RECHECK:
        return None
    t = _native.torch()
    L = _native.lib()
    n, d = Xt.shape
    k = Ct.shape[0]
    nv = L.emma_assign_variants(d)
    if nv <= 1:
        return 0
    key = _bucket(d, k)
    with _lock:
        if key in _tuned:
            return _tuned[key][0]
        Xs = Xt[: min(n, sample)]
        lab = t.empty(Xs.shape[0], dtype=t.int32, device=Xt.device)
        times = []
        for v in range(nv):
            _native.check(L.emma_assign_select(d, k, v), "emma_assign_select")
            _native.assign(Xs, Ct, mode=mode, labels=lab)  # warm-up
            e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                _native.assign(Xs, Ct, mode=mode, labels=lab)
            e1.record()
            e1.synchronize()
            times.append(e0.elapsed_time(e1) / reps)
        best = min(range(nv), key=lambda i: times[i])
        _native.check(L.emma_assign_select(d, k, best), "emma_assign_select")
        _tuned[key] = (best, times)
        return best
