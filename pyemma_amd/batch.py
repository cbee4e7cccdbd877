"""Many uneven trajectories in one captured launch sequence.

The discretization of a whole data set -- lagged moments of every trajectory
(LaggedCovariance over RunningCovar.add, coordinates/estimation/
covariance.py:219-244) and the assignment of every frame to a center set
(AbstractClustering._transform_array, clustering/interface.py:157-167) -- is
one fixed sequence of kernel launches over device-resident inputs:

  emma_moments_lagged: k_prep -> moment tiles (one launch per task kind over
                       all trajectories and pair slices) -> FP64 reductions
  emma_assign:         center mean / pack -> MFMA distances -> exact recheck

BatchedDiscretizer runs that sequence eagerly once (which also lets the
autotuner pick the tile shapes for these sizes), then captures it into a HIP
graph; replay() re-issues the whole sequence with one call.  The graph reads
the input buffers in place, so new frames of the same shape can be copied
into `X` between replays.  Nothing inside the sequence allocates memory, uses
a memset node or synchronizes the host.
"""
import numpy as np

from . import _native, autotune

__all__ = ["BatchedDiscretizer"]


class BatchedDiscretizer(object):
    """Lagged moments + assignment of all trajectories, captured as one graph.

    X          (n_rows, d) float32 device tensor, trajectories concatenated
    offsets    int64 trajectory offsets (ntraj + 1)
    lag        TICA lag time
    centers    (k, d) float32 cluster centers
    precision  moment tile arithmetic ('bf16x4-acc', 'bf16x4', 'fp32', 'fp64')
    """

    def __init__(self, X, offsets, lag, centers, precision="bf16x4-acc", reversible=True,
                 assign_mode=_native.EMMA_ASSIGN_BF16X3_RECHECK):
        from .coordinates.covariance import PRECISIONS

        t = _native.torch()
        dev = _native.device()
        self.X = _native.as_device_f32(X)
        self.offsets = t.as_tensor(np.asarray(offsets, dtype=np.int64), device=dev)
        self.C = _native.as_device_f32(centers)
        if self.C.shape[1] != self.X.shape[1]:
            raise ValueError("cluster centers have wrong dimension. Have dim=%i, but input has %i"
                             % (self.C.shape[1], self.X.shape[1]))
        self.lag = int(lag)
        self.reversible = bool(reversible)
        self.precision = PRECISIONS[precision]
        self.assign_mode = int(assign_mode)
        n, d = self.X.shape
        self.labels = t.empty(n, dtype=t.int32, device=dev)
        self.n_rechecked = t.zeros(1, dtype=t.int64, device=dev)
        self.graph = None
        self.out = None

    def _sequence(self):
        self.out = _native.moments_lagged(self.X, self.offsets, self.lag, reversible=self.reversible,
                                          remove_mean=True, c0t=True, precision=self.precision)
        _native.assign(self.X, self.C, mode=self.assign_mode, labels=self.labels, n_rechecked=self.n_rechecked)

    def run_eager(self):
        """One uncaptured pass (tunes tile shapes on first use); returns host results."""
        autotune.tune_assign(self.X, self.C, self.assign_mode)
        self._sequence()
        return self.results()

    def capture(self):
        """Eager warm-up, then capture the launch sequence into a HIP graph."""
        t = _native.torch()
        self.run_eager()
        t.cuda.synchronize()
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            self._sequence()
        self.graph = g
        return self

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    def results(self):
        """(w, mean, C00, C0t, labels) on the host (synchronizes)."""
        w, sx, sy, Mxx, Mxy, _ = self.out
        wh = float(w.item())
        return (wh, sx.cpu().numpy() / wh, Mxx.cpu().numpy() / wh, Mxy.cpu().numpy() / wh,
                self.labels.cpu().numpy())
