"""Featurized trajectories streamed through the GPU one chunk at a time.

Mirrors `pyemma.coordinates.source(trajfiles, features=feat, chunksize=...)`
(coordinates/api.py:256-408, the FeatureReader of
coordinates/data/feature_reader.py) for coordinate arrays: the source holds the
particle coordinates and a Featurizer, and every estimator that iterates it
sees the features of a few trajectory segments at a time, so the feature
matrix of the whole data set (12 GB for 10^7 frames of 300 distances) never
exists at once.  The coordinates themselves (3 GB for 10^7 frames of 25
particles) stay where they were given (device tensors, or host arrays
uploaded per chunk).

Segments and the lag: the lagged iteration of the reference
(coordinates/data/_base/iterable.py:182-288) pairs frame t with t + lag inside
one trajectory.  A segment that owns the pair starts [a, a + L) of trajectory
i is featurized over the frames [a, min(a + L + lag, T_i)): read as a
trajectory of its own with skip 0, its lagged pairs are exactly those pairs,
with the stride phase kept because a - skip and L are multiples of the stride.
Segments of one chunk are row-concatenated (DeviceTrajectories), so one
moments call per chunk covers all of them; the per-chunk moments are merged
with the Chan update of LaggedCovariance's storage, as the reference merges
its chunks (running_moments.py:35-148).
"""
import numpy as np

from .. import _native
from ._data import DeviceTrajectories

__all__ = ["FeatureSource", "source", "segment_plan"]

DEFAULT_CHUNK_FRAMES = 1 << 20


def segment_plan(lengths, chunk_frames, lag=0, stride=1, skip=0):
    """Chunks of trajectory segments: a list of chunks, each a list of
    (traj, first_frame, end_frame) frame ranges.

    lag == 0 and stride == 1, skip == 0: the segments partition every
    trajectory (frames for transform / assignment).  Otherwise the segments
    cover the lagged pair starts t = skip + j stride (t + lag < T) exactly once:
    a segment owning the starts [a, a + L) spans the frames
    [a, min(a + L + lag, T)); trajectories with no pair get no segment.
    """
    chunk_frames = max(int(chunk_frames), 1)
    lag, stride, skip = int(lag), max(int(stride), 1), int(skip)
    # pair starts per segment: a multiple of the stride, at least one stride
    L = max(stride, (max(chunk_frames - lag, 1) // stride) * stride)
    chunks, cur, cur_frames = [], [], 0
    for i, T in enumerate(np.asarray(lengths, dtype=np.int64)):
        T = int(T)
        a = skip
        while a + lag < T:
            end = min(a + L + lag, T)
            cur.append((i, a, end))
            cur_frames += end - a
            if cur_frames >= chunk_frames:
                chunks.append(cur)
                cur, cur_frames = [], 0
            a += L
    if cur:
        chunks.append(cur)
    return chunks


class FeatureSource(object):
    """Particle coordinates plus a Featurizer (coordinates.source with features=).

    xyz: one (T, n_atoms, 3) float32 array / torch tensor, or a list of them
    (host arrays are uploaded chunk by chunk).  chunksize: frames of features
    on the device at a time (None: 2^20).
    """

    def __init__(self, xyz, featurizer, chunksize=None):
        t = _native.torch()
        items = list(xyz) if isinstance(xyz, (list, tuple)) else [xyz]
        if not items:
            raise ValueError("no trajectories")
        self._xyz = []
        for x in items:
            if not isinstance(x, t.Tensor):
                x = np.ascontiguousarray(x, dtype=np.float32)
            if x.ndim != 3 or x.shape[1] != featurizer.n_atoms or x.shape[2] != 3:
                raise ValueError("coordinates must be of shape (n_frames, %d, 3)" % featurizer.n_atoms)
            self._xyz.append(x)
        self.featurizer = featurizer
        self.chunksize = int(chunksize) if chunksize else DEFAULT_CHUNK_FRAMES
        if featurizer.dimension() < 1:
            raise ValueError("no features added")

    # ---- the reader interface the estimators use ----
    @property
    def ntraj(self):
        return len(self._xyz)

    @property
    def number_of_trajectories(self):
        return self.ntraj

    def trajectory_lengths(self):
        return np.array([int(x.shape[0]) for x in self._xyz], dtype=np.int64)

    @property
    def lengths(self):
        return self.trajectory_lengths()

    def n_frames_total(self):
        return int(self.trajectory_lengths().sum())

    def dimension(self):
        return self.featurizer.dimension()

    @property
    def d(self):
        return self.dimension()

    def chunk_bytes(self):
        """Largest feature chunk held on the device (bytes)."""
        return int(self.chunksize) * self.dimension() * 4

    def _features(self, segs, off_dev=None):
        t = _native.torch()
        dev = _native.device()
        lengths = [b - a for _, a, b in segs]
        off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        # every segment's features written into its rows of one chunk buffer
        X = t.empty((int(off[-1]), self.dimension()), dtype=t.float32, device=dev)
        for (i, a, b), lo, hi in zip(segs, off[:-1], off[1:]):
            x = self._xyz[i][a:b]
            if not isinstance(x, t.Tensor):
                x = t.as_tensor(x, device=dev)
            elif x.device != dev:
                x = x.to(dev)
            self.featurizer.transform_device_into(x.contiguous(), X[int(lo):int(hi)])
        return DeviceTrajectories(X, off, offsets_dev=off_dev)

    def iter_chunks(self, lag=0, stride=1, skip=0):
        """(segments, DeviceTrajectories of their features) per chunk; see segment_plan.

        Non-finite features (from non-finite coordinates) raise
        InvalidDataInStreamException (datasource.py:1067-1075) once every chunk
        has been handed out: the count is kept on the device, so the chunks do
        not wait for the host."""
        from ..exceptions import InvalidDataInStreamException

        t = _native.torch()
        dev = _native.device()
        plan = segment_plan(self.trajectory_lengths(), self.chunksize, lag=lag, stride=stride, skip=skip)
        # every chunk's offsets uploaded before the first launch (an upload
        # from pageable host memory waits for the stream's queued work)
        offs = [np.concatenate([[0], np.cumsum([b - a for _, a, b in segs])]).astype(np.int64) for segs in plan]
        flat = t.as_tensor(np.concatenate(offs) if offs else np.zeros(0, np.int64), device=dev)
        bad = t.zeros(1, dtype=t.int64, device=dev)
        pos = 0
        for segs, off in zip(plan, offs):
            dt = self._features(segs, off_dev=flat[pos:pos + len(off)])
            pos += len(off)
            bad += _native.count_nonfinite(dt.X)
            yield segs, dt
        cnt = int(bad.item())
        if cnt:
            raise InvalidDataInStreamException("Non-finite values (%d) in input data" % cnt)

    def n_chunks(self, lag=0, stride=1, skip=0):
        return len(segment_plan(self.trajectory_lengths(), self.chunksize, lag=lag, stride=stride, skip=skip))

    def materialize(self):
        """All features at once (what the chunked iteration avoids); for estimators
        that need every frame together, such as k-means' Lloyd iterations."""
        segs = [(i, 0, int(T)) for i, T in enumerate(self.trajectory_lengths())]
        return self._features(segs)

    def per_trajectory(self, fn):
        """fn(DeviceTrajectories of a chunk) -> device tensor with one row per
        frame; returns one torch tensor per trajectory (rows in frame order)."""
        t = _native.torch()
        outs = [[] for _ in range(self.ntraj)]
        for segs, dt in self.iter_chunks():
            Y = fn(dt)
            for (i, a, b), lo, hi in zip(segs, dt.offsets[:-1], dt.offsets[1:]):
                outs[i].append(Y[int(lo):int(hi)])
        return [o[0] if len(o) == 1 else t.cat(o, dim=0) for o in outs]


def source(xyz, features=None, chunksize=None):
    """coordinates.source(...) for coordinate arrays with a distance Featurizer."""
    if features is None:
        raise ValueError("features (a Featurizer) are required for coordinate input")
    return FeatureSource(xyz, features, chunksize=chunksize)
