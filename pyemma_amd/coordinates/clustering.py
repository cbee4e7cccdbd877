"""k-means clustering and assignment to given centers on the GPU.

Mirrors pyemma/coordinates/clustering/interface.py:44-302 (AbstractClustering),
kmeans.py:48-338 (KmeansClustering) and assign.py:33-108 (AssignCenters).
"""
import logging
import math
import os
import random

import numpy as np

from .. import _native
from ._data import DeviceTrajectories, to_device
from .source import FeatureSource

log = logging.getLogger("pyemma_amd.coordinates.clustering")

ASSIGN_MODES = {"bf16x3": _native.EMMA_ASSIGN_BF16X3_RECHECK, "fp32": _native.EMMA_ASSIGN_FP32_PLAIN}


def get_n_jobs(logger=None):
    """_base/parallel.py:2-38: SLURM_CPUS_ON_NODE > PYEMMA_NJOBS > psutil.cpu_count(logical=True)."""
    def _from_env(var):
        try:
            v = int(os.environ[var])
        except (KeyError, ValueError):
            return None
        return v if v > 0 else None
    n = _from_env("SLURM_CPUS_ON_NODE") or _from_env("PYEMMA_NJOBS")
    if n is None:
        import psutil
        n = psutil.cpu_count(logical=True) or 1
    return n


def _strided(dt, stride, skip):
    """Rows skip, skip+stride, ... of every trajectory as a new DeviceTrajectories."""
    if stride == 1 and skip == 0:
        return dt
    t = _native.torch()
    parts, lens = [], []
    for i in range(dt.ntraj):
        p = dt.trajectory(i)[skip::stride]
        parts.append(p)
        lens.append(p.shape[0])
    X = t.cat(parts, dim=0).contiguous() if parts else dt.X[:0]
    return DeviceTrajectories(X, np.concatenate([[0], np.cumsum(lens)]).astype(np.int64))


class AbstractClustering(object):
    """Common interface: centers, assignment, dtrajs (interface.py:44-302)."""

    def __init__(self, metric="euclidean", n_jobs=None, assign_mode="bf16x3"):
        if metric != "euclidean":
            raise NotImplementedError("metric %r: only 'euclidean' is implemented on the GPU path" % metric)
        if assign_mode not in ASSIGN_MODES:
            raise ValueError("assign_mode must be one of %s" % sorted(ASSIGN_MODES))
        self.metric = metric
        self._clustercenters = None
        self._dtrajs = []
        self._previous_stride = -1
        self._data = None
        self.n_jobs = n_jobs
        self.assign_mode = assign_mode
        self.n_rechecked = 0
        self.chunksize = None
        self.stride = 1
        self.skip = 0

    @property
    def n_jobs(self):
        return self._n_jobs

    @n_jobs.setter
    def n_jobs(self, val):
        self._n_jobs = int(val) if val is not None else get_n_jobs()

    @property
    def clustercenters(self):
        return self._clustercenters

    @clustercenters.setter
    def clustercenters(self, val):
        self._clustercenters = np.asarray(val, dtype="float32", order="C")[:] if val is not None else None

    cluster_centers_ = clustercenters

    def dimension(self):
        return 1

    def output_type(self):
        return np.int32()

    def _assign_device(self, dt):
        """labels of all rows of dt (device int32)."""
        t = _native.torch()
        C = _native.as_device_f32(self.clustercenters)
        if C.shape[1] != dt.d:
            raise ValueError("cluster centers have wrong dimension. Have dim=%i, but input has %i"
                             % (C.shape[1], dt.d))
        nre = t.zeros(1, dtype=t.int64, device=dt.X.device)
        labels = _native.assign(dt.X, C, mode=ASSIGN_MODES[self.assign_mode], n_rechecked=nre)
        self.n_rechecked = int(nre.item())
        return labels

    def _assign_source(self, src):
        """Labels of a FeatureSource, one device tensor per trajectory, the
        features of one chunk at a time (rechecked frames summed over chunks
        on the device: the chunks do not wait for the host)."""
        t = _native.torch()
        C = _native.as_device_f32(self.clustercenters)
        if C.shape[1] != src.d:
            raise ValueError("cluster centers have wrong dimension. Have dim=%i, but input has %i"
                             % (C.shape[1], src.d))
        total = t.zeros(1, dtype=t.int64, device=C.device)
        nre = t.zeros(1, dtype=t.int64, device=C.device)

        def lab(dt):
            out = _native.assign(dt.X, C, mode=ASSIGN_MODES[self.assign_mode], n_rechecked=nre)
            total.add_(nre)
            return out

        out = src.per_trajectory(lab)
        self.n_rechecked = int(total.item())
        return out

    def assign_device(self, X):
        """Labels of X kept on the device (DeviceDtrajs), the input of pyemma_amd.dtraj's counts."""
        from ..dtraj import DeviceDtrajs

        if isinstance(X, FeatureSource):
            t = _native.torch()
            labs = self._assign_source(X)
            off = np.concatenate([[0], np.cumsum([len(l) for l in labs])]).astype(np.int64)
            return DeviceDtrajs(t.cat(labs).contiguous(), off)
        dt = to_device(X)
        return DeviceDtrajs(self._assign_device(dt), dt.offsets)

    def transform(self, X):
        """(T, 1) int32 label columns (interface.py:157-167)."""
        if isinstance(X, FeatureSource):
            return [l.cpu().numpy()[:, None] for l in self._assign_source(X)]
        single = not isinstance(X, (list, tuple, DeviceTrajectories))
        dt = to_device(X)
        out = [l[:, None] for l in dt.split(self._assign_device(dt))]
        return out[0] if single and dt.ntraj == 1 else out

    def get_output(self, stride=1, skip=0, chunk=None):
        if self._data is None:
            raise RuntimeError("no data: estimate() with data first")
        if isinstance(self._data, FeatureSource):
            return [l.cpu().numpy()[skip::stride, None] for l in self._assign_source(self._data)]
        dt = _strided(self._data, stride, skip)
        return [l[:, None] for l in dt.split(self._assign_device(dt))]

    def assign(self, X=None, stride=1):
        """interface.py:176-231."""
        if X is None:
            if self._previous_stride is stride and len(self._dtrajs) > 0:
                return self._dtrajs
            self._previous_stride = stride
            mapped = self.get_output(stride=stride, skip=self.skip)
            self._dtrajs = [np.transpose(m)[0] for m in mapped]
            return self._dtrajs
        if stride != 1:
            raise ValueError("assign accepts either X or stride parameters, but not both. If you want to map "
                             "only a subset of your data, extract the subset yourself and pass it as X.")
        mapped = self.transform(X)
        if isinstance(mapped, np.ndarray):
            return np.transpose(mapped)[0]
        return [np.transpose(m)[0] for m in mapped]

    @property
    def dtrajs(self):
        if len(self._dtrajs) == 0:
            self._dtrajs = self.assign(stride=1)
        return self._dtrajs

    def save_dtrajs(self, trajfiles=None, prefix="", output_dir=".", output_format="ascii", extension=".dtraj"):
        """interface.py:233-302."""
        if extension[0] != ".":
            extension = "." + extension
        if output_format == "ascii":
            def write_dtraj(fn, dtraj):
                with open(fn, "w") as f:
                    np.asarray(dtraj).tofile(f, sep="\n", format="%d")
        else:
            def write_dtraj(fn, dtraj):
                np.save(fn, np.asarray(dtraj))
        names = []
        if trajfiles is not None:
            for f in trajfiles:
                base = os.path.splitext(os.path.split(f)[1])[0]
                names.append("%s_%s%s" % (prefix, base, extension) if prefix else "%s%s" % (base, extension))
        else:
            for i in range(len(self.dtrajs)):
                names.append("%s_%i%s" % (prefix, i, extension) if prefix else "%i%s" % (i, extension))
        os.makedirs(output_dir, exist_ok=True)
        for fn, dtraj in zip(names, self.dtrajs):
            dest = os.path.join(output_dir, fn)
            if os.path.exists(dest) and not getattr(self, "overwrite_dtrajs", False):
                raise EnvironmentError("Attempted to write dtraj \"%s\" which already existed. To automatically"
                                       " overwrite existing files, set source.overwrite_dtrajs=True." % dest)
            write_dtraj(dest, dtraj)


class AssignCenters(AbstractClustering):
    """assign.py:33-108."""

    def __init__(self, clustercenters, metric="euclidean", stride=1, n_jobs=None, skip=0, assign_mode="bf16x3"):
        super().__init__(metric=metric, n_jobs=n_jobs, assign_mode=assign_mode)
        if isinstance(clustercenters, str):
            clustercenters = np.loadtxt(clustercenters, ndmin=2) if not clustercenters.endswith(".npy") \
                else np.load(clustercenters)
        clustercenters = np.array(clustercenters, dtype=np.float32, order="C")
        if not clustercenters.ndim == 2:
            raise ValueError("cluster centers have to be 2d")
        self.clustercenters = clustercenters
        self.stride = stride
        self.skip = skip
        self._estimated = True

    def describe(self):
        return "[AssignCenters centers shape=%s]" % (self.clustercenters.shape,)

    def estimate(self, X, **kwargs):
        # a FeatureSource is assigned chunk by chunk (its features never all exist at once)
        dt = X if isinstance(X, FeatureSource) else to_device(X)
        if dt.d != self.clustercenters.shape[1]:
            raise ValueError("cluster centers have wrong dimension. Have dim=%i, but input has %i"
                             % (self.clustercenters.shape[1], dt.d))
        self._data = dt
        self._dtrajs = []
        self._previous_stride = -1
        self.assign(None, self.stride)
        return self


class KmeansClustering(AbstractClustering):
    """kmeans.py:48-338: kmeans++ (or uniform) seeding and Lloyd on the GPU."""

    def __init__(self, n_clusters, max_iter=5, metric="euclidean", tolerance=1e-5, init_strategy="kmeans++",
                 fixed_seed=False, oom_strategy="memmap", stride=1, n_jobs=None, skip=0, clustercenters=None,
                 keep_data=False, assign_mode="bf16x3"):
        super().__init__(metric=metric, n_jobs=n_jobs, assign_mode=assign_mode)
        if clustercenters is None:
            clustercenters = []
        self._converged = False
        self.n_clusters = n_clusters
        self.max_iter = max_iter
        self.tolerance = tolerance
        self.init_strategy = init_strategy
        self.oom_strategy = oom_strategy
        self.fixed_seed = fixed_seed
        self.stride = stride
        self.skip = skip
        self.clustercenters = clustercenters
        self.keep_data = keep_data
        self.inertias = None
        self.initial_centers_ = None

    @property
    def init_strategy(self):
        return self._init_strategy

    @init_strategy.setter
    def init_strategy(self, value):
        valid = ("kmeans++", "uniform")
        if value not in valid:
            raise ValueError('invalid parameter "{}" for init_strategy. Should be one of {}'.format(value, valid))
        self._init_strategy = value

    @property
    def fixed_seed(self):
        return self._fixed_seed

    @fixed_seed.setter
    def fixed_seed(self, val):
        """kmeans.py:142-164."""
        if isinstance(val, bool) or val is None:
            self._fixed_seed = 42 if val else random.randint(0, 2 ** 32 - 1)
        elif isinstance(val, (int, np.integer)):
            if val < 0 or val > 2 ** 32 - 1:
                log.warning("seed has to be positive (or smaller than 2**32-1). Seed will be chosen randomly.")
                self.fixed_seed = False
            else:
                self._fixed_seed = int(val)
        else:
            raise ValueError("fixed seed has to be bool or integer")

    @property
    def converged(self):
        return self._converged

    def describe(self):
        return "[Kmeans, k=%i, inp_dim=%i]" % (self.n_clusters, self._data.d if self._data else -1)

    def _uniform_centers(self, dt):
        """kmeans.py:286-324: Python `random` samples per trajectory, taken in iteration order."""
        lengths = [int(x) for x in dt.lengths]
        total = sum(lengths)
        state = random.getstate()
        random.seed(self.fixed_seed)
        try:
            idx = {}
            for i, L in enumerate(lengths):
                idx[i] = random.sample(list(range(0, L)), int(math.ceil((L / float(total)) * self.n_clusters)))
        finally:
            random.setstate(state)
        rows = []
        for i, L in enumerate(lengths):
            chosen = set(idx[i])
            for l in range(L):
                if len(rows) < self.n_clusters and l in chosen:
                    rows.append(int(dt.offsets[i]) + l)
        return dt.X[rows].clone()

    def estimate(self, X, **kwargs):
        for k, v in kwargs.items():
            if k == "chunksize":
                self.chunksize = v
            elif k in ("clustercenters", "max_iter", "keep_data", "tolerance", "n_clusters", "stride", "skip",
                       "init_strategy", "fixed_seed"):
                setattr(self, k, v)
        # Lloyd needs every frame in every iteration: a FeatureSource is
        # featurized whole, as the reference copies all frames into one buffer
        # (kmeans.py:170-200)
        dt_full = X.materialize() if isinstance(X, FeatureSource) else to_device(X)
        dt = _strided(dt_full, self.stride, self.skip)
        n = dt.n_frames
        if not self.n_clusters:
            self.n_clusters = min(int(math.sqrt(n)), 5000)
            log.info("The number of cluster centers was not specified, using min(sqrt(N), 5000)=%s as "
                     "n_clusters." % self.n_clusters)
        resume = self.clustercenters is not None and self.clustercenters.size != 0
        if resume:
            if len(self.clustercenters) != self.n_clusters:
                raise RuntimeError("Passed clustercenters do not match n_clusters: {} vs. {}".format(
                    len(self.clustercenters), self.n_clusters))
            C = _native.as_device_f32(self.clustercenters).clone()
        elif self.init_strategy == "uniform":
            C = self._uniform_centers(dt)
        else:
            C, _ = _native.kmeanspp(dt.X, self.n_clusters, self.fixed_seed)
        self.initial_centers_ = C.cpu().numpy().copy()
        if self.max_iter > 0:
            _, inertias, n_iter, conv = _native.lloyd(dt.X, C, self.max_iter, self.tolerance,
                                                      mode=ASSIGN_MODES[self.assign_mode])
            it = int(n_iter.item())
            self.inertias = inertias[:it].cpu().numpy()
            self._converged = bool(conv.item())
        else:
            self.inertias = np.zeros(0)
            self._converged = False
        self.clustercenters = C.cpu().numpy()
        if not self._converged:
            log.warning("Algorithm did not reach convergence criterion of %g in %i iterations. Consider "
                        "increasing max_iter.", self.tolerance, self.max_iter)
        # dtrajs refer to the unstrided input (reference: assign(stride=1) over the data producer)
        self._data = dt_full
        self._dtrajs = []
        self._previous_stride = -1
        return self

    def fit(self, X, **kwargs):
        return self.estimate(X, **kwargs)
