"""Continuous-batching discretization service on one GPU.

A request asks for the discrete trajectory of one float32 trajectory under
one resident model: the TICA projection Y = (X - mean) V[:, :m]
(transform/_tica_base.py:117-133) followed by the assignment of every
projected frame to the model's cluster centers (clustering/interface.py:
157-167).  In the reference the same composition is the discretizer
pipeline (coordinates/api.py:534-611) run once per data set; here many
requests share the device:

  * models (mean, eigenvectors, centers) stay resident in device memory;
  * every admitted request takes a block of one preallocated device arena
    (input, projection, labels, assignment workspace) and one stream from a
    pool; requests are admitted first-in first-out while a stream is free
    and the arena has room (the HBM budget), also while others are running,
    so their copies and kernels overlap on the device;
  * completion is observed through per-request events; the block and the
    stream go back to the pool and the next queued request is admitted.

Results do not depend on the mix: every kernel is deterministic and a
request only touches its own arena block, so a request's labels are
bit-identical whether it runs alone or among others.
"""
import collections
import ctypes
import time

import numpy as np

from . import _native

__all__ = ["ResidentModel", "FirstFit", "Request", "DiscretizationServer", "latency_summary"]

_ALIGN = 256


class FirstFit(object):
    """First-fit allocator of byte ranges inside [0, size) (256-byte aligned)."""

    def __init__(self, size):
        self.size = int(size)
        self._free = [(0, self.size)]  # sorted, non-adjacent (offset, length)
        self._used = {}
        self.in_use = 0
        self.peak = 0

    @staticmethod
    def aligned(n):
        return (int(n) + _ALIGN - 1) // _ALIGN * _ALIGN

    def alloc(self, nbytes):
        n = self.aligned(max(int(nbytes), 1))
        for i, (off, ln) in enumerate(self._free):
            if ln >= n:
                if ln == n:
                    del self._free[i]
                else:
                    self._free[i] = (off + n, ln - n)
                self._used[off] = n
                self.in_use += n
                self.peak = max(self.peak, self.in_use)
                return off
        return None

    def free(self, off):
        n = self._used.pop(off)
        self.in_use -= n
        fl = self._free
        # insert keeping the list sorted, then coalesce with the neighbours
        i = 0
        while i < len(fl) and fl[i][0] < off:
            i += 1
        fl.insert(i, (off, n))
        if i + 1 < len(fl) and fl[i][0] + fl[i][1] == fl[i + 1][0]:
            fl[i] = (fl[i][0], fl[i][1] + fl[i + 1][1])
            del fl[i + 1]
        if i > 0 and fl[i - 1][0] + fl[i - 1][1] == fl[i][0]:
            fl[i - 1] = (fl[i - 1][0], fl[i - 1][1] + fl[i][1])
            del fl[i]

    def largest_free(self):
        return max((ln for _, ln in self._free), default=0)


class ResidentModel(object):
    """A (TICA transform, center set) pair held in device memory.

    mean (d,), eigenvectors (d, >= m) and centers (k, m) follow the
    reference's TICA.mean / TICA.eigenvectors (_tica_base.py:199-239) and
    KmeansClustering.clustercenters; m = dim is the projected dimension.
    """

    def __init__(self, mean, eigenvectors, centers, dim=None, assign_mode=_native.EMMA_ASSIGN_BF16X3_RECHECK):
        t = _native.torch()
        dev = _native.device()
        mean = np.asarray(mean, dtype=np.float64).ravel()
        V = np.asarray(eigenvectors, dtype=np.float64)
        C = np.require(centers, dtype=np.float32, requirements="C")
        m = int(dim) if dim is not None else C.shape[1]
        if V.ndim != 2 or V.shape[0] != mean.size or V.shape[1] < m:
            raise ValueError("eigenvectors must be (d, >= dim) with d = len(mean)")
        if C.ndim != 2 or C.shape[1] != m:
            raise ValueError("cluster centers have wrong dimension. Have dim=%i, but projection has %i"
                             % (C.shape[1] if C.ndim == 2 else -1, m))
        self.d = mean.size
        self.m = m
        self.k = C.shape[0]
        self.assign_mode = int(assign_mode)
        self.mean_dev = t.as_tensor(mean, device=dev)
        self.V_dev = t.as_tensor(np.ascontiguousarray(V[:, :m]), device=dev)
        self.C_dev = t.as_tensor(C, device=dev)

    @classmethod
    def from_estimators(cls, tica, clustering, **kw):
        return cls(tica.mean, tica.eigenvectors, clustering.clustercenters, dim=tica.dimension(), **kw)

    def block_layout(self, T):
        """Byte offsets inside one request block: X, Y, labels, n_rechecked, workspace."""
        L = _native.lib()
        a = FirstFit.aligned
        sizes = [T * self.d * 4, T * self.m * 4, T * 4, 8,
                 L.emma_assign_workspace_size(T, self.m, self.k, self.assign_mode)]
        offs, o = [], 0
        for s in sizes:
            offs.append(o)
            o += a(s)
        return offs, sizes, o


class Request(object):
    """One discretization request and its timing (seconds, time.perf_counter)."""

    __slots__ = ("rid", "model_key", "X", "T", "t_arrival", "t_admit", "t_done", "labels", "n_rechecked",
                 "_off", "_stream", "_event", "_labels_host", "_nrec_host")

    def __init__(self, rid, model_key, X, t_arrival):
        self.rid = rid
        self.model_key = model_key
        self.X = X
        self.T = int(X.shape[0])
        self.t_arrival = t_arrival
        self.t_admit = None
        self.t_done = None
        self.labels = None
        self.n_rechecked = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class DiscretizationServer(object):
    """Admit, run and complete discretization requests (see module docstring).

    hbm_budget  bytes of the device arena shared by the requests in flight
    n_streams   requests that may run at the same time
    """

    def __init__(self, hbm_budget, n_streams=4):
        t = _native.torch()
        self._t = t
        self.dev = _native.device()
        self.alloc = FirstFit(hbm_budget)
        self.arena = t.empty(self.alloc.size, dtype=t.uint8, device=self.dev)
        self.streams = [t.cuda.Stream(device=self.dev) for _ in range(int(n_streams))]
        self._free_streams = list(range(len(self.streams)))
        self.models = {}
        self.queue = collections.deque()
        self.running = []
        self.done = []
        self.max_running = 0
        self._next_id = 0

    # ---- models ----
    def add_model(self, key, model):
        self.models[key] = model

    # ---- requests ----
    def submit(self, model_key, X, t_arrival=None):
        """Queue a request; X is a (T, d) float32 array or (preferably pinned) CPU tensor."""
        if model_key not in self.models:
            raise KeyError("unknown model %r" % (model_key,))
        model = self.models[model_key]
        t = self._t
        if not isinstance(X, t.Tensor):
            X = t.from_numpy(np.require(X, dtype=np.float32, requirements="C"))
        if X.dim() != 2 or X.shape[1] != model.d or X.dtype != t.float32:
            raise ValueError("request data must be (T, %d) float32" % model.d)
        if X.shape[0] == 0:
            raise ValueError("empty trajectory")
        need = model.block_layout(X.shape[0])[2]
        if need > self.alloc.size:
            raise ValueError("request needs %d bytes of device memory, more than the whole budget (%d)"
                             % (need, self.alloc.size))
        r = Request(self._next_id, model_key, X, time.perf_counter() if t_arrival is None else t_arrival)
        self._next_id += 1
        self.queue.append(r)
        return r

    def _launch(self, r, si, off):
        t = self._t
        L = _native.lib()
        model = self.models[r.model_key]
        offs, sizes, _ = model.block_layout(r.T)
        base = off
        a = self.arena
        Xd = a[base + offs[0]: base + offs[0] + sizes[0]].view(t.float32).view(r.T, model.d)
        Yd = a[base + offs[1]: base + offs[1] + sizes[1]].view(t.float32).view(r.T, model.m)
        lab = a[base + offs[2]: base + offs[2] + sizes[2]].view(t.int32)
        nre = a[base + offs[3]: base + offs[3] + sizes[3]].view(t.int64)
        ws = a[base + offs[4]: base + offs[4] + max(sizes[4], 1)]
        s = self.streams[si]
        r._labels_host = t.empty(r.T, dtype=t.int32, pin_memory=True)
        r._nrec_host = t.empty(1, dtype=t.int64, pin_memory=True)
        sp = ctypes.c_void_p(s.cuda_stream)
        with t.cuda.stream(s):
            Xd.copy_(r.X, non_blocking=True)
            _native.check(L.emma_project(_native.ptr(Xd), r.T, model.d, _native.ptr(model.mean_dev),
                                         _native.ptr(model.V_dev), model.m, model.m, _native.ptr(Yd), sp),
                          "emma_project")
            _native.check(L.emma_assign(_native.ptr(Yd), r.T, model.m, _native.ptr(model.C_dev), model.k,
                                        _native.ptr(lab), model.assign_mode, _native.ptr(nre), _native.ptr(ws),
                                        sizes[4], sp), "emma_assign")
            r._labels_host.copy_(lab, non_blocking=True)
            r._nrec_host.copy_(nre, non_blocking=True)
            ev = t.cuda.Event()
            ev.record(s)
        r._event = ev
        r._off = off
        r._stream = si

    def _admit(self):
        admitted = 0
        while self.queue and self._free_streams:
            r = self.queue[0]
            need = self.models[r.model_key].block_layout(r.T)[2]
            off = self.alloc.alloc(need)
            if off is None:
                break  # first-in first-out: wait for memory to free up
            self.queue.popleft()
            si = self._free_streams.pop()
            r.t_admit = time.perf_counter()
            self._launch(r, si, off)
            self.running.append(r)
            admitted += 1
        self.max_running = max(self.max_running, len(self.running))
        return admitted

    def poll(self):
        """Complete finished requests, admit queued ones; returns the newly completed."""
        finished = []
        still = []
        for r in self.running:
            if r._event.query():
                r.t_done = time.perf_counter()
                r.labels = r._labels_host.numpy()
                r.n_rechecked = int(r._nrec_host[0])
                self.alloc.free(r._off)
                self._free_streams.append(r._stream)
                r._event = None
                finished.append(r)
            else:
                still.append(r)
        self.running = still
        self.done.extend(finished)
        self._admit()
        return finished

    def idle(self):
        return not self.queue and not self.running

    def run_until_idle(self, sleep=20e-6):
        self._admit()
        while not self.idle():
            if not self.poll():
                time.sleep(sleep)

    def replay(self, schedule, sleep=20e-6):
        """Submit (t_offset, model_key, X) items at their offsets from now and serve them all.

        Returns the requests in schedule order.  Latency = completion seen by
        the host minus the scheduled arrival time.
        """
        t0 = time.perf_counter()
        items = sorted(enumerate(schedule), key=lambda e: e[1][0])
        out = [None] * len(items)
        i = 0
        while i < len(items) or not self.idle():
            now = time.perf_counter() - t0
            while i < len(items) and items[i][1][0] <= now:
                idx, (toff, key, X) = items[i]
                out[idx] = self.submit(key, X, t_arrival=t0 + toff)
                i += 1
            if not self.poll() and (i >= len(items) or items[i][1][0] > time.perf_counter() - t0):
                time.sleep(sleep)
        return out


def latency_summary(requests):
    lat = np.array([r.latency for r in requests], dtype=np.float64) * 1e3
    frames = sum(r.T for r in requests)
    t_first = min(r.t_arrival for r in requests)
    t_last = max(r.t_done for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "frames": int(frames), "frames_per_s": frames / (t_last - t_first)}
