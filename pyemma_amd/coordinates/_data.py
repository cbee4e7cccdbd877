"""Input handling: trajectories as one device-resident, row-concatenated matrix.

Mirrors what the reference's DataInMemory source provides to the estimators
(coordinates/data/data_in_memory.py:34-153): a list of (T_i, d) arrays with a
common dimension, float input kept as float32 for the device (the reference
keeps the input dtype, data_in_memory.py:108, and converts to float32 for
clustering, clustering/interface.py:159).  The non-finite check of
DataSourceIterator.__next__ (datasource.py:1067-1075) runs on the device.
"""
import numpy as np

from .. import _native
from ..exceptions import InvalidDataInStreamException


class DeviceTrajectories(object):
    """Row-concatenated float32 trajectories on the GPU plus CSR offsets."""

    def __init__(self, X, offsets, offsets_dev=None):
        t = _native.torch()
        self.X = X
        self.offsets = np.asarray(offsets, dtype=np.int64)
        # offsets_dev: the same offsets already on the device (an upload from
        # pageable host memory waits for the stream's queued work)
        self.offsets_dev = t.as_tensor(self.offsets, device=X.device) if offsets_dev is None else offsets_dev
        self.d = int(X.shape[1])

    @property
    def ntraj(self):
        return len(self.offsets) - 1

    @property
    def lengths(self):
        return np.diff(self.offsets)

    @property
    def n_frames(self):
        return int(self.offsets[-1])

    def trajectory(self, i):
        return self.X[int(self.offsets[i]):int(self.offsets[i + 1])]

    def split(self, Y):
        """Split a row-aligned device result back into per-trajectory host arrays."""
        Yh = Y.cpu().numpy()
        return [Yh[self.offsets[i]:self.offsets[i + 1]] for i in range(self.ntraj)]

    def check_finite(self):
        cnt = int(_native.count_nonfinite(self.X).item())
        if cnt:
            raise InvalidDataInStreamException("Non-finite values (%d) in input data" % cnt)
        return self


def _as_list(data):
    t = _native.torch()
    if isinstance(data, DeviceTrajectories):
        return data
    if isinstance(data, (np.ndarray, t.Tensor)):
        return [data]
    if isinstance(data, (list, tuple)):
        if len(data) == 0:
            raise ValueError("empty data")
        return list(data)
    raise ValueError("unsupported data type %s: give an ndarray, a list of ndarrays or torch tensors" % type(data))


def to_device(data, check_finite=True):
    """ndarray / list of ndarrays / torch tensors -> DeviceTrajectories."""
    t = _native.torch()
    items = _as_list(data)
    if isinstance(items, DeviceTrajectories):
        return items.check_finite() if check_finite else items
    dev = _native.device()
    arrs = []
    for a in items:
        if isinstance(a, t.Tensor):
            a2 = a.to(device=dev, dtype=t.float32)
        else:
            a2 = np.asarray(a)
            if a2.dtype.kind not in "fiub":
                raise ValueError("unsupported dtype %s" % a2.dtype)
        if a2.ndim == 1:
            a2 = a2.reshape(-1, 1)
        if a2.ndim != 2:
            raise ValueError("trajectories must be 1- or 2-dimensional")
        arrs.append(a2)
    d = arrs[0].shape[1]
    if any(a.shape[1] != d for a in arrs):
        raise ValueError("all trajectories need the same number of dimensions")
    lengths = [a.shape[0] for a in arrs]
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    if all(isinstance(a, np.ndarray) for a in arrs):
        host = np.concatenate([np.require(a, dtype=np.float32) for a in arrs], axis=0) if len(arrs) > 1 \
            else np.require(arrs[0], dtype=np.float32, requirements="C")
        X = t.from_numpy(np.ascontiguousarray(host)).to(dev)
    else:
        X = t.cat([a if isinstance(a, t.Tensor) else t.from_numpy(np.require(a, dtype=np.float32)).to(dev)
                   for a in arrs], dim=0).contiguous()
    dt = DeviceTrajectories(X, offsets)
    return dt.check_finite() if check_finite else dt
