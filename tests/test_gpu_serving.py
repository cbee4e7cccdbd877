"""Continuous batching: each request's labels are the same alone and in the mix.

Models: three (TICA transform, center set) pairs with d in {16, 48, 130},
projected dimension 8 and k in {8, 64, 300}.  40 requests of 100 - 20,000
frames arrive as a Poisson stream into a server whose device budget holds
only a few requests at a time, so later requests are admitted while earlier
ones run.  Labels must be bit-identical to running every request alone, and
equal to the sequential FP32 oracle applied to the request's projection.
"""
import numpy as np
import pytest

from oracle import native as on
from pyemma_amd import _native
from pyemma_amd.serving import DiscretizationServer, ResidentModel, latency_summary

pytestmark = pytest.mark.gpu


def _models(rng):
    out = {}
    for key, (d, k) in enumerate([(16, 8), (48, 64), (130, 300)]):
        mean = rng.standard_normal(d)
        V, _ = np.linalg.qr(rng.standard_normal((d, 8)))
        C = rng.standard_normal((k, 8)).astype(np.float32) * 2.0
        out[key] = (mean, V, C)
    return out


def _requests(rng, models, n=40):
    reqs = []
    for i in range(n):
        key = int(rng.integers(len(models)))
        d = models[key][0].size
        T = int(np.exp(rng.uniform(np.log(100), np.log(20000))))
        X = (rng.standard_normal((T, d)) * 1.5 + models[key][0]).astype(np.float32)
        reqs.append((key, X))
    return reqs


def test_alone_equals_mix_and_oracle():
    t = _native.torch()
    rng = np.random.default_rng(5)
    models = _models(rng)
    reqs = _requests(rng, models)

    def server(budget, n_streams):
        s = DiscretizationServer(budget, n_streams=n_streams)
        for key, (mean, V, C) in models.items():
            s.add_model(key, ResidentModel(mean, V, C))
        return s

    alone = []
    s1 = server(64 << 20, 1)
    for key, X in reqs:
        r = s1.submit(key, X)
        s1.run_until_idle()
        alone.append(r.labels.copy())

    budget = 24 << 20  # a few of the larger requests at a time
    s2 = server(budget, 4)
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, size=len(reqs)))
    mixed = s2.replay([(float(a), key, X) for a, (key, X) in zip(arrivals, reqs)])
    assert s2.max_running >= 2, "no overlap happened"
    assert s2.alloc.peak <= budget
    for i, (r, lab) in enumerate(zip(mixed, alone)):
        assert np.array_equal(r.labels, lab), "request %d differs between alone and mix" % i

    # labels = sequential FP32 oracle on the request's own projection
    for (key, X), lab in zip(reqs[:10], alone[:10]):
        mean, V, C = models[key]
        m = s1.models[key]
        Y = _native.project(_native.as_device_f32(X), m.mean_dev, m.V_dev, 8)
        t.cuda.synchronize()
        ref = on.assign(Y.cpu().numpy(), C, nthreads=4)
        assert np.array_equal(lab, ref)
    summ = latency_summary(mixed)
    assert summ["requests"] == len(reqs) and summ["p99_ms"] >= summ["p50_ms"] > 0


def test_request_larger_than_budget_rejected():
    rng = np.random.default_rng(1)
    mean = rng.standard_normal(32)
    V, _ = np.linalg.qr(rng.standard_normal((32, 4)))
    C = rng.standard_normal((16, 4)).astype(np.float32)
    s = DiscretizationServer(1 << 20, n_streams=2)
    s.add_model("m", ResidentModel(mean, V, C))
    with pytest.raises(ValueError):
        s.submit("m", np.zeros((100000, 32), np.float32))
    with pytest.raises(ValueError):
        s.submit("m", np.zeros((10, 31), np.float32))


def test_alone_equals_mix_serving_model_shapes():
    """The serving workload's model shapes: 8 resident models with d cycling over
    {64, 128, 300, 512}, k over {256, 1000, 4096}, m = 32 projected dims;
    24 Poisson-arriving requests of 1,000 - 50,000 Gaussian-mixture frames.
    Labels alone == in the mix, and == the sequential FP32 oracle on the
    request's projection (every request)."""
    from pyemma_amd.synthetic import gaussian_mixture

    t = _native.torch()
    rng = np.random.default_rng(55)
    dims, ks = [64, 128, 300, 512], [256, 1000, 4096]
    pools = {d: gaussian_mixture(60000, d, seed=70 + d, device=_native.device()).cpu().numpy() for d in dims}
    models = {}
    for mi in range(8):
        d, k = dims[mi % 4], ks[mi % 3]
        pool = pools[d]
        mean = pool[:5000].astype(np.float64).mean(0)
        V, _ = np.linalg.qr(rng.standard_normal((d, 32)))
        rows = np.sort(rng.choice(len(pool), size=k, replace=False))
        C = ((pool[rows].astype(np.float64) - mean) @ V).astype(np.float32)
        models[mi] = (mean, V, C)
    reqs = []
    for _ in range(24):
        key = int(rng.integers(8))
        d = models[key][0].size
        T = int(np.exp(rng.uniform(np.log(1000), np.log(50000))))
        s = int(rng.integers(0, 60000 - T + 1))
        reqs.append((key, np.ascontiguousarray(pools[d][s:s + T])))

    def server(budget, n_streams):
        srv = DiscretizationServer(budget, n_streams=n_streams)
        for key, (mean, V, C) in models.items():
            srv.add_model(key, ResidentModel(mean, V, C))
        return srv

    s1 = server(512 << 20, 1)
    alone = []
    for key, X in reqs:
        r = s1.submit(key, X)
        s1.run_until_idle()
        alone.append(r.labels.copy())
    budget = 320 << 20
    s2 = server(budget, 8)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, size=len(reqs)))
    mixed = s2.replay([(float(a), key, X) for a, (key, X) in zip(arrivals, reqs)])
    assert s2.max_running >= 2, "no overlap happened"
    assert s2.alloc.peak <= budget
    for i, (r, lab) in enumerate(zip(mixed, alone)):
        assert np.array_equal(r.labels, lab), "request %d differs between alone and mix" % i
    for (key, X), lab in zip(reqs, alone):
        m = s1.models[key]
        Y = _native.project(_native.as_device_f32(X), m.mean_dev, m.V_dev, 32)
        t.cuda.synchronize()
        ref = on.assign(Y.cpu().numpy(), models[key][2], nthreads=8)
        assert np.array_equal(lab, ref)
