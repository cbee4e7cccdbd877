"""Benchmark of the TICA-moments + assignment path on one MI355X.

Workload (default): 200 trajectories x 50,000 frames x 300 float32 features
(pairwise distances of 25 Brownian particles in a 3-basin bias, generated on
the GPU), TICA lag 20 and assignment to k = 1000 centers (kmeans++ on a
200k-frame subsample, outside the timed region).

One step = lagged moments of all pairs (bf16 hi/lo split on MFMA with two-level
FP32 accumulation, FP64 merge: eigenvalues within 1e-6 of FP64), then the assignment of every frame (bf16x3 MFMA + exact recheck) on a second
stream while the host solves the FP64 eigenproblem of the moments.  Prints
one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s TICA+assign (d=300,k=1000) vs plain GPU & CPU; serving p99 ms"
PEAK_BF16_TFLOPS = 2516.6
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_d300_k1000.md")


KERNEL_SOURCES = {"moments": ["moments.hip", "emma_common.hpp"], "assign": ["assign.hip", "emma_common.hpp"]}


def source_digest(kernel):
    """sha256 (first 16 hex digits) of the HIP sources a kernel is built from."""
    import hashlib

    h = hashlib.sha256()
    for name in KERNEL_SOURCES[kernel]:
        with open(os.path.join(ROOT, "pyemma_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel, default_workload):
    """HBM bytes per call of `kernel` from the committed FETCH_SIZE/WRITE_SIZE summary
    (calibrated against a plain copy).  None when the summary does not cover it, when
    the run is not the profiled workload, or when the kernel's sources changed since
    the summary was taken (its `sources:` line records their digest)."""
    if not default_workload:
        return None
    vals, digests = {}, {}
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                for key, dst in (("traffic_bytes:", vals), ("sources:", digests)):
                    if line.startswith(key):
                        for item in line.split(":", 1)[1].split(";"):
                            name, _, val = item.strip().partition("=")
                            dst[name] = val
    except OSError:
        return None
    if kernel not in vals or digests.get(kernel) != source_digest(kernel):
        return None
    return int(vals[kernel])


def mom_executed_flops(d, P):
    """bf16 MFMA flops the reversible split moments kernel executes: the
    64 x 64 blocks on and above the diagonal of the d x d result (blocks
    wholly past d skipped), two matrices (UU, VV), four bf16 products (hh, hl,
    lh, ll) per element, 2 flops per product and pair."""
    nb = -(-d // 64)
    return nb * (nb + 1) // 2 * 64 * 64 * 2 * 4 * 2 * float(P)


def asg_first_pass_products(d, variant):
    """MFMA products per dimension of the assignment's first pass for a tile
    variant: 1 for the coarse fp16 pass (variants below the shape count, except
    the resident fine kernel, variant 0 for d <= 128), 3 for the bf16 split."""
    from pyemma_amd import _native

    shapes = _native.lib().emma_assign_variants(d) // 2
    if variant is None or variant >= shapes or (variant == 0 and d <= 128):
        return 3
    return 1


def asg_executed_flops(d, k, n, products=3):
    """MFMA flops of the assignment's first pass over all n frames: `products`
    products (1: coarse fp16; 3: bf16 hh, hl, lh) over d padded to the K-step
    width and k padded to 32.  The passes over the listed frames come on top."""
    ks = -(-d // 16)
    for s in (2, 4, 8, 12, 16, 19, 24, 32):
        if ks <= s:
            ks = s
            break
    return products * 2.0 * n * (ks * 16) * (-(-k // 32) * 32)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ntraj", type=int, default=200)
    ap.add_argument("--nframes", type=int, default=50000)
    ap.add_argument("--k", type=int, default=1000)
    ap.add_argument("--lag", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 GPU comparison")
    ap.add_argument("--no-hbm-legs", action="store_true", help="skip the small-d assignment legs")
    ap.add_argument("--no-large-k", action="store_true", help="skip the 10^7 x 512, k = 4096 k-means leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the 4096-trajectory batched leg")
    ap.add_argument("--no-legs", action="store_true",
                    help="only the timed steps: skip the other moments modes and the adjacent, projected and "
                         "k-means legs (for profiler runs)")
    ap.add_argument("--cpu-frames", type=int, default=100000, help="frames in the CPU moments sample")
    ap.add_argument("--cpu-assign-frames", type=int, default=10000, help="frames in the CPU assign sample")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="arrival window of the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate of requests per second")
    ap.add_argument("--serve-pool", type=int, default=1000000, help="frames per d in the pinned host data pool")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    from pyemma_amd import _native, _solvers
    from pyemma_amd.synthetic import brownian_distances

    dev = _native.device()
    default_workload = (args.ntraj, args.nframes, args.k, args.lag) == (200, 50000, 1000, 20)
    t0 = time.time()
    X, offsets, xyz = brownian_distances(args.ntraj, args.nframes, seed=2, device=dev, return_positions=True)
    torch.cuda.synchronize()
    N, d = X.shape
    log("data %d x %d generated in %.1f s" % (N, d, time.time() - t0))
    off_dev = torch.as_tensor(offsets, device=dev)
    sub = X[::max(1, N // 200000)].contiguous()
    C, _ = _native.kmeanspp(sub, args.k, 12345)
    torch.cuda.synchronize()
    log("centers: kmeans++ k=%d on %d frames" % (args.k, sub.shape[0]))

    from pyemma_amd import autotune
    best = autotune.tune_assign(X, C)
    log("assign tile variant %s (ms per variant on a sample: %s)" % (best, autotune.tuned()))

    s_mom = torch.cuda.current_stream()
    s_asg = torch.cuda.Stream()
    labels = torch.empty(N, dtype=torch.int32, device=dev)
    nre = torch.zeros(1, dtype=torch.int64, device=dev)
    ev = {k: torch.cuda.Event(enable_timing=True) for k in ("m0", "m1")}
    ea = {k: torch.cuda.Event(enable_timing=True) for k in ("a0", "a1")}

    eig_s = []
    last = {}

    def step(prec=_native.EMMA_PREC_BF16X4_ACC, amode=_native.EMMA_ASSIGN_BF16X3_RECHECK, labels=labels):
        # moments first (each kernel fills the GPU on its own), then the
        # assignment on the second stream while the host solves the d x d
        # eigenproblem of the moments
        ev["m0"].record(s_mom)
        w, sx, sy, Mxx, Mxy, _ = _native.moments_lagged(X, off_dev, args.lag, reversible=True, remove_mean=True,
                                                        c0t=True, precision=prec)
        ev["m1"].record(s_mom)
        s_asg.wait_stream(s_mom)
        with torch.cuda.stream(s_asg):
            ea["a0"].record(s_asg)
            _native.assign(X, C, mode=amode, labels=labels, n_rechecked=nre)
            ea["a1"].record(s_asg)
        wh = w.item()  # waits for the moments only
        C00 = Mxx.cpu().numpy() / wh
        C0t = Mxy.cpu().numpy() / wh
        te = time.perf_counter()
        l, R = _solvers.eig_corr(C00, C0t, 1e-6, sign_maxelement=True)
        eig_s.append(time.perf_counter() - te)
        torch.cuda.current_stream().wait_stream(s_asg)
        torch.cuda.synchronize()
        last["C00"], last["C0t"] = C00, C0t
        return l, ev["m0"].elapsed_time(ev["m1"]), ea["a0"].elapsed_time(ea["a1"])

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times, tm, ta = [], [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        l, mms, ams = step()
        times.append(time.perf_counter() - t1)
        tm.append(mms)
        ta.append(ams)
    torch.cuda.synchronize()
    step_s = float(np.mean(times))
    value = N / step_s
    n_re = int(nre.item())
    P = int(sum(max(0, (b - a) - args.lag) for a, b in zip(offsets[:-1], offsets[1:])))
    mom_ms, asg_ms = float(np.mean(tm)), float(np.mean(ta))
    f_mom = 6.0 * d * d * P
    f_asg = 2.0 * d * args.k * N
    kern = {
        "moments": {"ms": mom_ms, "algorithmic_tflop": f_mom / 1e12, "tflops": f_mom / (mom_ms * 1e-3) / 1e12},
        "assign": {"ms": asg_ms, "algorithmic_tflop": f_asg / 1e12, "tflops": f_asg / (asg_ms * 1e-3) / 1e12,
                   "rechecked_frames": n_re, "tile_variant": best},
    }
    dom = "assign" if asg_ms >= mom_ms else "moments"
    ach = kern[dom]["tflops"]
    roofline = {"bound": "mfma", "kernel": dom, "achieved": ach, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_BF16_TFLOPS, "traffic": measured_traffic(dom, default_workload)}
    eig_ms = float(np.mean(eig_s[-args.steps:])) * 1e3
    kern["eig_host"] = {"ms": eig_ms, "threads": _solvers.blas_threads()}
    log("step %.2f ms (moments %.2f ms, assign %.2f ms, host eig %.2f ms, rechecked %d)"
        % (step_s * 1e3, mom_ms, asg_ms, eig_ms, n_re))

    # the other moments modes, one call each on the same stream: FP64 tiles,
    # and the bf16 split with single-level FP32 accumulation; their matrices
    # are kept for the eigenvalue comparison against the FP64 oracle
    mats = {"bf16x4_acc": (last["C00"], last["C0t"])}
    other_modes = () if args.no_legs else (("moments_fp64", "fp64", _native.EMMA_PREC_FP64),
                                           ("moments_bf16x4", "bf16x4", _native.EMMA_PREC_BF16X4))
    for key, name, prec in other_modes:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        w, _, _, Mxx, Mxy, _ = _native.moments_lagged(X, off_dev, args.lag, reversible=True, remove_mean=True,
                                                      c0t=True, precision=prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        wh = w.item()
        mats[name] = (Mxx.cpu().numpy() / wh, Mxy.cpu().numpy() / wh)
        kern[key] = {"ms": ms, "algorithmic_tflop": f_mom / 1e12, "tflops": f_mom / (ms * 1e-3) / 1e12}
        log("%s %.2f ms" % (key, ms))

    if not args.no_legs:
        kern.update(adjacent_legs(args, X, off_dev, offsets, labels, xyz))
        kern["streamed_features"] = streamed_leg(args, xyz, offsets, C, labels)
        kern["tica_projected"] = projected_leg(args, X, off_dev, mom_ms)
        kern["kmeans"] = kmeans_leg(args, X, sub, cpu=not args.no_cpu)
    del xyz

    plain = None
    labels_plain = None
    if not args.no_plain:
        # the plain path writes its own label buffer: the headline labels stay
        # untouched for the parity checks below
        labels_plain = torch.empty_like(labels)
        t1 = time.perf_counter()
        step(_native.EMMA_PREC_FP32, _native.EMMA_ASSIGN_FP32_PLAIN, labels=labels_plain)
        tp = time.perf_counter() - t1
        plain = {"value": N / tp, "unit": "frames/s", "ms_per_step": tp * 1e3,
                 "moments_ms": ev["m0"].elapsed_time(ev["m1"]), "assign_ms": ea["a0"].elapsed_time(ea["a1"]),
                 "kind": "FP32 MFMA moments + sequential FP32 assign loop"}
        log("plain GPU step %.1f ms" % (tp * 1e3))

    parity = {}
    if labels_plain is not None:
        mism = int((labels != labels_plain).sum().item())
        parity.update({"labels_identical_vs_plain_gpu": mism == 0, "label_mismatches_vs_plain_gpu": mism,
                       "label_frames_vs_plain_gpu": N})
        log("headline labels vs plain FP32 loop: %d mismatches over %d frames" % (mism, N))
        del labels_plain
    cpu = None
    if not args.no_cpu:
        cpu, par = cpu_baseline(args, X, offsets, C, labels, mats)
        parity.update(par)
    kern["moments"]["executed_tflop"] = mom_executed_flops(d, P) / 1e12
    kern["assign"]["executed_tflop"] = asg_executed_flops(d, args.k, N, asg_first_pass_products(d, best)) / 1e12
    ex = kern[dom]["executed_tflop"] * 1e12 / (kern[dom]["ms"] * 1e-3) / 1e12
    roofline.update({"executed_tflops": ex, "executed_frac": ex / PEAK_BF16_TFLOPS})

    serving = None
    if args.serve_seconds > 0 or not args.no_hbm_legs or not args.no_large_k or not args.no_batched:
        del X, sub, labels
        torch.cuda.empty_cache()
    if not args.no_hbm_legs:
        kern.update(assign_hbm_legs())
    if not args.no_large_k:
        kern["large_k"] = large_k_leg(args)
    if not args.no_batched:
        kern["batched"] = batched_leg(args)
    if args.serve_seconds > 0:
        serving = serving_leg(args)

    out = {
        "metric": METRIC, "value": value, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16x4 two-level moments / bf16x3+fp32-recheck assign, fp64 merge",
        "data": "synthetic",
        "config": {"workload": "%d traj x %d frames x %d f32 (Brownian pairwise distances), TICA lag %d, "
                               "assign k=%d" % (args.ntraj, args.nframes, d, args.lag, args.k),
                   "frames": N, "lagged_pairs": P},
        "roofline": roofline, "kernels": kern, "plain_gpu": plain, "cpu_baseline": cpu, "parity": parity,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


def _timed(fn, reps=3):
    """Average device time of fn() over reps calls after one warm-up, HIP events on the current stream."""
    import torch

    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def adjacent_legs(args, X, off_dev, offsets, labels, xyz):
    """The producer and consumers either side of the path, on the same workload:
    pairwise-distance featurization of the particle positions (HBM-bound),
    transition counts of the assigned labels at the TICA lag (HBM/atomic-bound)
    and the VAMP covariances C00, C0t, Ctt (non-reversible moments)."""
    import torch

    from pyemma_amd import _native
    from pyemma_amd.coordinates.featurizer import Featurizer

    N, d = X.shape
    out = {}
    na = xyz.shape[1]
    f = Featurizer(na)
    f.add_distances(np.arange(na))
    Y = f.transform_device(xyz)
    diff = float((Y - X).abs().max().item())
    ms = _timed(lambda: f.transform_device(xyz))
    nbytes = N * (12 * na + 4 * d)
    out["featurize"] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                        "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "max_abs_diff_vs_fp64_features": diff}
    del Y
    ms = _timed(lambda: _native.count_transitions(labels, off_dev, args.lag, True, args.k))
    C, _, bad = _native.count_transitions(labels, off_dev, args.lag, True, args.k)
    total = int(C.sum().item())
    expect = int(sum(max(0, (b - a) - args.lag) for a, b in zip(offsets[:-1], offsets[1:])))
    nbytes = 4 * N + 8 * args.k * args.k
    out["counts"] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                     "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "pairs_counted": total, "pairs_expected": expect, "bad": int(bad.item())}
    ms = _timed(lambda: _native.moments_lagged(X, off_dev, args.lag, reversible=False, remove_mean=True, c0t=True,
                                               ctt=True, precision=_native.EMMA_PREC_BF16X4_ACC), reps=1)
    out["vamp_moments"] = {"ms": ms, "precision": "bf16x4-acc", "matrices": "C00, C0t, Ctt (separate means)"}
    log("featurize %.2f ms, counts %.2f ms, vamp moments %.2f ms"
        % (out["featurize"]["ms"], out["counts"]["ms"], out["vamp_moments"]["ms"]))
    torch.cuda.synchronize()
    return out


def streamed_leg(args, xyz, offsets, C, labels_ref, chunk=1 << 20):
    """The TICA moments and the assignment with the features streamed from the
    particle positions (coordinates.source(..., features=...)): every chunk of
    about 2^20 frames is featurized on the device, its moments merged with the
    Chan update and its frames assigned, so the 12 GB feature matrix never
    exists (peak: one chunk of features).  Timed against the same estimators
    on the materialized features of the same featurizer (featurize all,
    moments, assign); eigenvalues and labels compared between the two, and the
    labels against the headline's (whose features the generator computes in
    FP64, so near-ties can differ)."""
    import torch

    from pyemma_amd import _solvers
    from pyemma_amd.coordinates import AssignCenters, Featurizer, LaggedCovariance, source
    from pyemma_amd.coordinates._data import DeviceTrajectories

    na = xyz.shape[1]
    f = Featurizer(na)
    f.add_distances(np.arange(na))
    trajs = [xyz[int(a):int(b)] for a, b in zip(offsets[:-1], offsets[1:])]
    src = source(trajs, features=f, chunksize=chunk)
    Cc = C.cpu().numpy()
    cov_kw = dict(c00=True, c0t=True, remove_data_mean=True, reversible=True, bessel=False, lag=args.lag,
                  precision="bf16x4-acc")

    def run(data):
        cov = LaggedCovariance(**cov_kw)
        cov.estimate(data)
        lab = AssignCenters(Cc).assign_device(data)
        torch.cuda.synchronize()
        return cov, lab

    def materialized():
        dt = DeviceTrajectories(f.transform_device(xyz), offsets)
        return run(dt)

    def timed(fn):
        fn()
        t0 = time.perf_counter()
        r = fn()
        return r, (time.perf_counter() - t0) * 1e3

    (cov_m, lab_m), ms_m = timed(materialized)
    (cov_s, lab_s), ms_s = timed(lambda: run(src))
    l_s, _ = _solvers.eig_corr(cov_s.C00_, cov_s.C0t_, 1e-6, sign_maxelement=True)
    l_m, _ = _solvers.eig_corr(cov_m.C00_, cov_m.C0t_, 1e-6, sign_maxelement=True)
    n = min(len(l_s), len(l_m))
    mism = int((lab_s.S != lab_m.S).sum().item())
    mism_ref = int((lab_s.S.reshape(-1) != labels_ref.reshape(-1)).sum().item())
    out = {"ms": ms_s, "materialized_ms": ms_m, "chunk_frames": chunk, "n_chunks": src.n_chunks(lag=args.lag),
           "peak_feature_bytes": src.chunk_bytes(),
           "materialized_feature_bytes": int(lab_m.S.numel()) * f.dimension() * 4,
           "eig_rel_diff_vs_materialized": float(np.max(np.abs(l_s[:n] - l_m[:n]) / np.abs(l_m[:n]))),
           "label_mismatches_vs_materialized": mism,
           "label_mismatches_vs_headline_features": mism_ref}
    log("streamed features: moments + assign %.1f ms in %d chunks (features on the device at a time: %.2f GB "
        "instead of %.1f GB; materialized featurize + moments + assign %.1f ms); eig rel diff %.1e, label "
        "mismatches %d (vs the headline's FP64-generated features: %d)"
        % (ms_s, out["n_chunks"], out["peak_feature_bytes"] / 1e9, out["materialized_feature_bytes"] / 1e9, ms_m,
           out["eig_rel_diff_vs_materialized"], mism, mism_ref))
    return out


def projected_leg(args, X, off_dev, mom_ms, m=32, n_seed=200_000):
    """The TICA-projected variant of the pipeline (tica, then k-means and
    assignment on its output): every frame projected onto the m leading
    kinetic-map TICA components (Y = (X - mean) R[:, :m] with R scaled by the
    eigenvalues, FP64 math, float32 out), then assigned to k centers that
    kmeans++ picks from a subsample of Y.  Reports the projection and
    assignment times and the pipeline rate N / (moments + projection +
    assignment) beside the headline."""
    import torch

    from pyemma_amd import _native, _solvers, autotune

    N, d = X.shape
    w, sx, _, Mxx, Mxy, _ = _native.moments_lagged(X, off_dev, args.lag, reversible=True, remove_mean=True,
                                                    c0t=True, precision=_native.EMMA_PREC_BF16X4_ACC)
    wh = float(w.item())
    l, R = _solvers.eig_corr(Mxx.cpu().numpy() / wh, Mxy.cpu().numpy() / wh, 1e-6, sign_maxelement=True)
    m = min(m, R.shape[1])
    V = np.ascontiguousarray(R[:, :m] * l[None, :m])  # kinetic map
    mean_dev = (sx / w).contiguous()
    V_dev = torch.as_tensor(V, device=X.device)
    Y = _native.project(X, mean_dev, V_dev, m)
    ms_proj = _timed(lambda: _native.project(X, mean_dev, V_dev, m))
    rng = np.random.default_rng(3)
    sub = Y[torch.as_tensor(np.sort(rng.choice(N, size=min(n_seed, N), replace=False)), device=X.device)]
    C, _ = _native.kmeanspp(sub.contiguous(), args.k, 7)
    autotune.tune_assign(Y, C)
    lab = torch.empty(N, dtype=torch.int32, device=X.device)
    nre = torch.zeros(1, dtype=torch.int64, device=X.device)
    ms_asg = _timed(lambda: _native.assign(Y, C, labels=lab, n_rechecked=nre))
    torch.cuda.synchronize()
    total = mom_ms + ms_proj + ms_asg
    proj_bytes = 4 * (d + m) * N
    out = {"dim": m, "k": args.k, "project_ms": ms_proj,
           "project_gbs": proj_bytes / (ms_proj * 1e-3) / 1e9,
           "assign_ms": ms_asg, "rechecked_frames": int(nre.item()),
           "pipeline_ms": total, "pipeline_frames_per_s": N / (total * 1e-3)}
    log("tica-projected (m=%d): project %.2f ms, assign %.2f ms, pipeline %.1f M frames/s"
        % (m, ms_proj, ms_asg, out["pipeline_frames_per_s"] / 1e6))
    del Y, sub, lab
    torch.cuda.empty_cache()
    return out


def kmeans_leg(args, X, sub, cpu=True, seed=12345, n_lloyd=1_000_000, iters=5, n_cpu=20_000):
    """kmeans++ seeding and Lloyd iterations on the bench data.  kmeans++:
    k centers from the subsample (each round after the first reads the n
    frames once and the D^2 state: B = (4 d + 8) n per round).  Lloyd:
    `iters` iterations with tol = 0 on the first n_lloyd frames from the
    kmeans++ centers (each iteration an assignment, 2 d k n flops, then the
    FP64 cost and the center update).  With cpu: the oracle on n_cpu frames
    of the subsample (kmeans++ with the same k and seed, one Lloyd
    iteration), scaled linearly in the frames, and whether the GPU's
    kmeans++ picks on those frames equal the oracle's."""
    from pyemma_amd import _native

    n, d = sub.shape
    k = args.k
    ms_pp = _timed(lambda: _native.kmeanspp(sub, k, seed), reps=1)
    C0, _ = _native.kmeanspp(sub, k, seed)
    Xl = X[:n_lloyd]
    res = {}

    def run_lloyd():
        res["out"] = _native.lloyd(Xl, C0.clone(), iters, 0.0)

    ms_ll = _timed(run_lloyd, reps=1)
    n_iter = int(res["out"][2].item())
    pp_bytes = (4 * d + 8) * n * (k - 1)
    f_ll = 2.0 * d * k * Xl.shape[0] * n_iter
    out = {"kmeanspp": {"frames": n, "dim": d, "k": k, "ms": ms_pp, "bytes": pp_bytes,
                        "gbs": pp_bytes / (ms_pp * 1e-3) / 1e9, "hbm_frac": pp_bytes / (ms_pp * 1e-3) / 8e12},
           "lloyd": {"frames": int(Xl.shape[0]), "dim": d, "k": k, "iterations": n_iter, "ms": ms_ll,
                     "ms_per_iteration": ms_ll / max(n_iter, 1), "assign_tflops": f_ll / (ms_ll * 1e-3) / 1e12}}
    log("kmeans++ k=%d on %d frames %.1f ms (%.0f GB/s); lloyd %d iterations on %d frames %.1f ms"
        % (k, n, ms_pp, out["kmeanspp"]["gbs"], n_iter, Xl.shape[0], ms_ll))
    if cpu:
        from oracle import native as on

        cores = len(os.sched_getaffinity(0))
        Xs_dev = sub[:n_cpu].contiguous()
        Xs = Xs_dev.cpu().numpy()
        t0 = time.perf_counter()
        Cc, idx_c = on.kmeanspp(Xs, k, seed, nthreads=cores)
        t_pp = time.perf_counter() - t0
        _, idx_g = _native.kmeanspp(Xs_dev, k, seed)
        same = bool(np.array_equal(idx_g.cpu().numpy(), idx_c))
        t0 = time.perf_counter()
        on.lloyd(Xs, Cc, 1, 0.0, nthreads=cores)
        t_ll = time.perf_counter() - t0
        out["cpu"] = {"kind": "port", "cores": cores,
                      "sample": "kmeans++ k=%d on %d frames (%.2f s); one Lloyd iteration on %d frames (%.2f s); "
                                "scaled linearly in the frames" % (k, n_cpu, t_pp, n_cpu, t_ll),
                      "kmeanspp_ms": t_pp * 1e3 * n / n_cpu,
                      "lloyd_ms": t_ll * 1e3 * Xl.shape[0] / n_cpu * max(n_iter, 1),
                      "kmeanspp_picks_identical": same}
        log("kmeans cpu (oracle, %d cores): kmeans++ %.2f s, lloyd iteration %.2f s on %d frames; picks identical %s"
            % (cores, t_pp, t_ll, n_cpu, same))
    return out


def assign_hbm_legs(n=10_000_000):
    """Assignment where it is HBM-bound by its algorithmic work (small d and k: the
    serving models' d = 64, k = 256 and the batched workload's d = 128, k = 512),
    10^7 Gaussian-mixture frames each; bytes = (4d + 4) N + 4 k d per call."""
    import torch

    from pyemma_amd import _native, autotune
    from pyemma_amd.synthetic import gaussian_mixture

    out = {}
    for d, k in ((64, 256), (128, 512)):
        X = gaussian_mixture(n, d, seed=60 + d, device=_native.device())
        rng = np.random.default_rng(d)
        C = X[torch.as_tensor(np.sort(rng.choice(n, size=k, replace=False)), device=X.device)].contiguous()
        variant = autotune.tune_assign(X, C)
        lab = torch.empty(n, dtype=torch.int32, device=X.device)
        nre = torch.zeros(1, dtype=torch.int64, device=X.device)
        ms = _timed(lambda: _native.assign(X, C, labels=lab, n_rechecked=nre))
        nbytes = (4 * d + 4) * n + 4 * k * d
        gbs = nbytes / (ms * 1e-3) / 1e9
        out["assign_d%d_k%d" % (d, k)] = {"ms": ms, "bytes": nbytes, "gbs": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
                                          "tile_variant": variant, "rechecked_frames": int(nre.item()),
                                          "algorithmic_tflops": 2.0 * d * k * n / (ms * 1e-3) / 1e12}
        log("assign d=%d k=%d: %.2f ms, %.0f GB/s (%.0f %% of HBM)" % (d, k, ms, gbs, 100 * gbs / PEAK_HBM_GBS))
        del X, C, lab
        torch.cuda.empty_cache()
    return out


def large_k_leg(args, n=10_000_000, d=512, k=4096, n_seed=200_000, iters=5, n_plain=1_000_000, n_cpu=20_000):
    """k-means at the largest configured shape: 10^7 x 512 Gaussian-mixture
    frames (64 components, means N(0, 4 I), diagonal variances log-uniform in
    [0.1, 2]) generated on the device, kmeans++ with k = 4096 on a seeded
    n_seed-frame subsample, the assignment of all frames to those centers,
    and `iters` Lloyd iterations (tol 0) on all frames.  Parity: the first
    n_plain headline labels against the plain FP32 loop on the device, and
    n_cpu frames spread over the data against the oracle on the host."""
    import torch

    from pyemma_amd import _native, autotune
    from pyemma_amd.synthetic import gaussian_mixture

    t0 = time.time()
    X = gaussian_mixture(n, d, seed=3, device=_native.device())
    torch.cuda.synchronize()
    log("large-k data %d x %d generated in %.1f s" % (n, d, time.time() - t0))
    rng = np.random.default_rng(31)
    sub = X[torch.as_tensor(np.sort(rng.choice(n, size=n_seed, replace=False)), device=X.device)].contiguous()
    res = {}

    def seed():
        res["C"], _ = _native.kmeanspp(sub, k, 4242)

    ms_pp = _timed(seed, reps=1)
    C = res["C"]
    variant = autotune.tune_assign(X, C)
    lab = torch.empty(n, dtype=torch.int32, device=X.device)
    nre = torch.zeros(1, dtype=torch.int64, device=X.device)
    ms_asg = _timed(lambda: _native.assign(X, C, labels=lab, n_rechecked=nre))
    n_re = int(nre.item())

    def run_lloyd():
        res["out"] = _native.lloyd(X, C.clone(), iters, 0.0)

    ms_ll = _timed(run_lloyd, reps=1)
    n_iter = int(res["out"][2].item())
    f_asg = 2.0 * d * k * n
    ex_asg = asg_executed_flops(d, k, n, asg_first_pass_products(d, variant))
    ach = f_asg / (ms_asg * 1e-3) / 1e12
    pp_bytes = (4 * d + 8) * n_seed * (k - 1)
    out = {"workload": "%d x %d f32 Gaussian mixture, k=%d" % (n, d, k),
           "kmeanspp": {"frames": n_seed, "k": k, "ms": ms_pp, "bytes": pp_bytes,
                        "gbs": pp_bytes / (ms_pp * 1e-3) / 1e9},
           "assign": {"ms": ms_asg, "algorithmic_tflop": f_asg / 1e12, "tflops": ach,
                      "executed_tflop": ex_asg / 1e12, "rechecked_frames": n_re, "tile_variant": variant,
                      "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                                   "frac": ach / PEAK_BF16_TFLOPS,
                                   "executed_frac": ex_asg / (ms_asg * 1e-3) / 1e12 / PEAK_BF16_TFLOPS}},
           "lloyd": {"frames": n, "iterations": n_iter, "ms": ms_ll, "ms_per_iteration": ms_ll / max(n_iter, 1)}}
    log("large-k: kmeans++ k=%d on %d frames %.0f ms; assign %.1f ms (%.0f TFLOP/s algorithmic, %d rechecked); "
        "lloyd %d iterations %.0f ms" % (k, n_seed, ms_pp, ms_asg, ach, n_re, n_iter, ms_ll))
    # parity of the headline labels
    lab_plain = torch.empty(n_plain, dtype=torch.int32, device=X.device)
    Xp = X[:n_plain]
    ms_plain = _timed(lambda: _native.assign(Xp, C, mode=_native.EMMA_ASSIGN_FP32_PLAIN, labels=lab_plain), reps=1)
    mism = int((lab[:n_plain] != lab_plain).sum().item())
    out["parity"] = {"label_mismatches_vs_plain_gpu": mism, "label_frames_vs_plain_gpu": n_plain,
                     "plain_gpu_ms_per_1e6_frames": ms_plain * 1e6 / n_plain}
    if not args.no_cpu:
        from oracle import native as on

        idx = torch.as_tensor(np.sort(rng.choice(n, size=n_cpu, replace=False)), device=X.device)
        cores = len(os.sched_getaffinity(0))
        t1 = time.perf_counter()
        ref = on.assign(X[idx].cpu().numpy(), C.cpu().numpy(), nthreads=cores)
        t_cpu = time.perf_counter() - t1
        cm = int(np.count_nonzero(lab[idx].cpu().numpy() != ref))
        out["parity"].update({"label_mismatches_vs_oracle": cm, "label_frames_vs_oracle": n_cpu})
        out["cpu"] = {"kind": "port", "cores": cores, "assign_ms_scaled": t_cpu * 1e3 * n / n_cpu,
                      "sample": "assign of %d frames x k=%d (C oracle, %d threads, %.2f s), scaled to %d frames"
                                % (n_cpu, k, cores, t_cpu, n)}
    log("large-k parity: %s" % out["parity"])
    del X, sub, lab, lab_plain, res
    torch.cuda.empty_cache()
    return out


def batched_leg(args, ntraj=4096, d=128, lag=5, k=512, n_seed=1_000_000, n_plain=2_000_000, reps=3):
    """The batched workload at full size: 4096 linear-Langevin trajectories with
    log-uniform lengths 1k - 200k frames (about 1.5e8 frames x 128 float32,
    about 79 GB, generated on the device), TICA lag 5 moments and the
    assignment of every frame to k = 512 centers (kmeans++ on a seeded 1M-frame
    subsample), run as one captured HIP graph (BatchedDiscretizer).  Reports
    the replay time, the per-call times of the eager pass, the rooflines
    (moments: executed bf16 MFMA rate; assignment: HBM bytes (4 d + 4) N + 4 k d
    against 8 TB/s) and parity: graph replay == eager bit for bit, labels vs
    the plain FP32 loop on the first n_plain frames, every TICA eigenvalue of
    the default mode vs the FP64-tile moments of the whole data."""
    import torch

    from oracle import eig as oeig
    from pyemma_amd import _native, autotune
    from pyemma_amd.batch import BatchedDiscretizer
    from pyemma_amd.synthetic import langevin_trajectories

    dev = _native.device()
    t0 = time.time()
    X, off = langevin_trajectories(ntraj, d, lmin=1000, lmax=200000, seed=4, device=dev)
    torch.cuda.synchronize()
    N = int(off[-1])
    log("batched data: %d trajectories, %d frames x %d generated in %.1f s" % (ntraj, N, d, time.time() - t0))
    rng = np.random.default_rng(41)
    sub = X[torch.as_tensor(np.sort(rng.choice(N, size=n_seed, replace=False)), device=dev)].contiguous()
    C, _ = _native.kmeanspp(sub, k, 4343)
    del sub
    b = BatchedDiscretizer(X, off, lag, C)
    w0, mean0, C00_0, C0t_0, lab0 = b.run_eager()  # eager pass (also tunes the tile shapes)
    b.capture()  # the graph's outputs are filled by its replays
    # per-call times of the eager sequence
    off_dev = torch.as_tensor(np.asarray(off, dtype=np.int64), device=dev)
    ms_mom = _timed(lambda: _native.moments_lagged(X, off_dev, lag, reversible=True, remove_mean=True, c0t=True,
                                                   precision=_native.EMMA_PREC_BF16X4_ACC), reps=1)
    nre = torch.zeros(1, dtype=torch.int64, device=dev)
    lab = torch.empty(N, dtype=torch.int32, device=dev)
    ms_asg = _timed(lambda: _native.assign(X, C, labels=lab, n_rechecked=nre), reps=1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.replay()
    e1.record()
    e1.synchronize()
    ms_graph = e0.elapsed_time(e1) / reps
    w1, mean1, C00_1, C0t_1, lab1 = b.results()
    same = bool(w1 == w0 and np.array_equal(mean1, mean0) and np.array_equal(C00_1, C00_0)
                and np.array_equal(C0t_1, C0t_0) and np.array_equal(lab1, lab0))
    P = int(sum(max(0, int(e) - int(a) - lag) for a, e in zip(off[:-1], off[1:])))
    ex_mom = mom_executed_flops(d, P)
    asg_bytes = (4 * d + 4) * N + 4 * k * d
    out = {"workload": "%d Langevin trajectories, lengths log-uniform [1e3, 2e5], %d frames x %d f32, lag %d, k=%d"
                       % (ntraj, N, d, lag, k),
           "frames": N, "lagged_pairs": P, "graph_replay_ms": ms_graph, "frames_per_s": N / (ms_graph * 1e-3),
           "moments": {"ms": ms_mom, "executed_tflop": ex_mom / 1e12,
                       "executed_frac": ex_mom / (ms_mom * 1e-3) / 1e12 / PEAK_BF16_TFLOPS,
                       "algorithmic_tflop": 6.0 * d * d * P / 1e12},
           "assign": {"ms": ms_asg, "bytes": asg_bytes, "gbs": asg_bytes / (ms_asg * 1e-3) / 1e9,
                      "hbm_frac": asg_bytes / (ms_asg * 1e-3) / 1e9 / PEAK_HBM_GBS,
                      "rechecked_frames": int(nre.item()),
                      "executed_frac": asg_executed_flops(d, k, N, asg_first_pass_products(
                          d, autotune.tuned().get(autotune._bucket(d, k), (None,))[0])) / (ms_asg * 1e-3) / 1e12
                                       / PEAK_BF16_TFLOPS}}
    # parity
    lab_plain = torch.empty(n_plain, dtype=torch.int32, device=dev)
    _native.assign(X[:n_plain], C, mode=_native.EMMA_ASSIGN_FP32_PLAIN, labels=lab_plain)
    mism = int(np.count_nonzero(lab_plain.cpu().numpy() != lab0[:n_plain]))
    w, _, _, Mxx, Mxy, _ = _native.moments_lagged(X, off_dev, lag, reversible=True, remove_mean=True, c0t=True,
                                                  precision=_native.EMMA_PREC_FP64)
    wh = w.item()
    ev64 = oeig.generalized_eigenvalues(Mxx.cpu().numpy() / wh, Mxy.cpu().numpy() / wh)
    ev = oeig.generalized_eigenvalues(C00_0, C0t_0)
    big = np.abs(ev64) >= 0.1
    out["parity"] = {"graph_replay_equals_eager": same, "label_mismatches_vs_plain_gpu": mism,
                     "label_frames_vs_plain_gpu": n_plain,
                     "eig_rel_err_max_vs_fp64_tiles": float(np.max(np.abs(ev - ev64) / np.abs(ev64))),
                     "eig_rel_err_max_abs_ge_0.1_vs_fp64_tiles": float(np.max(np.abs(ev - ev64)[big] / np.abs(ev64[big])))
                     if big.any() else None,
                     "eig_abs_err_max_vs_fp64_tiles": float(np.max(np.abs(ev - ev64))),
                     "eig_count": int(len(ev)), "eig_count_abs_ge_0.1": int(big.sum())}
    log("batched: graph replay %.1f ms (%.3g frames/s); moments %.1f ms, assign %.1f ms (%.0f GB/s); parity %s"
        % (ms_graph, out["frames_per_s"], ms_mom, ms_asg, out["assign"]["gbs"], out["parity"]))
    del X, b, lab, lab_plain, C
    torch.cuda.empty_cache()
    return out


def serving_leg(args):
    """Poisson arrivals of discretization requests against 8 resident models.

    Models: d cycles over {64, 128, 300, 512}, k over {256, 1000, 4096},
    projected dimension m = min(d, 32); request lengths log-uniform in
    [1k, min(1M, pool)] frames, data windows of a pinned host pool per d
    (Gaussian-mixture frames generated on the device).  Latency is the time
    from the scheduled arrival to the host seeing the labels.
    """
    import torch

    from pyemma_amd.serving import DiscretizationServer, ResidentModel, latency_summary
    from pyemma_amd.synthetic import gaussian_mixture

    dev = torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(5)
    dims, ks = [64, 128, 300, 512], [256, 1000, 4096]
    pool_n = int(args.serve_pool)
    pools, models = {}, {}
    t0 = time.time()
    for i, d in enumerate(dims):
        Xd = gaussian_mixture(pool_n, d, seed=50 + i, device=dev)
        host = torch.empty((pool_n, d), dtype=torch.float32, pin_memory=True)
        host.copy_(Xd)
        pools[d] = (host, Xd[: min(pool_n, 20000)].double().mean(0).cpu().numpy())
        del Xd
    torch.cuda.synchronize()
    for mi in range(8):
        d, k = dims[mi % 4], ks[mi % 3]
        m = min(d, 32)
        host, mean = pools[d]
        V, _ = np.linalg.qr(rng.standard_normal((d, m)))
        rows = rng.choice(pool_n, size=k, replace=False)
        C = ((host[torch.as_tensor(np.sort(rows))].double().numpy() - mean) @ V).astype(np.float32)
        models[mi] = (d, ResidentModel(mean, V, C))
    log("serving: pools and 8 models ready in %.1f s" % (time.time() - t0))
    budget = 48 << 30
    server = DiscretizationServer(budget, n_streams=8)
    for key, (_, mdl) in models.items():
        server.add_model(key, mdl)
    # warm-up: one small request per model
    for key, (d, _) in models.items():
        server.submit(key, pools[d][0][:4096])
    server.run_until_idle()
    server.done.clear()
    n_req = max(1, int(rng.poisson(args.serve_rate * args.serve_seconds)))
    arrivals = np.sort(rng.uniform(0.0, args.serve_seconds, size=n_req))
    lens = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(min(1e6, pool_n)), size=n_req))).astype(np.int64)
    sched = []
    for a, T in zip(arrivals, lens):
        key = int(rng.integers(8))
        d = models[key][0]
        start = int(rng.integers(0, pool_n - T + 1))
        sched.append((float(a), key, pools[d][0][start:start + T]))
    t1 = time.perf_counter()
    reqs = server.replay(sched)
    wall = time.perf_counter() - t1
    out = latency_summary(reqs)
    out.update({"offered_req_per_s": args.serve_rate, "arrival_window_s": args.serve_seconds, "wall_s": wall,
                "max_in_flight": server.max_running, "hbm_budget_gb": budget / 2 ** 30,
                "peak_arena_gb": server.alloc.peak / 2 ** 30,
                "offered_gb_per_s": float(sum(r.T * models[r.model_key][0] * 4 for r in reqs)) / 1e9
                                    / args.serve_seconds,
                "models": "8 resident: d in {64,128,300,512}, k in {256,1000,4096}, m = min(d, 32)",
                "requests_lengths": "log-uniform [1e3, %d] frames" % min(10 ** 6, pool_n)})
    log("serving: %d requests, p50 %.2f ms, p99 %.2f ms, %.3g frames/s, max in flight %d"
        % (out["requests"], out["p50_ms"], out["p99_ms"], out["frames_per_s"], server.max_running))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, X, offsets, C, labels_gpu, mats, per_traj=500):
    """The oracle (CPU restatement of the reference path) beside the GPU, and
    the parity of the headline results against it.

    Moments: the FP64 NumPy restatement over every trajectory of the
    workload (not a sample), timed whole; its C00 and C0t are the FP64
    reference for the eigenvalues of every GPU precision mode.  Assignment:
    the C oracle with all host cores on `per_traj` frames drawn from each
    trajectory (seeded), scaled to the N frames; those frames' headline GPU
    labels are compared with it.  One-thread rates: the moments of one
    trajectory with BLAS limited to one thread, the assignment of 2,000
    frames with one OpenMP thread."""
    import torch
    from threadpoolctl import threadpool_limits

    from oracle import eig as oeig
    from oracle import moments as om
    from oracle import native as on
    from pyemma_amd import _solvers

    cores = len(os.sched_getaffinity(0))
    N, d = X.shape
    ntraj = len(offsets) - 1
    t0 = time.perf_counter()
    trajs = [X[offsets[i]:offsets[i + 1]].cpu().numpy() for i in range(ntraj)]
    t_copy = time.perf_counter() - t0
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        ref = om.covariance_lagged(trajs, args.lag, remove_data_mean=True, reversible=True, bessel=False)
        t_mom = time.perf_counter() - t0
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        om.covariance_lagged(trajs[:1], args.lag, remove_data_mean=True, reversible=True, bessel=False)
        t_mom1 = (time.perf_counter() - t0) * N / len(trajs[0])
    t0 = time.perf_counter()
    l_ref, _ = _solvers.eig_corr(ref["C00"], ref["C0t"], 1e-6, sign_maxelement=True)
    t_eig = time.perf_counter() - t0
    log("oracle FP64 moments of all %d trajectories: %.1f s (host copy %.1f s)" % (ntraj, t_mom, t_copy))
    # assignment sample: per_traj frames of every trajectory
    rng = np.random.default_rng(11)
    idx = np.concatenate([np.sort(rng.choice(offsets[i + 1] - offsets[i], size=min(per_traj, offsets[i + 1]
                                                                                  - offsets[i]), replace=False))
                          + offsets[i] for i in range(ntraj)]).astype(np.int64)
    idx_dev = torch.as_tensor(idx, device=X.device)
    Xa = X[idx_dev].cpu().numpy()
    Ch = C.cpu().numpy()
    t0 = time.perf_counter()
    lab_ref = on.assign(Xa, Ch, nthreads=cores)
    t_asg = time.perf_counter() - t0
    n1 = 2000
    t0 = time.perf_counter()
    on.assign(Xa[:n1], Ch, nthreads=1)
    t_asg1 = time.perf_counter() - t0
    na = len(idx)
    t_full = t_mom + t_eig + t_asg * N / na
    t_full1 = t_mom1 + t_eig + t_asg1 * N / n1
    cpu = {"value": N / t_full, "unit": "frames/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
           "sample": "moments: all %d trajectories (%d frames, numpy FP64 dgemm restatement, %d BLAS threads, "
                     "%.1f s, measured whole); eig d=%d (%.3f s); assign: %d frames (%d per trajectory) x k=%d "
                     "(C oracle, OpenMP %d threads, %.2f s), scaled to %d frames"
                     % (ntraj, N, cores, t_mom, d, t_eig, na, per_traj, Ch.shape[0], cores, t_asg, N),
           "one_thread": {"value": N / t_full1, "unit": "frames/s",
                          "sample": "moments of 1 trajectory with 1 BLAS thread, assign of %d frames with 1 "
                                    "thread, scaled to %d frames" % (n1, N)}}
    lab_gpu = labels_gpu[idx_dev].cpu().numpy()
    mism = int(np.count_nonzero(lab_gpu != lab_ref))
    parity = {"labels_identical_vs_oracle": mism == 0, "label_mismatches_vs_oracle": mism,
              "label_frames_vs_oracle": na, "eig_frames": N,
              "eig_reference": "FP64 oracle moments of all %d trajectories" % ntraj}
    # every eigenvalue of the generalized problem, and every eigenvalue TICA
    # returns (eig_corr with epsilon 1e-6: the same count is required)
    ev_ref = oeig.generalized_eigenvalues(ref["C00"], ref["C0t"])
    parity["eig_count"] = int(len(ev_ref))
    parity["eig_returned_count_oracle"] = int(len(l_ref))
    for name, (C00, C0t) in mats.items():
        ev = oeig.generalized_eigenvalues(C00, C0t)
        rel = np.abs(ev - ev_ref) / np.abs(ev_ref)
        parity["eig_%s_rel_err_max" % name] = float(rel.max())
        parity["eig_%s_abs_err_max" % name] = float(np.abs(ev - ev_ref).max())
        l_gpu, _ = _solvers.eig_corr(C00, C0t, 1e-6, sign_maxelement=True)
        if len(l_gpu) == len(l_ref):
            parity["eig_%s_returned_rel_err_max" % name] = float(np.max(np.abs(l_gpu - l_ref) / np.abs(l_ref)))
        else:
            parity["eig_%s_returned_count" % name] = int(len(l_gpu))
    log("cpu baseline %.3g frames/s on %d cores; parity %s" % (cpu["value"], cores, parity))
    del trajs
    return cpu, parity


if __name__ == "__main__":
    main()
