"""Seeded synthetic trajectories of the benchmark workloads.

hmm_trajectory       4-state hidden Markov chain, Gaussian emissions (100k x 30)
brownian_distances   pairwise distances of Brownian particles in a 3-basin bias
gaussian_mixture     mixture of Gaussians with anisotropic diagonal variances
langevin_trajectories many linear Langevin trajectories of log-uniform length

Large sets are generated on the GPU with torch (plumbing only; seeded torch
generators), small ones on the host.  Every generator is deterministic for a
given seed and device type.
"""
import math

import numpy as np


def _torch():
    import torch
    return torch


def hmm_trajectory(n, d, seed=1, nstates=4, p_stay=0.99, mean_scale=5.0):
    """Hidden chain with P_ii = p_stay and uniform off-diagonal moves; x_t = mu_{s_t} + N(0, I).

    mu_s ~ N(0, mean_scale^2 I).  Returns (n, d) float32.
    """
    rng = np.random.default_rng(seed)
    mu = rng.normal(0.0, mean_scale, size=(nstates, d))
    s0 = rng.integers(nstates)
    switch = rng.random(n) < (1.0 - p_stay)
    switch[0] = False
    inc = rng.integers(1, nstates, size=n)
    states = (s0 + np.cumsum(np.where(switch, inc, 0))) % nstates
    X = mu[states] + rng.normal(size=(n, d))
    return X.astype(np.float32)


def brownian_distances(ntraj, nframes, nparticles=25, sigma=0.05, kappa_dt=0.01, burn_in=1000, seed=2,
                       device=None, basin_scale=1.5, return_positions=False):
    """Pairwise distances (i < j) of particles doing Euler-Maruyama Brownian motion
    under U(r) = kappa/2 min_b |r - R_b|^2 (3 basins R_b ~ N(0, basin_scale^2 I)),
    step r <- r - kappa_dt (r - R_b*) + sigma xi.  Distances in FP64, stored as
    float32.  Returns (X (ntraj*nframes, P) tensor, offsets ndarray), and with
    return_positions the float32 positions (ntraj*nframes, nparticles, 3) as well."""
    t = _torch()
    dev = device or ("cuda" if t.cuda.is_available() else "cpu")
    g = t.Generator(device=dev).manual_seed(seed)
    R = t.randn((3, 3), generator=g, device=dev, dtype=t.float64) * basin_scale
    start = t.randint(0, 3, (ntraj, nparticles), generator=g, device=dev)
    r = R[start] + 0.3 * t.randn((ntraj, nparticles, 3), generator=g, device=dev, dtype=t.float64)
    iu, ju = np.triu_indices(nparticles, k=1)
    iu = t.as_tensor(iu, device=dev)
    ju = t.as_tensor(ju, device=dev)
    P = len(iu)
    out = t.empty((ntraj, nframes, P), dtype=t.float32, device=dev)
    xyz = t.empty((ntraj, nframes, nparticles, 3), dtype=t.float32, device=dev) if return_positions else None

    def step(r, noise):
        diff = r[:, :, None, :] - R[None, None, :, :]          # (ntraj, np, 3, 3)
        b = (diff * diff).sum(-1).argmin(-1)                  # nearest basin
        return r - kappa_dt * (r - R[b]) + sigma * noise

    blk = 256
    for t0 in range(0, burn_in, blk):
        nb = min(blk, burn_in - t0)
        noise = t.randn((nb, ntraj, nparticles, 3), generator=g, device=dev, dtype=t.float64)
        for q in range(nb):
            r = step(r, noise[q])
    pos = t.empty((blk, ntraj, nparticles, 3), dtype=t.float64, device=dev)
    for t0 in range(0, nframes, blk):
        nb = min(blk, nframes - t0)
        noise = t.randn((nb, ntraj, nparticles, 3), generator=g, device=dev, dtype=t.float64)
        for q in range(nb):
            r = step(r, noise[q])
            pos[q] = r
        p = pos[:nb]
        dist = (p[:, :, iu, :] - p[:, :, ju, :]).norm(dim=-1)  # (nb, ntraj, P)
        out[:, t0:t0 + nb, :] = dist.permute(1, 0, 2).to(t.float32)
        if xyz is not None:
            xyz[:, t0:t0 + nb] = p.permute(1, 0, 2, 3).to(t.float32)
    offsets = np.arange(ntraj + 1, dtype=np.int64) * nframes
    if return_positions:
        return out.reshape(ntraj * nframes, P), offsets, xyz.reshape(ntraj * nframes, nparticles, 3)
    return out.reshape(ntraj * nframes, P), offsets


def gaussian_mixture(n, d, ncomp=64, mean_scale=2.0, var_lo=0.1, var_hi=2.0, seed=3, device=None, chunk=1 << 20):
    """x = mu_c + sqrt(v_c) * N(0, I), c uniform over ncomp components,
    mu_c ~ N(0, mean_scale^2 I), v_c log-uniform in [var_lo, var_hi] per dimension.
    Returns an (n, d) float32 tensor."""
    t = _torch()
    dev = device or ("cuda" if t.cuda.is_available() else "cpu")
    g = t.Generator(device=dev).manual_seed(seed)
    mu = t.randn((ncomp, d), generator=g, device=dev) * mean_scale
    lv = math.log(var_lo) + (math.log(var_hi) - math.log(var_lo)) * t.rand((ncomp, d), generator=g, device=dev)
    sd = t.exp(0.5 * lv)
    X = t.empty((n, d), dtype=t.float32, device=dev)
    for lo in range(0, n, chunk):
        m = min(chunk, n - lo)
        c = t.randint(0, ncomp, (m,), generator=g, device=dev)
        X[lo:lo + m] = mu[c] + sd[c] * t.randn((m, d), generator=g, device=dev)
    return X


def log_uniform_lengths(ntraj, lmin, lmax, seed):
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lmin), math.log(lmax), size=ntraj))).astype(np.int64)


def langevin_trajectories(ntraj, d, lmin=1000, lmax=200000, seed=4, device=None, lengths=None):
    """x_{t+1} = (I + A) x_t + xi, A = Q diag(l) Q^T, l ~ U[-1, -0.01], Q Haar,
    xi ~ N(0, I), started from the stationary distribution.  Trajectory
    lengths are floor(exp(U[ln lmin, ln lmax])).  Returns (X tensor (sum T_i, d),
    offsets ndarray)."""
    t = _torch()
    dev = device or ("cuda" if t.cuda.is_available() else "cpu")
    if lengths is None:
        lengths = log_uniform_lengths(ntraj, lmin, lmax, seed)
    lengths = np.asarray(lengths, dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    g = t.Generator(device=dev).manual_seed(seed)
    Z = t.randn((d, d), generator=g, device=dev, dtype=t.float64)
    Q, Rq = t.linalg.qr(Z)
    Q = Q * t.sign(t.diagonal(Rq))[None, :]
    lam = -1.0 + 0.99 * t.rand(d, generator=g, device=dev, dtype=t.float64)
    M = (Q * (1.0 + lam)[None, :]) @ Q.T                       # I + A
    sd_stat = 1.0 / t.sqrt(1.0 - (1.0 + lam) ** 2)              # stationary std in the eigenbasis
    x = (t.randn((ntraj, d), generator=g, device=dev, dtype=t.float64) * sd_stat[None, :]) @ Q.T
    X = t.empty((int(offsets[-1]), d), dtype=t.float32, device=dev)
    off_t = t.as_tensor(offsets[:-1], device=dev)
    len_t = t.as_tensor(lengths, device=dev)
    order = np.argsort(-lengths, kind="stable")
    Tmax = int(lengths.max())
    # trajectories sorted by length: the active ones at step s are a prefix
    ord_t = t.as_tensor(order, device=dev)
    x = x[ord_t]
    lens_sorted = lengths[order]
    neg = -lens_sorted
    for s in range(Tmax):
        nact = int(np.searchsorted(neg, -s, side="left"))  # number of lengths > s
        if nact == 0:
            break
        rows = off_t[ord_t[:nact]] + s
        X[rows] = x[:nact].to(t.float32)
        x = x[:nact] @ M.T + t.randn((nact, d), generator=g, device=dev, dtype=t.float64)
    return X, offsets
