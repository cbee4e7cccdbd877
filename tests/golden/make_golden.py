"""Regenerate the committed fixtures under tests/golden/.

koopman_states.npz
    The input trajectories of the reference's own TICA/Koopman tests
    (pyemma/coordinates/tests/data/test_data_koopman.npz: 1000 integer state
    trajectories of 200 frames, states 0..19, read with allow_pickle=False),
    stored as one (1000, 200) uint8 array in the file's key order.  Inputs
    only: the reference holds no expected outputs for this path.
hmm_small.npz
    A small seeded 4-state HMM trajectory (2000 x 30 float32, the config-1
    generator) with the FP64 oracle's TICA moments and eigenvalues at lag 10,
    the oracle's kmeans++ picks (k=16, seed 42), 5 Lloyd iterations from them
    and the sequential-FP32 labels.  These pin the oracle between rounds.
oom_counts.npz
    The discrete trajectories and the reference count tensor of the
    reference's OOM-MSM tests (pyemma/msm/tests/data/TestData_OOM_MSM.npz,
    read with allow_pickle=False; used at msm/tests/test_oom_msm.py:94-120):
    1000 trajectories of 200 frames over 5 states (uint8, in key order
    arr_0 .. arr_999) and C2t (5, 5, 5).  C2t summed over its last axis is
    the sliding-window count matrix at lag 10 of those trajectories.

Run from the repository root: python tests/golden/make_golden.py
(the Koopman file is only re-read when /root/reference is present).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import eig as oeig  # noqa: E402
from oracle import moments as om  # noqa: E402
from oracle import native as on  # noqa: E402
from pyemma_amd.synthetic import hmm_trajectory  # noqa: E402

REF_KOOPMAN = "/root/reference/pyemma/coordinates/tests/data/test_data_koopman.npz"
REF_OOM = "/root/reference/pyemma/msm/tests/data/TestData_OOM_MSM.npz"


def make_koopman():
    if not os.path.exists(REF_KOOPMAN):
        print("reference fixture not present; keeping", os.path.join(HERE, "koopman_states.npz"))
        return
    f = np.load(REF_KOOPMAN, allow_pickle=False)
    keys = list(f.keys())
    states = np.stack([f[k] for k in keys]).astype(np.uint8)
    assert states.shape == (1000, 200) and states.max() < 20
    np.savez_compressed(os.path.join(HERE, "koopman_states.npz"), states=states,
                        keys=np.array(keys))


def make_oom_counts():
    if not os.path.exists(REF_OOM):
        print("reference fixture not present; keeping", os.path.join(HERE, "oom_counts.npz"))
        return
    f = np.load(REF_OOM, allow_pickle=False)
    dtrajs = np.stack([f["arr_%d" % k] for k in range(1000)])
    assert dtrajs.shape == (1000, 200) and dtrajs.min() >= 0 and dtrajs.max() < 5
    np.savez_compressed(os.path.join(HERE, "oom_counts.npz"), dtrajs=dtrajs.astype(np.uint8), C2t=f["C2t"])


def make_hmm():
    X = hmm_trajectory(2000, 30, seed=11)
    lag = 10
    m = om.covariance_lagged([X], lag, c00=True, c0t=True, remove_data_mean=True, reversible=True, bessel=False)
    ev = oeig.generalized_eigenvalues(m["C00"], m["C0t"])
    C0, idx = on.kmeanspp(X, 16, 42)
    C5, labels5, inert5, conv5 = on.lloyd(X, C0, 5, 0.0)
    labels = on.assign(X, C5)
    np.savez_compressed(os.path.join(HERE, "hmm_small.npz"), X=X, lag=lag, mean=m["mean"], C00=m["C00"],
                        C0t=m["C0t"], eigenvalues=ev, kmpp_idx=idx, kmpp_centers=C0, lloyd_centers=C5,
                        lloyd_inertias=inert5, labels=labels)


if __name__ == "__main__":
    make_koopman()
    make_oom_counts()
    make_hmm()
    print("ok")
