import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libemma_amd.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def golden_hmm():
    return dict(np.load(os.path.join(GOLDEN, "hmm_small.npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def koopman_data():
    """Input trajectories of the reference's Koopman/TICA tests, one-hot mapped to
    10 features like test_tica.py:443-452 (chi[2n:2n+2, n] = 1)."""
    f = np.load(os.path.join(GOLDEN, "koopman_states.npz"), allow_pickle=False)
    states = f["states"].astype(np.int64)
    nf = 10
    chi = np.zeros((20, nf), dtype=float)
    for n in range(nf):
        chi[2 * n:2 * (n + 1), n] = 1.0
    return [chi[s, :] for s in states]
