"""The C-ABI library loads, exports every symbol include/emma_amd.h declares,
and rejects bad arguments on the host (no GPU work is issued by these calls)."""
import ctypes
import os
import re

import pytest

from pyemma_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "emma_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emma_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("emma_moments_lagged", "emma_assign", "emma_kmeanspp", "emma_lloyd", "emma_project",
                     "emma_count_nonfinite", "emma_last_error", "emma_version"):
        assert required in names


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_version_and_workspace_sizes():
    L = _native.lib()
    assert b"gfx950" in L.emma_version()
    assert L.emma_moments_workspace_size(1000, 30, 1, 7, 1) > 0
    assert L.emma_assign_workspace_size(1000, 300, 1000, 1) > L.emma_assign_workspace_size(1000, 300, 1000, 0)
    assert L.emma_kmeanspp_workspace_size(1000, 30, 10) > 0
    assert L.emma_lloyd_workspace_size(1000, 30, 10, 1) > 0
    # workspace grows with the number of frames for the recheck list
    assert L.emma_assign_workspace_size(10 ** 6, 30, 10, 1) > L.emma_assign_workspace_size(10 ** 3, 30, 10, 1)


@pytest.mark.parametrize("call,expect", [
    ("assign_k0", "bad sizes"),
    ("assign_mode", "unknown mode"),
    ("moments_ctt_sym", "meaningless"),
    ("moments_lag0_c0t", "lag must be positive"),
    ("moments_prec", "unknown precision"),
    ("kmpp_k0", "bad sizes"),
    ("lloyd_mode", "unknown assign mode"),
    ("project_sizes", "bad sizes"),
    ("nonfinite_neg", "n < 0"),
    ("counts_lag0", "bad sizes"),
    ("counts_big", "too large"),
    ("maxstate_null", "bad arguments"),
])
def test_invalid_arguments_rejected_on_host(call, expect):
    L = _native.lib()
    p = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    z = ctypes.c_void_p(0)
    if call == "assign_k0":
        rc = L.emma_assign(p, 10, 3, p, 0, p, 0, z, p, 1 << 20, z)
    elif call == "assign_mode":
        rc = L.emma_assign(p, 10, 3, p, 2, p, 7, z, p, 1 << 20, z)
    elif call == "moments_ctt_sym":
        rc = L.emma_moments_lagged(p, 10, 3, p, 1, 1, 1, 0, 1 | 8, 1, p, p, p, p, p, p, p, 1 << 20, z)
    elif call == "moments_lag0_c0t":
        rc = L.emma_moments_lagged(p, 10, 3, p, 1, 0, 1, 0, 4, 1, p, p, p, p, p, p, p, 1 << 20, z)
    elif call == "moments_prec":
        rc = L.emma_moments_lagged(p, 10, 3, p, 1, 1, 1, 0, 3, 9, p, p, p, p, p, p, p, 1 << 20, z)
    elif call == "kmpp_k0":
        rc = L.emma_kmeanspp(p, 10, 3, 0, 1, p, p, p, 1 << 20, z)
    elif call == "lloyd_mode":
        rc = L.emma_lloyd(p, 10, 3, p, 2, 5, 0.0, 9, p, p, p, p, p, 1 << 20, z)
    elif call == "project_sizes":
        rc = L.emma_project(p, 10, 0, p, p, 1, 1, p, z)
    elif call == "counts_lag0":
        rc = L.emma_count_transitions(p, 10, p, 1, 0, 1, 5, p, p, p, z)
    elif call == "counts_big":
        rc = L.emma_count_transitions(p, 10, p, 1, 1, 1, 50000, p, p, p, z)
    elif call == "maxstate_null":
        rc = L.emma_max_state(p, 10, z, z)
    else:
        rc = L.emma_count_nonfinite(p, -1, p, z)
    assert rc == -1
    assert expect in L.emma_last_error().decode()


def test_workspace_too_small_rejected():
    L = _native.lib()
    p = ctypes.c_void_p(16)
    rc = L.emma_assign(p, 1000, 300, p, 1000, p, 1, ctypes.c_void_p(0), p, 8, ctypes.c_void_p(0))
    assert rc == -1 and b"workspace too small" in L.emma_last_error()


def test_api_fails_loudly_without_gpu():
    """No CPU fallback: without a visible GPU the product path raises."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import numpy as np
    from pyemma_amd.coordinates import tica
    with pytest.raises(_native.EmmaError):
        tica(np.random.rand(100, 3).astype(np.float32), lag=2)


def test_assign_variant_table():
    """Variant queries and selection are host-only calls: two first-pass
    pipelines (coarse fp16 + fine, or fine only) times the tile shapes."""
    L = _native.lib()
    assert L.emma_assign_variants(32) == 12
    assert L.emma_assign_variants(128) == 6
    assert L.emma_assign_variants(300) == 4
    assert L.emma_assign_variants(512) == 2
    assert L.emma_assign_variants(513) == 0
    assert L.emma_assign_select(300, 1000, 5) == -1
    assert L.emma_assign_select(600, 1000, 0) == -1
    assert L.emma_assign_select(300, 1000, 1) == 0
    assert L.emma_assign_select(300, 1000, -1) == 0
