"""MI355X-native implementation of PyEMMA's coordinate-discretization hot path.

TICA lagged moments, kmeans++ / Lloyd and assignment to centers run as
hand-written HIP kernels for gfx950 (libemma_amd.so, C ABI in
include/emma_amd.h); the host code mirrors pyemma.coordinates.
"""
from . import coordinates
from ._native import version
from .exceptions import InvalidDataInStreamException, ZeroRankError

__all__ = ["coordinates", "version", "ZeroRankError", "InvalidDataInStreamException"]
__version__ = "0.1.0"
