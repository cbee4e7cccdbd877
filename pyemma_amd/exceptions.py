"""Exceptions with the reference's names and bases."""
import numpy as np


class ZeroRankError(np.linalg.LinAlgError):
    """Input matrix has rank zero (pyemma/_ext/variational/util.py:8)."""


class InvalidDataInStreamException(Exception):
    """Non-finite values in the input (pyemma/coordinates/data/_base/datasource.py:1075)."""


class NotConvergedWarning(UserWarning):
    """k-means did not reach its tolerance within max_iter."""
