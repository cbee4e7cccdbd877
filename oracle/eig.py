"""FP64 closed form of the TICA eigenproblem.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference solves C_t r = C_0 r l through a rank-truncated whitening
(pyemma/_ext/variational/solvers/direct.py:225-295, eig_corr).  When C_0 is
well conditioned (no eigenvalue below epsilon) that is the generalized
symmetric-definite problem, which LAPACK solves directly; the reference's
own test (solvers/tests/test_direct.py:36-71) checks eig_corr against
scipy.linalg.eig(Ct, C0) in the same way.  This module is that closed form.

Pinned by: the 3x3 known-answer matrices of test_direct.py:36-71 and the
identity K = C0^-1 Ct = R diag(l) R^-1 (tests/test_oracle_eig.py).
"""
import numpy as np
import scipy.linalg as sla

__all__ = ["generalized_eigenvalues", "generalized_eig"]


def generalized_eig(C0, Ct):
    """Eigenvalues (descending |l|) and eigenvectors of Ct r = C0 r l, FP64."""
    C0 = np.asarray(C0, dtype=np.float64)
    Ct = np.asarray(Ct, dtype=np.float64)
    if np.allclose(Ct, Ct.T):
        l, R = sla.eigh(0.5 * (Ct + Ct.T), 0.5 * (C0 + C0.T))
    else:
        l, R = sla.eig(Ct, C0)
    order = np.argsort(np.abs(l), kind="stable")[::-1]
    return l[order], R[:, order]


def generalized_eigenvalues(C0, Ct):
    return generalized_eig(C0, Ct)[0]
