"""CPU oracle for the TICA-moments + k-means discretization path.

TEST INFRASTRUCTURE ONLY.  Everything under ``oracle/`` is the checker that
the HIP path is compared against.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it; the product
package ``pyemma_amd`` never does, and fails loudly when its HIP library is
missing instead of falling back to anything here.

Contents
--------
moments   FP64 NumPy restatement of the reference's lagged, mean-free moment
          estimator (pyemma/_ext/variational/estimators/moments.py:699-839,
          running_moments.py:9-148, coordinates/estimation/covariance.py:160-301)
          and of the lag pairing of coordinates/data/_base/iterable.py:95-160.
eig       FP64 generalized eigenproblem C_t r = C_0 r l solved directly with
          LAPACK, the closed form that the reference's eig_corr
          (pyemma/_ext/variational/solvers/direct.py:225-295) approximates.
counts    transition count matrix and state histogram of discrete
          trajectories (deeptime's count_matrix / count_states as called at
          pyemma/msm/estimators/_dtraj_stats.py:60-64, 138-145).
featurize FP32 pairwise distances of particle coordinates (mdtraj's
          non-periodic compute_distances as DistanceFeature.transform calls it,
          coordinates/data/featurization/distances.py:53-54).
native    ctypes binding of emma_oracle.c: the sequential FP32 assign loop,
          kmeans++ with a documented RNG/summation order, and Lloyd.

Pinning status (what each function is checked against) is listed in the
module docstrings and in the tests under tests/test_oracle_*.py.
"""
