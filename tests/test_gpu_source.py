"""Streamed featurization (coordinates.source with features=): the estimators
see the pairwise-distance features one chunk of trajectory segments at a time.

Cases: trajectories shorter than the lag, longer than a chunk (split into
segments with lag overlap), stride and skip; chunk sizes from a few frames to
all of them.  Streamed == materialized: FP64-tile moments to rounding (the
chunks are merged with the Chan update), TICA eigenvalues within 1e-6 of the
FP64 oracle in the default precision, assignment labels identical, TICA
projections identical for the same model.  Non-finite coordinates raise.
"""
import numpy as np
import pytest

from oracle import eig as oeig
from oracle import moments as om
from pyemma_amd import _native
from pyemma_amd.coordinates import assign_to_centers, covariance_lagged, featurizer, source, tica
from pyemma_amd.exceptions import InvalidDataInStreamException

pytestmark = pytest.mark.gpu

LENGTHS = [3000, 50, 7000, 1200, 9, 4000]
N_ATOMS = 8


def _xyz(seed=3, lengths=LENGTHS):
    rng = np.random.default_rng(seed)
    out = []
    for T in lengths:
        steps = rng.standard_normal((T, N_ATOMS, 3)) * 0.05
        out.append((np.cumsum(steps, axis=0) + rng.standard_normal((1, N_ATOMS, 3))).astype(np.float32))
    return out


def _feat():
    f = featurizer(N_ATOMS)
    f.add_distances(f.pairs(np.arange(N_ATOMS)))
    return f


@pytest.mark.parametrize("chunk", [37, 1000, 100000])
@pytest.mark.parametrize("lag,stride,skip", [(7, 1, 0), (7, 2, 3), (1, 3, 5)])
def test_streamed_moments_equal_materialized(chunk, lag, stride, skip):
    xyz, f = _xyz(), _feat()
    src = source(xyz, features=f, chunksize=chunk)
    feats = f.transform(xyz)
    kw = dict(c00=True, c0t=True, ctt=True, lag=lag, stride=stride, skip=skip, remove_data_mean=True,
              reversible=False, precision="fp64")
    a = covariance_lagged(src, **kw)
    b = covariance_lagged(feats, **kw)
    for m in ("C00_", "C0t_", "Ctt_"):
        np.testing.assert_allclose(getattr(a, m), getattr(b, m), rtol=1e-9, atol=1e-12 * np.abs(getattr(b, m)).max())
    np.testing.assert_allclose(a.mean, b.mean, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(a.mean_tau, b.mean_tau, rtol=1e-10, atol=1e-12)


def test_streamed_tica_vs_oracle_and_outputs():
    xyz, f = _xyz(seed=5), _feat()
    src = source(xyz, features=f, chunksize=1500)
    feats = f.transform(xyz)
    lag = 10
    T = tica(src, lag=lag, dim=4, kinetic_map=False)
    Tm = tica(feats, lag=lag, dim=4, kinetic_map=False)
    ref = om.covariance_lagged(feats, lag, c00=True, c0t=True, remove_data_mean=True, reversible=True,
                               bessel=False)
    l_ref, _ = oeig.generalized_eig(ref["C00"], ref["C0t"])
    l_ref = np.sort(np.real(l_ref))[::-1]
    l = np.sort(T.eigenvalues)[::-1]
    np.testing.assert_allclose(l, l_ref, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(np.sort(Tm.eigenvalues)[::-1], l_ref, rtol=1e-6, atol=1e-6)
    # the same model projects streamed and materialized data identically
    Ys = Tm.transform(src)
    Ym = Tm.transform(feats)
    assert len(Ys) == len(Ym) == len(xyz)
    for ys, ym in zip(Ys, Ym):
        np.testing.assert_array_equal(ys, ym)
    # get_output of a streamed estimate: one array per trajectory, every frame
    out = T.get_output()
    assert [len(o) for o in out] == LENGTHS


@pytest.mark.parametrize("chunk", [64, 2500, 100000])
def test_streamed_assignment_labels_identical(chunk):
    xyz, f = _xyz(seed=7), _feat()
    feats = f.transform(xyz)
    allf = np.concatenate(feats)
    rng = np.random.default_rng(1)
    C = allf[rng.choice(len(allf), 40, replace=False)]
    a = assign_to_centers(source(xyz, features=f, chunksize=chunk), centers=C)
    b = assign_to_centers(feats, centers=C)
    assert len(a) == len(b) == len(xyz)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_streamed_device_coordinates():
    """Coordinates already on the device (torch tensors) are sliced there."""
    t = _native.torch()
    xyz, f = _xyz(seed=11, lengths=[2000, 3000]), _feat()
    dev = [t.as_tensor(x, device=_native.device()) for x in xyz]
    a = covariance_lagged(source(dev, features=f, chunksize=700), lag=5, precision="fp64")
    b = covariance_lagged(f.transform(xyz), lag=5, precision="fp64")
    np.testing.assert_allclose(a.C0t_, b.C0t_, rtol=1e-9, atol=1e-12 * np.abs(b.C0t_).max())


def test_streamed_nonfinite_raises():
    xyz, f = _xyz(seed=13, lengths=[500, 600]), _feat()
    xyz[1][321, 2, 0] = np.nan
    with pytest.raises(InvalidDataInStreamException):
        covariance_lagged(source(xyz, features=f, chunksize=200), lag=3)
