"""The fluence-rate-dependent Monte Carlo iteration (runMonteCarlo.m:80-113)
around the GPU operators: per-voxel parameters from the fluence rate, GPU
binning into sub-media with the callers' sortrows relabelling, the media
table and map rebuilt, and the packet ramp.  Host logic over GPU runs;
no reference output exists to compare with (MATLAB), so the checks are the
structure of the split media and the physics of a saturable absorber."""
from __future__ import annotations

import numpy as np
import pytest

import models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def test_fr_independent_medium_splits_into_equal_submedia(api):
    m = models.base(n=20, nph=2e4)
    nB = 4
    r = api.fr_dependent_MCmatlab(m, 2, lambda FR: (20.0, 200.0, 0.9), nB, 3, seed=5)
    mp = r.MC.mediaProperties
    assert mp["mua"].shape[0] == 4 - 1 + nB
    np.testing.assert_array_equal(mp["mua"][1:1 + nB, 0], np.float32(20.0))
    np.testing.assert_array_equal(mp["mus"][1:1 + nB, 0], np.float32(200.0))
    M0, M = np.asarray(m.MC.M), np.asarray(r.MC.M)
    assert np.all(M[M0 == 1] == 1) and np.all(M[M0 == 2] == 2)  # identical sets all take label 1
    assert np.all(M[M0 == 3] == 3 + nB - 1) and np.all(M[M0 == 4] == 4 + nB - 1)
    assert r.MC.nPhotons == 20000  # the last run of the ramp launches the full request


def test_saturable_absorber_bins_follow_the_fluence_rate(api):
    m = models.base(n=20, nph=4e4)
    m.MC.P = 1.0
    plain = api.MCmatlab(m, 1, seed=1)
    M0 = np.asarray(m.MC.M)
    vox = (M0 == 3).ravel(order="F")
    FR1 = np.asarray(plain.MC.NFR).ravel(order="F")[vox]
    Fsat = float(np.median(FR1))
    mua0 = 5.0
    nB = 6
    r = api.fr_dependent_MCmatlab(m, 3, lambda FR: (mua0 / (1 + FR / Fsat), 200.0, 0.9), nB, 3, seed=9)
    sub = r.MC.mediaProperties["mua"][2:2 + nB, 0]
    assert np.all(np.diff(sub) >= 0) and sub.max() <= mua0 and sub.min() > 0  # sorted ascending, saturated
    assert sub.max() - sub.min() > 0.2 * mua0
    M = np.asarray(r.MC.M).ravel(order="F")[vox]
    FR = np.asarray(r.MC.NFR).ravel(order="F")[vox]
    lo, hi = M == 3, M == 3 + nB - 1  # lowest-mua and highest-mua sub-media
    assert lo.any() and hi.any()
    assert FR[lo].mean() > FR[hi].mean()  # strongly lit voxels are the most bleached


def test_split_remaps_deposition_criteria_bounds(api):
    """Criteria bounds on the split medium and on later media follow the
    split (getOpticalMediaProperties.m:126-133): a lower bound on medium 2
    stays at its first sub-medium, an upper bound on medium 3 moves by
    nBins - 1.  The split run then counts scatterings in media 2..3 only,
    like the oracle on the same split model."""
    import oracle_lib as OL
    from mcstats import compare_scalars

    nB = 3
    m = models.base(n=20, nph=2e4)
    dc = m.MC.depositionCriteria
    dc.minScatterings, dc.evaluateOnlyAtEndOfLife = 3, False
    dc.minSubmediaIdx, dc.maxSubmediaIdx = 2, 3
    r = api.fr_dependent_MCmatlab(m, 2, lambda FR: (20.0, 200.0, 0.9), nB, 1, seed=21)
    rd = r.MC.depositionCriteria
    assert (rd.minSubmediaIdx, rd.maxSubmediaIdx) == (2, 3 + nB - 1)
    split = r.copy()
    G, O = [], []
    for s in range(6):
        g = api.MCmatlab(split, 1, seed=300 + s).MC
        G.append({"NFR": g.NFR, "NI_zneg": g.NI_zneg, "NI_zpos": g.NI_zpos})
        O.append(OL.oracle_mc(split, seed=s))
    bad = [row for row in compare_scalars(G, O, split) if not abs(row[1]) <= row[2]]
    assert not bad, bad
    t = api.t_dependent_heat  # the heat-side helper validates its medium count the same way
    with pytest.raises(ValueError):
        t(np.zeros((2, 2, 2), np.float32), np.float32(np.nan), {"A": [0], "E": [0]}, np.ones((2, 2, 2), np.uint8), [3.7], [0.0037],
          (1e-3,) * 3, 1, lambda T: (3.7, 0.0037), 256, [(1, 1e-3, 0)])
