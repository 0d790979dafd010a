"""GPU Monte Carlo parity: the HIP transport (through the C ABI) against the
FP64 oracle, statistically, on the feature matrix of tests/models.py.

The two sides use different random streams (Philox on the GPU, dSFMT in the
oracle), so the bar is statistical (north-star): every scalar summary
(absorbed fraction, escape per face, collected light, far field) agrees
within 3 sigma -- Welch's t at the two-sided 3-sigma tail probability -- from
S seeds per side, and per-voxel NFR agrees within its standard error (at most
2% of voxels beyond 3 sigma, mean t^2 < 1.5).  Photon counts are exact.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import models
import oracle_lib as OL
from mcstats import check_scalars_two_stage, compare_scalars, compare_voxels

pytestmark = pytest.mark.gpu

SEEDS = 8
NPH = 20000


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1, "no gfx950 device visible to libmcm"
    return A


def gpu_runs(api, m, seeds, simType=1):
    out = []
    for s in seeds:
        r = api.MCmatlab(m, simType, seed=1000 + s)
        mc = r.FMC if simType == 2 else r.MC
        d = {k: getattr(mc, k) for k in ("NFR", "farField", "NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg",
                                         "nPhotons", "nPhotonsCollected")}
        d = {k: v for k, v in d.items() if not (np.isscalar(v) and isinstance(v, float) and math.isnan(v))}
        if mc.useLightCollector:
            d["image"] = np.asarray(mc.LC.image)
        out.append(d)
    return out


CASES = ["pencil_bt0", "pencil_bt1", "pencil_bt2", "pencil_bt3", "pencil_oblique", "fresnel_bt1", "fresnel_bt3",
         "isotropic_point", "isotropic_line", "planewave_bt1", "planewave_bt2", "radial_tophat_gauss", "radial_gauss_lambert",
         "radial_custom", "xy_tophat_gauss", "xy_custom_lambert", "xy_custom_angular", "lens_timebins", "fiber", "farfield",
         "criteria_scatter", "criteria_collected", "criteria_now", "custom_phase", "source3d", "lg01", "fmc_source3d"]
SIM_TYPE = {"fmc_source3d": 2}  # the fluorescence route: the MEX reads model.FMC


@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_oracle(api, name):
    print(f"[gpu-mc] case {name}", flush=True)
    st = SIM_TYPE.get(name, 1)
    m = models.cases()[name]()
    (m.FMC if st == 2 else m.MC).nPhotonsRequested = NPH
    G = gpu_runs(api, m, range(SEEDS), simType=st)
    O = [OL.oracle_mc(m, simType=st, seed=s) for s in range(SEEDS)]
    for g in G:
        assert g["nPhotons"] == NPH
    bad = check_scalars_two_stage(lambda s: gpu_runs(api, m, [s], simType=st)[0], lambda s: OL.oracle_mc(m, simType=st, seed=s), m,
                                  simType=st, runs=(G, O))
    assert not bad, f"scalar summaries outside 3 sigma: {bad}"
    frac, mt2, n = compare_voxels(G, O)
    assert n > 100
    assert frac < 0.02 and mt2 < 1.5, (frac, mt2, n)


def test_gpu_wavelengths(api):
    """Three wavelengths in one call: the photon budget splits as in
    MCmatlab.c:635 and each wavelength's slice matches the oracle."""
    m = models.cases()["wavelengths3"]()
    m.MC.nPhotonsRequested = 3 * NPH + 1
    G = gpu_runs(api, m, range(SEEDS))
    O = [OL.oracle_mc(m, seed=s) for s in range(SEEDS)]
    assert G[0]["nPhotons"] == 3 * NPH + 1
    for iL in range(3):
        bad = check_scalars_two_stage(lambda s: gpu_runs(api, m, [s])[0], lambda s: OL.oracle_mc(m, seed=s), m, iL=iL,
                                      runs=(G, O))
        assert not bad, (iL, bad)


def test_gpu_deterministic_given_seed(api):
    """Packets draw from per-packet counter streams, so a seed fixes the
    result up to FP64 summation order."""
    m = models.cases()["fresnel_bt1"]()
    m.MC.nPhotonsRequested = 5000
    a = api.MCmatlab(m, 1, seed=7).MC
    b = api.MCmatlab(m, 1, seed=7).MC
    np.testing.assert_allclose(a.NFR, b.NFR, rtol=1e-9, atol=0)
    np.testing.assert_allclose(a.NI_zneg, b.NI_zneg, rtol=1e-9, atol=0)


def test_gpu_example_paths(api):
    m = models.cases()["paths"]()
    r = api.MCmatlab(m, 1, seed=3).MC
    P = np.asarray(r.examplePaths)
    assert P.shape[0] == 4 and P.shape[1] > 8
    # paths are separated by NaN columns; the first path starts with two
    nan_cols = np.isnan(P[0])
    assert nan_cols[0] and nan_cols[1]
    assert np.all(P[3][~nan_cols] > 0)  # weights positive


def test_gpu_invalid_input_is_an_error(api):
    from mcmatlab_amd._lib import MCMError

    m = models.cases()["pencil_bt1"]()
    m.MC.M = m.MC.M.copy()
    m.MC.M[0, 0, 0] = 9  # medium index beyond the media table
    with pytest.raises(MCMError):
        api.MCmatlab(m, 1, seed=1)


@pytest.mark.parametrize("name", ["pencil_bt1", "fresnel_bt1"])
def test_gpu_fluence_copies_keep_results(api, name):
    """The autotuner's XCD-private fluence copies are summed before any read,
    so a tuned plan gives the untuned plan's outputs (up to FP64 summation
    order), and tuning leaves no trial deposits behind."""
    m = models.cases()[name]()
    a = api.MCPlan(m)
    b = api.MCPlan(m)
    try:
        ms = b.autotune(nPackets=20000)
        assert ms[1] > 0 and ms[8] > 0
        for plan in (a, b):
            plan.simulate(30000, seed=5, firstPhoton=0)
        ra, rb = a.normalize(), b.normalize()
    finally:
        a.close()
        b.close()
    for k in ("NFR", "NI_zneg", "NI_zpos"):
        np.testing.assert_allclose(rb[k], ra[k], rtol=1e-5, atol=1e-6 * float(np.max(np.abs(ra[k]))))


@pytest.mark.parametrize("copies", [2, 8])
def test_gpu_set_fluence_copies(api, copies):
    """Copies set by hand: deposits made before the change are folded and
    kept, and the outputs equal a single-copy plan's."""
    m = models.cases()["pencil_bt1"]()
    a = api.MCPlan(m)
    b = api.MCPlan(m)
    try:
        with pytest.raises(RuntimeError):
            b.set_fluence_copies(3)
        for plan in (a, b):
            plan.simulate(10000, seed=6, firstPhoton=0)
        b.set_fluence_copies(copies)
        for plan in (a, b):
            plan.simulate(20000, seed=6, firstPhoton=10000)
        b.set_fluence_copies(1)
        b.set_fluence_copies(copies)
        ra, rb = a.normalize(), b.normalize()
    finally:
        a.close()
        b.close()
    for k in ("NFR", "NI_zneg", "NI_zpos"):
        np.testing.assert_allclose(rb[k], ra[k], rtol=1e-5, atol=1e-6 * float(np.max(np.abs(ra[k]))))
