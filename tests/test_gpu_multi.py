"""Multi-grid launches (mcm_mc_multi_run) against plain launches.

Items on different grids (16^3 .. 28^3; Fresnel and matched interfaces,
boundary types 1-3, light collector with time bins, far field, custom
phase function, several wavelengths) run in one sequence of persistent
launches per interface class; each lane reads its packet's grid.  Item
(plan, iL, n, seed, stream = iL) draws exactly the packets of that plan's
mcm_mc_plan_simulate(iL, n, seed), so its normalised outputs equal the
plain launch's up to the order of the FP64 atomic sums (then rounded to
single): compared at 1e-5 relative, and the collected-packet counts
exactly.  Plain launches are themselves checked against the oracle in
test_gpu_mc.py.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from models import _custom_cdf_media, _dc, _lc, _set, base
from mcmatlab_amd.model import SourceIntensityDistribution as SID
from test_gpu_batch import check_equal, plain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def grids():
    return [
        _set(base(n=24, matched=False), LS_theta=0.3, LS_phi=0.5),
        base(n=16, nL=3, nph=3001),
        _lc(base(n=20), f=0.5, diam=0.5, fieldSize=0.1, res=8, nTimeBins=6, tEnd=2e-11, z=-0.3),
        _set(base(n=28), farFieldRes=12),
        _set(base(n=18, matched=False, bt=3), sourceType=4, FPID=SID(radialDistr=0, radialWidth=0.02),
             AID=SID(radialDistr=1, radialWidth=0.2)),
        _set(base(n=22, bt=2), sourceType=2),
        _custom_cdf_media(base(n=26)),
    ]


ITEMS = [  # (plan, iL, nPhotons, seed, stream)
    (0, 0, 6000, 11, 0), (1, 0, 4000, 12, 0), (1, 2, 7000, 12, 2), (2, 0, 5000, 13, 0), (3, 0, 6000, 14, 0),
    (4, 0, 5000, 15, 0), (5, 0, 4000, 16, 0), (6, 0, 5000, 17, 0), (0, 0, 3000, 18, 0), (1, 1, 2000, 19, 1),
]


def test_multi_items_equal_plain_launches(api):
    ms = grids()
    plans = [api.MCPlan(m) for m in ms]
    try:
        got = api.multi_run(plans, ITEMS)
    finally:
        for p in plans:
            p.close()
    for k, ((pi, iL, n, seed, stream), g) in enumerate(zip(ITEMS, got)):
        print(f"[gpu-multi] item {k} plan {pi}", flush=True)
        ref = plain(api, ms[pi], [(iL, n, seed, stream)])[0]
        assert g["nPhotons"] == n
        if "image" in g:
            assert g["nPhotonsCollected"] == ref[1]["collected"]
        assert g["NFR"].shape[:3] == ms[pi].MC.M.shape
        check_equal(g, ref, iL)


def test_multi_single_plan_equals_batch(api):
    """On one plan, a multi-grid launch is a batched launch."""
    m = base(n=20, nL=3, nph=3001)
    plan = api.MCPlan(m)
    try:
        items = [(0, 0, 5000, 3, 0), (0, 2, 3000, 4, 2)]
        got = api.multi_run([plan], items)
        ref = plan.batch([(iL, n, s, st) for _, iL, n, s, st in items])
    finally:
        plan.close()
    for g, r, it in zip(got, ref, items):
        a, b = np.asarray(g["NFR"], np.float64), np.asarray(r["NFR"], np.float64)
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.max(np.abs(b))), it


def test_multi_unsupported_and_invalid(api):
    ms = [base(n=16), _dc(base(n=18), minScatterings=3, maxScatterings=40)]
    plans = [api.MCPlan(m) for m in ms]
    try:
        with pytest.raises(RuntimeError, match="criteria"):
            api.multi_run(plans, [(0, 0, 100, 1, 0), (1, 0, 100, 1, 0)])
        with pytest.raises(RuntimeError, match="plan index"):
            api.multi_run(plans, [(2, 0, 100, 1, 0)])
        with pytest.raises(RuntimeError, match="wavelength"):
            api.multi_run(plans, [(0, 1, 100, 1, 0)])
        r = api.multi_run(plans[:1], [(0, 0, 1000, 1, 0)])
        assert r[0]["nPhotons"] == 1000
    finally:
        for p in plans:
            p.close()
