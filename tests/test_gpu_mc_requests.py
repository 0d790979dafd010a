"""Run-length requests of the MC MEX (MCmatlab.c:334-336, 674-714): a
simulation time instead of a packet count (nPhotonsRequested = NaN), and a
count of collected packets (requestCollectedPhotons) instead of launched
ones.  The outputs are normalised per launched packet either way, so a
timed run estimates the same fluence as a counted one: the absorbed
fraction (sum of NFR * mua * dV) is compared at 5 sigma of the counted
runs' spread."""
from __future__ import annotations

import math

import numpy as np
import pytest

import models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def absorbed(m, r):
    mua = np.asarray(m.MC.mediaProperties["mua"], np.float64)[:, 0][m.MC.M - 1]
    d = m.G.dx * m.G.dy * m.G.dz
    return float(np.sum(np.asarray(r.NFR, np.float64).reshape(m.MC.M.shape) * mua) * d)


def test_time_limited_run(api):
    m = models.cases()["fresnel_bt1"]()
    m.MC.nPhotonsRequested = math.nan
    m.MC.simulationTimeRequested = 0.004  # minutes
    r = api.MCmatlab(m, 1, seed=5).MC
    assert r.nPhotons > 1e5
    assert 0.5 * 0.004 <= r.simulationTime <= 4 * 0.004, r.simulationTime
    m2 = models.cases()["fresnel_bt1"]()
    m2.MC.nPhotonsRequested = 2e5
    fr = [absorbed(m2, api.MCmatlab(m2, 1, seed=10 + s).MC) for s in range(4)]
    sd = np.std(fr, ddof=1) * math.sqrt(2e5 / r.nPhotons + 1)
    assert abs(absorbed(m, r) - np.mean(fr)) <= 5 * sd + 1e-6, (absorbed(m, r), fr)


def test_time_limited_run_several_wavelengths(api):
    m = models.cases()["wavelengths3"]()
    m.MC.nPhotonsRequested = math.nan
    m.MC.simulationTimeRequested = 0.003
    r = api.MCmatlab(m, 1, seed=6).MC
    assert r.nPhotons > 3e4
    nfr = np.asarray(r.NFR)
    assert nfr.shape[-1] == 3 and np.all(np.isfinite(nfr)) and np.all(nfr.reshape(-1, 3).sum(0) > 0)


@pytest.mark.parametrize("name", ["lens_timebins", "fiber"])
def test_collected_packet_request(api, name):
    m = models.cases()[name]()
    m.MC.requestCollectedPhotons = True
    m.MC.nPhotonsRequested = 3000
    r = api.MCmatlab(m, 1, seed=7).MC
    assert r.nPhotonsCollected >= 3000
    assert r.nPhotons > r.nPhotonsCollected
    # the image is normalised per launched packet: the same as a counted run of as many packets
    m2 = models.cases()[name]()
    m2.MC.nPhotonsRequested = float(r.nPhotons)
    r2 = api.MCmatlab(m2, 1, seed=8).MC
    a, b = float(np.sum(r.LC.image)), float(np.sum(r2.LC.image))
    rel = 5 / math.sqrt(r.nPhotonsCollected)
    assert abs(a - b) <= rel * max(a, b), (a, b)
