"""Batched persistent launches (mcm_mc_batch_run) against plain launches.

An item (wavelength iL, nPhotons, seed, stream = iL) draws exactly the
packets of mcm_mc_plan_simulate(iL, nPhotons, seed), so its normalised
outputs equal the plain launch's up to the order of the FP64 atomic sums
(then rounded to single): compared at 1e-5 relative.  Plain launches are
themselves checked against the oracle in test_gpu_mc.py.
"""
from __future__ import annotations

import numpy as np
import pytest

import models

pytestmark = pytest.mark.gpu

KEYS = ("NFR", "image", "farField", "NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg")


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def plain(api, m, items):
    plan = api.MCPlan(m)
    try:
        out = []
        for iL, n, seed, stream in items:
            assert stream == iL
            plan.simulate(n, seed=seed, iL=iL, firstPhoton=0)
            c = plan.counters()
            r = plan.normalize(iL)
            out.append((r, c))
        return out
    finally:
        plan.close()


def check_equal(batch, ref, iL):
    r, _ = ref
    for k in KEYS:
        if k not in batch:
            continue
        a = np.asarray(batch[k], np.float64).ravel(order="F")
        b = np.asarray(r[k], np.float64).reshape(-1, r[k].shape[-1] if r[k].ndim > 2 else 1, order="F")
        b = b[:, iL] if b.shape[1] > 1 else b[:, 0]
        scale = max(np.max(np.abs(b)), 1e-30)
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * scale), k


@pytest.mark.parametrize("name,items", [
    ("wavelengths3", [(0, 7000, 11, 0), (1, 3000, 11, 1), (2, 12000, 11, 2), (1, 5000, 12, 1)]),
    ("fresnel_bt1", [(0, 5000, 21, 0), (0, 9000, 22, 0)]),
    ("lens_timebins", [(0, 4000, 31, 0), (0, 6000, 32, 0)]),
    ("farfield", [(0, 6000, 41, 0)]),
    ("planewave_bt2", [(0, 4000, 51, 0), (0, 4000, 52, 0)]),
])
def test_batch_items_equal_plain_launches(api, name, items):
    print(f"[gpu-batch] {name}", flush=True)
    m = models.cases()[name]()
    plan = api.MCPlan(m)
    try:
        got = plan.batch(items)
    finally:
        plan.close()
    for (iL, n, seed, stream), g in zip(items, got):
        ref = plain(api, m, [(iL, n, seed, stream)])[0]
        assert g["nPhotons"] == n
        if "image" in g:
            assert g["nPhotonsCollected"] == ref[1]["collected"]
        check_equal(g, ref, iL)


def test_batch_many_small_items(api):
    """64 items with 1e3..1e4 packets each: every item's absorbed fraction is
    finite and the packets of the whole batch are all accounted for."""
    m = models.cases()["pencil_bt1"]()
    rng = np.random.default_rng(5)
    items = [(0, int(n), 100 + k, 0) for k, n in enumerate(rng.integers(1000, 10000, 64))]
    plan = api.MCPlan(m)
    try:
        got = plan.batch(items)
    finally:
        plan.close()
    assert [g["nPhotons"] for g in got] == [n for _, n, _, _ in items]
    for g in got:
        assert np.isfinite(g["NFR"]).all() and g["NFR"].max() > 0


@pytest.mark.parametrize("name", ["criteria_scatter", "source3d"])
def test_batch_unsupported_inputs_are_errors(api, name):
    from mcmatlab_amd._lib import MCMError

    m = models.cases()[name]()
    plan = api.MCPlan(m)
    try:
        with pytest.raises(MCMError):
            plan.batch([(0, 100, 1, 0)])
    finally:
        plan.close()
