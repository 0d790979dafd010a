"""The Monte Carlo closed forms of tests/closed_forms.py run directly on the
GPU (through the C ABI): launch distributions of every source type, plane
wave normalisation, Beer-Lambert, Fresnel at normal and oblique incidence,
total internal reflection, the cyclic wrap and the too-old kill, light
collectors with time bins, deposition criteria, tabulated phase functions
and example paths.  These pin the GPU against the mathematics, independently
of the oracle it is compared with elsewhere.  FP32 transport: exact
deposits are compared at 5e-5 relative."""
from __future__ import annotations

import math

import numpy as np
import pytest

import closed_forms as CF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_run():
    from mcmatlab_amd import _lib, api

    assert _lib.lib().mcm_device_count() >= 1

    def run(m, seed):
        mc = api.MCmatlab(m, 1, seed=9000 + seed).MC
        keys = ("NFR", "farField", "NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg", "nPhotons",
                "nPhotonsCollected", "examplePaths")
        d = {k: getattr(mc, k) for k in keys}
        if mc.useLightCollector:
            d["image"] = np.asarray(mc.LC.image)
        return {k: v for k, v in d.items() if not (isinstance(v, float) and math.isnan(v))}

    return run


CHECKS = ["beer_lambert", "fresnel_normal", "tir", "energy", "point_source", "planewave", "radial_focal_profiles",
          "radial_angular_profiles", "xy_profiles", "lg01", "cyclic_wrap", "too_old", "fresnel_oblique", "collectors",
          "unscattered_criteria", "example_paths", "source3d", "line_source", "tabulated_phase", "hg_single"]


@pytest.mark.parametrize("check", CHECKS)
def test_gpu_closed_form(gpu_run, check):
    print(f"[gpu-closed-form] {check}", flush=True)
    if check == "beer_lambert":
        CF.beer_lambert(gpu_run, rtol=3e-5)
    else:
        getattr(CF, check)(gpu_run)
