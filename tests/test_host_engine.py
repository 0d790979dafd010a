"""The GPU kernel's per-packet engine (mc_engine.h / mc_photon.h, FP32),
compiled for the host in tests/_build/libmc_host.so, against the oracle and
the closed forms -- on the CPU, so every change to the GPU physics is checked
before it reaches a GPU.  The wavefront scheduling of the real kernel is
covered by the GPU tests."""
import ctypes as C
import os

import numpy as np
import pytest

import closed_forms as CF
import models
import oracle_lib as OL
from mcmatlab_amd import abi
from mcstats import compare_scalars, compare_voxels

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_build", "libmc_host.so")
pytestmark = pytest.mark.skipif(not os.path.exists(HARNESS), reason="host harness not built")

_lib = None


def host_run(m, seed, records=False):
    """records: launch through LaunchRecord and tally escapes through
    EscapeRecord (the GPU's launch / transport / tally kernel split)."""
    global _lib
    if _lib is None:
        _lib = C.CDLL(HARNESS)
        _lib.mc_host_run.argtypes = [C.POINTER(abi.mcm_mc_input), C.POINTER(abi.mcm_mc_output), C.c_uint64, C.c_void_p]
        _lib.mc_host_set_mode.argtypes = [C.c_int]
    inp = abi.MCInputs(m, seed=seed)
    out = abi.MCOutputs(inp)
    _lib.mc_host_set_mode(1 if records else 0)
    try:
        assert _lib.mc_host_run(C.byref(inp.struct), C.byref(out.struct), seed, None) == 0
    finally:
        _lib.mc_host_set_mode(0)
    return out.result()


@pytest.mark.parametrize("name", ["pencil_bt0", "pencil_bt2", "fresnel_bt1", "fresnel_bt3", "radial_tophat_gauss", "planewave_bt2",
                                  "lens_timebins", "fiber", "farfield", "source3d", "xy_custom_angular", "isotropic_line"])
def test_launch_and_escape_records_reproduce_inline_path(name):
    """The launch record carries the packet and its random stream exactly, and
    the deferred escape tallies add the same weights in the same order, so
    both paths give bit-identical outputs."""
    m = models.cases()[name]()
    m.MC.nPhotonsRequested = 3000
    a = host_run(m, 77)
    b = host_run(m, 77, records=True)
    assert a["nPhotons"] == b["nPhotons"] and a["nPhotonsCollected"] == b["nPhotonsCollected"]
    for k, v in a.items():
        if isinstance(v, np.ndarray) and v.dtype.kind == "f":
            assert np.array_equal(v, b[k], equal_nan=True), k


@pytest.mark.parametrize("name", ["pencil_bt1", "fresnel_bt1", "fresnel_bt3", "radial_tophat_gauss", "planewave_bt2", "lens_timebins",
                                  "criteria_scatter", "source3d", "custom_phase", "xy_custom_angular"])
def test_host_engine_matches_oracle(name):
    m = models.cases()[name]()
    m.MC.nPhotonsRequested = 15000
    H = [host_run(m, 1000 + s) for s in range(8)]
    O = [OL.oracle_mc(m, seed=s) for s in range(8)]
    bad = [r for r in compare_scalars(H, O, m) if not abs(r[1]) <= r[2]]
    assert not bad, bad
    frac, mt2, n = compare_voxels(H, O)
    assert frac < 0.02 and mt2 < 1.5, (frac, mt2, n)


def test_host_engine_beer_lambert():
    CF.beer_lambert(host_run, rtol=3e-5)  # FP32 transport


@pytest.mark.parametrize("check", ["fresnel_normal", "tir", "energy", "point_source"])
def test_host_engine_closed_forms(check):
    getattr(CF, check)(host_run)


@pytest.mark.slow
def test_host_engine_hg_single():
    CF.hg_single(host_run)
