"""GPU heat propagator parity through the C ABI.

FP64 (default): temperature and damage within 1e-9 relative of the FP64 CPU
restatement of the reference scheme, for DG-ADI and the
explicit scheme, over boundary types, lights on/off, heterogeneous media,
sensors and per-medium maxima.  FP32: the reference's own arithmetic; the
temperature field is compared with the FP32 oracle (itself bit-identical to
the reference build) to 2 ulp, damage to 1e-5 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as OL
from mcmatlab_amd.media import heat_coefficients
from mcmatlab_amd.workloads import layered_tissue_heat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def heat_case(n=16, bt=1, lights=True, steps=12, damage=True, hetero=True, seed=0, shape=None):
    T0, Om0, Pon, Poff = layered_tissue_heat(n=n, steps_on=steps, steps_off=steps)
    P = dict(Pon if lights else Poff)
    P["heatBoundaryType"] = bt
    d = 0.1 / n
    if hetero:
        P["dTdtperdeltaT"] = heat_coefficients([1.2e-3, 3.7, 3.9, 3.8], [2.6e-4, 0.0037, 0.0052, 0.0045], d, d, d)
    rng = np.random.default_rng(seed)
    T0 = (37 + 8 * rng.random(T0.shape)).astype(np.float32)
    if hetero:
        P["dTdt_abs"] = (np.asarray(P["dTdt_abs"]) * 50).astype(np.float32)
    P["dt"] = 2e-5 if hetero else P["dt"]
    if shape is not None:
        nx, ny, nz = shape
        T0 = T0[:nx, :ny, :nz].copy(order="F")
        Om0 = Om0[:nx, :ny, :nz].copy(order="F")
        P["M"] = np.asarray(P["M"])[:nx, :ny, :nz].copy(order="F")
        P["dTdt_abs"] = np.asarray(P["dTdt_abs"])[:nx, :ny, :nz].copy(order="F")
        from mcmatlab_amd.media import sensor_weights

        li, w = sensor_weights([[0.0, 0.0, 0.01], [0.005, -0.003, 0.02]], (np.arange(nx) - (nx - 1) / 2) * d,
                               (np.arange(ny) - (ny - 1) / 2) * d, (np.arange(nz) + 0.5) * d)
        P["tempSensorCornerIdxs"], P["tempSensorInterpWeights"] = li, w
    return T0, (Om0 if damage else np.float32(np.nan)), P


CASES = {
    "sinked_on": dict(),
    "insulated_on": dict(bt=0),
    "lights_off": dict(lights=False),
    "no_damage": dict(damage=False),
    "one_step": dict(steps=1),
    "homogeneous": dict(hetero=False),
    "odd_grid": dict(shape=(13, 9, 11)),
    "tall_grid": dict(shape=(5, 7, 16), bt=0),
    # several x/y tiles and z runs of the explicit stencil, several tiles of the y/z sweeps
    "multi_tile": dict(n=72, shape=(70, 19, 37), steps=5),
    "multi_tile_insulated": dict(n=72, shape=(33, 17, 49), steps=4, bt=0),
}


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("scheme", ["dgadi", "explicit"])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_heat_fp64_matches_oracle(api, name, scheme):
    print(f"[gpu-heat] {name} {scheme}", flush=True)
    T0, Om0, P = heat_case(**CASES[name])
    T, Om, sens, mx = api.finiteElementHeatPropagator(T0, Om0, P, precision="fp64", scheme=scheme)
    T64, Om64 = api.finiteElementHeatPropagator.last_state
    ref = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP64_EXPLICIT if scheme == "explicit" else OL.HEAT_FP64_DGADI)
    assert rel(T64, ref["T_f64"]) <= 1e-9
    if CASES[name].get("damage", True):
        assert rel(Om64, ref["Omega_f64"]) <= 1e-9
        assert np.array_equal(Om, ref["Omega"]) or rel(Om, ref["Omega"]) <= 1e-6
    else:
        assert np.isnan(Om).all() and Om.size == 1
    np.testing.assert_allclose(sens, ref["sensorTemps"], rtol=1e-9)
    np.testing.assert_allclose(mx, ref["maxMediaTemps"], rtol=1e-7)
    np.testing.assert_allclose(T, ref["T"], rtol=2.4e-7)


@pytest.mark.parametrize("name", ["sinked_on", "insulated_on", "lights_off", "odd_grid"])
def test_gpu_heat_fp32_matches_reference_arithmetic(api, name):
    T0, Om0, P = heat_case(**CASES[name])
    T, Om, sens, mx = api.finiteElementHeatPropagator(T0, Om0, P, precision="fp32")
    ref = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP32_DGADI)
    ulp = np.spacing(np.abs(ref["T"]).astype(np.float32))
    assert np.max(np.abs(T - ref["T"]) / ulp) <= 2
    assert rel(Om, ref["Omega"]) <= 1e-5
    np.testing.assert_allclose(sens, ref["sensorTemps"], rtol=1e-6)
    np.testing.assert_allclose(mx, ref["maxMediaTemps"], rtol=1e-6)


def test_gpu_heat_chained_fp64_state(api):
    """Two calls (lights on, then off) carrying the FP64 state equal one
    oracle run of the same schedule."""
    T0, Om0, Pon, Poff = layered_tissue_heat(n=20, steps_on=15, steps_off=25)
    api.finiteElementHeatPropagator(T0, Om0, Pon)
    st = api.finiteElementHeatPropagator.last_state
    T, Om, _, _ = api.finiteElementHeatPropagator(T0, Om0, Poff, state_f64=st)
    T64, Om64 = api.finiteElementHeatPropagator.last_state
    r1 = OL.oracle_heat(T0, Om0, Pon)
    r2 = OL.oracle_heat(T0, Om0, Poff, T_f64=r1["T_f64"], Omega_f64=r1["Omega_f64"])
    assert rel(T64, r2["T_f64"]) <= 1e-9
    assert rel(Om64, r2["Omega_f64"]) <= 1e-9


def test_gpu_heat_invalid_medium_is_an_error(api):
    from mcmatlab_amd._lib import MCMError

    T0, Om0, P = heat_case()
    P["M"] = np.asarray(P["M"]).copy()
    P["M"][0, 0, 0] = 7  # beyond nM - 1
    with pytest.raises(MCMError):
        api.finiteElementHeatPropagator(T0, Om0, P)


def line_case(nx, ny, nz, bt, seed=0, steps=6):
    """Random 4-medium block of nx x ny x nz voxels."""
    rng = np.random.default_rng(seed)
    d = 0.1 / 64
    M = rng.integers(0, 4, (nx, ny, nz)).astype(np.uint8)
    T0 = (37 + 10 * rng.random((nx, ny, nz))).astype(np.float32)
    Om0 = np.zeros((nx, ny, nz), np.float32)
    P = {"M": M, "A": np.array([0, 0, 0, 7.6e66]), "E": np.array([0, 0, 0, 422.5e3]),
         "dTdtperdeltaT": heat_coefficients([1.2e-3, 3.7, 3.9, 3.8], [2.6e-4, 0.0037, 0.0052, 0.0045], d, d, d),
         "dTdt_abs": (2e4 * rng.random((nx, ny, nz))).astype(np.float32), "useAllCPUs": True, "heatBoundaryType": bt,
         "tempSensorCornerIdxs": np.zeros(0), "tempSensorInterpWeights": np.zeros((0, 3)), "lightsOn": True,
         "steps": steps, "dt": 2e-5}
    return T0, Om0, P


@pytest.mark.parametrize("bt", [0, 1])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("n", [1, 32, 33, 64, 65, 128, 129, 200, 256, 257, 300, 512, 513])
def test_gpu_heat_line_lengths(api, n, axis, bt):
    """FP64 lines of up to 512 voxels are solved by the wavefront solver, a
    line per segment of 64, 32 or 16 lanes with 1, 2, 4 or 8 rows per lane
    (the lengths cover every shape and both ends of its range), longer ones
    by the lane-per-line recurrence.  The other two extents are odd, so
    some wavefronts hold a segment with no line of its own."""
    print(f"[gpu-heat] axis {axis} line {n} bt {bt}", flush=True)
    shape = [3, 5, 7]
    shape[axis] = n
    T0, Om0, P = line_case(*shape, bt, seed=n + 1000 * axis)
    api.finiteElementHeatPropagator(T0, Om0, P, precision="fp64")
    T64, Om64 = api.finiteElementHeatPropagator.last_state
    ref = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP64_DGADI)
    assert rel(T64, ref["T_f64"]) <= 1e-9
    assert rel(Om64, ref["Omega_f64"]) <= 1e-9


@pytest.mark.parametrize("zrun", [8, 12, 16, 24, 32])
@pytest.mark.parametrize("name", ["multi_tile", "multi_tile_insulated"])
def test_gpu_heat_explicit_every_zrun_matches_oracle(api, name, zrun):
    """Every launch shape the explicit plan can choose gives the oracle's
    result (the shape only splits the grid into workgroup tiles)."""
    T0, Om0, P = heat_case(**CASES[name])
    with api.HeatPlan(T0, Om0, P, scheme="explicit") as h:
        h.set_explicit_zrun(zrun)
        assert h.explicit_zrun == zrun
        h.advance(P["steps"], P["dt"], P["lightsOn"])
        r = h.download(f64=True)
    ref = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP64_EXPLICIT)
    assert rel(r["T_f64"], ref["T_f64"]) <= 1e-9
    assert rel(r["Omega_f64"], ref["Omega_f64"]) <= 1e-9
