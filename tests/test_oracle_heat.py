"""Pinning the heat oracle.

1. FP32 Douglas-Gunn ADI bit for bit against the reference C
   (+MCmatlab/src/finiteElementHeatPropagator.c compiled single-threaded in
   oracle/_ref): T, Omega, sensor temperatures and per-medium maxima, over
   boundary types, lights, damage on/off, heterogeneous media, odd grids and
   1..50 steps.  Skipped where the reference build is absent.
2. FP64 DG-ADI and FP64 explicit against closed forms
   (tests/heat_closed_forms.py).
3. FP64 DG-ADI against FP32 DG-ADI: same scheme, so they agree to FP32
   precision.
"""
import numpy as np
import pytest

import heat_closed_forms as HC
import oracle_lib as OL

ref_only = pytest.mark.skipif(not OL.ref_available(), reason="reference build (oracle/_ref) not present")

CASES = {
    "sinked_on": dict(),
    "insulated_on": dict(bt=0),
    "lights_off": dict(lights=False),
    "no_damage": dict(damage=False),
    "one_step": dict(steps=1),
    "fifty_steps": dict(steps=50),
    "homogeneous": dict(hetero=False),
    "odd_grid": dict(shape=(13, 9, 11)),
    "flat_grid": dict(shape=(16, 16, 3), bt=0),
}


def heat_case(**kw):
    from test_gpu_heat import heat_case as hc

    return hc(**kw)


@pytest.fixture(scope="module")
def ref_lib():
    from mexdriver import MexLib

    return MexLib(OL.REF_HEAT_LIB)


@ref_only
@pytest.mark.parametrize("name", list(CASES))
def test_fp32_oracle_bitwise_equals_reference(ref_lib, name):
    from mexdriver import run_heat_mex

    T0, Om0, P = heat_case(**CASES[name])
    r = run_heat_mex(ref_lib, T0, Om0, P)
    o = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP32_DGADI)
    for k in ("T", "Omega", "sensorTemps", "maxMediaTemps"):
        a = np.asarray(r[k]).ravel(order="F")
        b = np.asarray(o[k]).ravel(order="F")
        assert a.shape == b.shape, k
        assert np.array_equal(a, b, equal_nan=True), k


@ref_only
def test_reference_zero_steps_quirk(ref_lib):
    """With steps == 0 the reference returns a zero-filled T (:703-705: the
    step loop never writes the output); the oracle restates it."""
    from mexdriver import run_heat_mex

    T0, Om0, P = heat_case()
    P["steps"] = 0
    r = run_heat_mex(ref_lib, T0, Om0, P)
    o = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP32_DGADI)
    assert np.all(r["T"] == 0) and np.all(o["T"] == 0)
    assert r["sensorTemps"].shape == o["sensorTemps"].shape


def _run(variant):
    def run(T0, Om0, P, T_f64=None):
        return OL.oracle_heat(T0, Om0, P, variant=variant, T_f64=T_f64)

    return run


SCHEMES = {"dgadi": OL.HEAT_FP64_DGADI, "explicit": OL.HEAT_FP64_EXPLICIT}


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_fp64_uniform(scheme):
    HC.uniform(_run(SCHEMES[scheme]), scheme)


@pytest.mark.parametrize("lights", [False, True])
@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_fp64_conservation(scheme, lights):
    HC.conservation(_run(SCHEMES[scheme]), scheme, lights)


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_fp64_eigenmode(scheme):
    HC.eigenmode(_run(SCHEMES[scheme]), scheme)


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_fp64_frozen_boundary(scheme):
    HC.frozen(_run(SCHEMES[scheme]))


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_fp64_arrhenius(scheme):
    HC.arrhenius(_run(SCHEMES[scheme]))


def test_fp32_eigenmode():
    """The FP32 variant is the same scheme: the eigenmode holds to FP32 precision."""
    def run(T0, Om0, P):
        r = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP32_DGADI)
        r["T_f64"] = r["T"].astype(np.float64)
        return r

    import math

    n = (12, 10, 8)
    P = HC._params(n, d=0.005, steps=1, dt=5e-2)
    c = float(np.asarray(P["dTdtperdeltaT"], dtype=np.float64)[0, 0, 0])
    k = (1, 2, 1)
    v = np.ones(n)
    lam = []
    for a in range(3):
        sh = [1, 1, 1]
        sh[a] = n[a]
        v = v * np.cos(math.pi * k[a] * (np.arange(n[a]) + 0.5) / n[a]).reshape(sh)
        lam.append(-4 * c * math.sin(math.pi * k[a] / (2 * n[a])) ** 2)
    dt = 5e-2
    lx, ly, lz = lam
    g = (1 + dt * (lx / 2 + ly + lz)) / (1 - dt * lx / 2)
    g = (g - dt * ly / 2) / (1 - dt * ly / 2)
    g = (g - dt * lz / 2) / (1 - dt * lz / 2)
    r = run(v.astype(np.float32), np.zeros(n, np.float32), P)
    np.testing.assert_allclose(r["T_f64"], g * v.astype(np.float32), atol=2e-6)


def test_fp64_dgadi_close_to_fp32():
    T0, Om0, P = heat_case()
    a = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP64_DGADI)
    b = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP32_DGADI)
    np.testing.assert_allclose(a["T_f64"], b["T"], rtol=1e-5)
