"""Pinning the heat oracle without the reference (whose MEX source needs
MATLAB's mex.h and mx libraries, absent here).

1. FP64 DG-ADI and FP64 explicit against closed forms
   (tests/heat_closed_forms.py): uniform field, conservation of sum(VHC T)
   over heterogeneous media, cosine eigenmodes (the scheme's amplification
   factor), frozen boundaries, the Arrhenius integral at constant and at the
   step's mean temperature, the series-conductance steady state of three
   media (coefficient indexing), trilinear sensors on a linear field.
2. The FP32 DG-ADI variant (the reference's arithmetic) against the same
   eigenmode, the series steady state and conservation to FP32 precision,
   and against the FP64 variant of the same scheme.
3. Every oracle value is independent of its OpenMP thread count.
"""
import numpy as np
import pytest

import heat_closed_forms as HC
import oracle_lib as OL

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


def _run(variant):
    def run(T0, Om0, P, T_f64=None):
        r = OL.oracle_heat(T0, Om0, P, variant=variant, T_f64=T_f64)
        if variant == OL.HEAT_FP32_DGADI:
            r["T_f64"] = np.asarray(r["T"], np.float64)
        return r

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


@pytest.mark.parametrize("check", ["arrhenius_midpoint", "steady_series", "sensors_linear"])
@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_fp64_more_closed_forms(check, scheme):
    getattr(HC, check)(_run(SCHEMES[scheme]), scheme)


def test_fp32_steady_series_and_conservation():
    """The FP32 variant's coefficient indexing and fluxes: the series steady
    state holds and sum(VHC T) is conserved to FP32 precision."""
    import math

    run = _run(OL.HEAT_FP32_DGADI)
    nx = 14
    n = (nx, 3, 3)
    VHC, TC = np.array([3.7, 1.2, 4.1]), np.array([0.0037, 0.0011, 0.0052])
    media = np.array([0] * 4 + [1] * 5 + [2] * 5, np.uint8)
    M = np.repeat(np.repeat(media[:, None, None], 3, axis=1), 3, axis=2).copy(order="F")
    P = HC._params(n, VHC=VHC, TC=TC, M=M, bt=1, steps=20, dt=1e-3)
    a, b = media[:-1], media[1:]
    G = 2 * TC[a] * TC[b] / (TC[a] + TC[b])
    R = np.concatenate([[0.0], np.cumsum(1 / G)])
    prof = 20.0 + 30.0 * R / R[-1]
    T0 = np.repeat(np.repeat(prof[:, None, None], 3, axis=1), 3, axis=2).astype(np.float32)
    r = run(T0, np.zeros(n, np.float32), P)
    np.testing.assert_allclose(r["T_f64"], T0, rtol=0, atol=5e-5)
    rng = np.random.default_rng(1)
    n = (10, 9, 8)
    M = rng.integers(0, 3, n).astype(np.uint8)
    P = HC._params(n, VHC=VHC, TC=TC, M=M, steps=9, dt=2e-3)
    T0 = (30 + 20 * rng.random(n)).astype(np.float32)
    r = run(T0, np.zeros(n, np.float32), P)
    vhc = VHC.astype(np.float32).astype(np.float64)[M]
    e0, e1 = np.sum(vhc * T0), np.sum(vhc * r["T_f64"])
    assert abs(e1 - e0) <= 1e-6 * e0 * math.sqrt(M.size) / 10, (e1 - e0)


@pytest.mark.parametrize("variant", [OL.HEAT_FP32_DGADI, OL.HEAT_FP64_DGADI, OL.HEAT_FP64_EXPLICIT])
def test_threads_do_not_change_values(variant):
    from mcmatlab_amd.workloads import layered_tissue_heat

    T0, Om0, Pon, _ = layered_tissue_heat(n=23, steps_on=7, steps_off=5)
    a = OL.oracle_heat(T0, Om0, Pon, variant=variant)
    b = OL.oracle_heat(T0, Om0, Pon, variant=variant, nthreads=4)
    for k, v in a.items():
        if v is not None:
            assert np.array_equal(np.asarray(v), np.asarray(b[k]), equal_nan=True), k


def test_zero_steps_quirk():
    """With steps == 0 the reference returns a zero-filled T
    (finiteElementHeatPropagator.c:703-705: the output array is created
    zero-filled and only the step loop writes it); the oracle restates it."""
    T0, Om0, P = heat_case()
    P["steps"] = 0
    o = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP32_DGADI)
    assert np.all(o["T"] == 0)
