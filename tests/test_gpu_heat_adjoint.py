"""Discrete adjoint of the FP64 DG-ADI heat propagation
(mcm_heat_adjoint_run).  No reference counterpart: pinned by the
discrete-adjoint identity <dJ/dx, dx> = dJ against central differences of
the same FP64 forward propagation (the forward itself is pinned to the
oracle in test_gpu_heat.py).

- dJ/dT0 along a random direction, with FP64 initial states: 1e-6.
- dJ/dQ along the uniform scaling of the heating rate (a single-precision
  input, so the differences are Richardson-extrapolated): 1e-4.
- dJ/d dTdtperdeltaT (the conduction coefficients, through which the
  thermal conductivities enter) along a random direction of the table
  (Richardson-extrapolated central differences of the representable
  single-precision perturbation): 1e-5; dJ/dTC through the chain rule and
  the torch.autograd.Function against central differences in TC: 1e-3.
- checkpoint segments of 5 steps over 23 steps (a partial last segment),
  fixed-temperature and insulating boundaries, custom target masks;
  lights-off schedules give dJ/dQ = 0.
"""
from __future__ import annotations

import numpy as np
import pytest

from mcmatlab_amd.media import heat_coefficients
from mcmatlab_amd.workloads import layered_tissue_heat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def case(bt=1, n=16, shape=(14, 12, 10), seed=0):
    T0, Om0, Pon, _ = layered_tissue_heat(n=n, steps_on=10, steps_off=10)
    P = dict(Pon)
    P["heatBoundaryType"] = bt
    d = 0.1 / n
    P["dTdtperdeltaT"] = heat_coefficients([1.2e-3, 3.7, 3.9, 3.8], [2.6e-4, 0.0037, 0.0052, 0.0045], d, d, d)
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    M0 = np.asarray(P["M"])
    reps = [-(-e // m) for e, m in zip(shape, M0.shape)]
    M = np.tile(M0, reps)[:nx, :ny, :nz].copy(order="F")
    M[rng.random(M.shape) < 0.25] = 3  # more blood (E != 0)
    P["M"] = M
    P["dTdt_abs"] = (3e4 * rng.random((nx, ny, nz))).astype(np.float32)
    P["tempSensorCornerIdxs"] = np.zeros(0)
    P["tempSensorInterpWeights"] = np.zeros((0, 3))
    P["dt"] = 2e-5
    T0 = (45 + 10 * rng.random((nx, ny, nz))).astype(np.float32)
    Om0 = np.zeros((nx, ny, nz), np.float32)
    segs = [(15, 2e-5, True), (8, 3e-5, False)]
    return T0, Om0, P, segs


@pytest.mark.parametrize("bt", [0, 1])
def test_adjoint_dT0_matches_central_differences(api, bt):
    T0, Om0, P, segs = case(bt)
    r = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=5)
    assert r["J"] > 0
    rng = np.random.default_rng(1)
    d = rng.standard_normal(T0.shape)
    h = 1e-3
    base = T0.astype(np.float64)
    Jp = api.heat_adjoint(T0, Om0, P, segments=segs, grads=False, T_f64=base + h * d)["J"]
    Jm = api.heat_adjoint(T0, Om0, P, segments=segs, grads=False, T_f64=base - h * d)["J"]
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(r["dJ_dT0"] * d))
    assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-12 * r["J"], (fd, ad)


@pytest.mark.parametrize("bt", [0, 1])
def test_adjoint_dQ_matches_central_differences(api, bt):
    T0, Om0, P, segs = case(bt, seed=3)
    r = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=5)
    Q = np.asarray(P["dTdt_abs"], np.float64)

    def J_scaled(f):
        P2 = dict(P)
        P2["dTdt_abs"] = (Q * f).astype(np.float32)
        q32 = np.asarray(P2["dTdt_abs"], np.float64)
        return api.heat_adjoint(T0, Om0, P2, segments=segs, grads=False)["J"], q32

    def D(h):
        (jp, qp), (jm, qm) = J_scaled(1 + h), J_scaled(1 - h)
        # directional derivative along the actually representable perturbation
        return (jp - jm), qp - qm

    dJ1, dq1 = D(4e-3)
    dJ2, dq2 = D(2e-3)
    ad1, ad2 = float(np.sum(r["dJ_dQ"] * dq1)), float(np.sum(r["dJ_dQ"] * dq2))
    # dJ / ad = 1 + c h^2 + O(h^4) for central differences: Richardson
    rich = (4 * (dJ2 / ad2) - dJ1 / ad1) / 3
    assert abs(rich - 1) <= 1e-4, (dJ1 / ad1, dJ2 / ad2, rich)


def fd_check_shape(api, shape):
    """The adjoint's dJ/dT0 along a random direction against central
    differences of J on a block of the given shape."""
    T0, Om0, P, segs = case(1, shape=shape, seed=5)
    segs = [(6, 2e-5, True), (3, 3e-5, False)]
    r = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=4)
    rng = np.random.default_rng(2)
    d = rng.standard_normal(T0.shape)
    h = 1e-3
    base = T0.astype(np.float64)
    Jp = api.heat_adjoint(T0, Om0, P, segments=segs, grads=False, T_f64=base + h * d)["J"]
    Jm = api.heat_adjoint(T0, Om0, P, segments=segs, grads=False, T_f64=base - h * d)["J"]
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(r["dJ_dT0"] * d))
    assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-12 * r["J"], (fd, ad)


@pytest.mark.parametrize("shape", [(520, 4, 3), (3, 600, 4), (4, 3, 513), (5, 512, 3)])
def test_adjoint_long_lines(api, shape):
    """Lines beyond one wavefront's reach (more than 512 voxels) take the
    lane-per-line recurrence; 512 is the longest wavefront solve."""
    fd_check_shape(api, shape)


@pytest.mark.parametrize("shape", [(40, 7, 5), (100, 5, 3), (200, 3, 5), (5, 40, 7), (3, 100, 5), (7, 5, 200), (3, 5, 129)])
def test_adjoint_segment_shapes(api, shape):
    """Lines of 33..256 voxels are solved on 16- or 32-lane segments, several
    lines per wavefront; odd line counts leave some segments without a line."""
    print(f"[gpu-adjoint] shape {shape}", flush=True)
    fd_check_shape(api, shape)


def test_adjoint_lights_off_has_no_source_gradient(api):
    T0, Om0, P, _ = case(1)
    r = api.heat_adjoint(T0, Om0, P, segments=[(12, 2e-5, False)], checkpointEvery=4)
    assert np.all(r["dJ_dQ"] == 0)
    assert np.any(r["dJ_dT0"] != 0)


def test_adjoint_target_mask_and_checkpoint_period(api):
    T0, Om0, P, segs = case(0, seed=7)
    M = np.asarray(P["M"])
    blood = (M == 3)
    a = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=5)
    b = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=100, targetMask=blood)
    assert np.isclose(a["J"], b["J"], rtol=1e-12)
    np.testing.assert_allclose(a["dJ_dQ"], b["dJ_dQ"], rtol=1e-9, atol=1e-30)
    half = blood.copy()
    half[: half.shape[0] // 2] = False
    c = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=7, targetMask=half)
    assert 0 < c["J"] < a["J"]
    assert np.isclose(c["J"], float(np.sum(a["Omega"][half])), rtol=1e-12)


def test_adjoints_in_threads_equal_one_at_a_time(api):
    """heat_adjoints runs independent models from several host threads; each
    result equals the same call made alone (same kernels, same order per
    plan: bit for bit)."""
    cases = []
    for k, shape in enumerate([(14, 12, 10), (9, 16, 11), (12, 12, 12)]):
        T0, Om0, P, segs = case(bt=k % 2, shape=shape, seed=10 + k)
        cases.append((T0, Om0, P, segs))
    alone = [api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=5) for T0, Om0, P, segs in cases]
    many = api.heat_adjoints(cases, workers=3, checkpointEvery=5)
    for a, b in zip(alone, many):
        assert a["J"] == b["J"]
        assert np.array_equal(a["dJ_dQ"], b["dJ_dQ"]) and np.array_equal(a["dJ_dT0"], b["dJ_dT0"])


@pytest.mark.parametrize("bt", [0, 1])
def test_adjoint_dc_matches_central_differences(api, bt):
    T0, Om0, P, segs = case(bt, seed=11)
    r = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=5, coefGrad=True)
    g = r["dJ_dc"]
    assert g.shape == (4, 4, 3) and np.all(np.isfinite(g)) and np.any(g != 0)
    c = np.asarray(P["dTdtperdeltaT"], np.float64)
    rng = np.random.default_rng(4)
    d = rng.standard_normal(c.shape)

    def D(h):
        out = []
        for sgn in (1, -1):
            cp = (c * (1 + sgn * h * d)).astype(np.float32)
            out.append((api.heat_adjoint(T0, Om0, dict(P, dTdtperdeltaT=cp), segments=segs, grads=False)["J"], cp.astype(np.float64)))
        (jp, cp), (jm, cm) = out
        return jp - jm, float(np.sum(g * (cp - cm)))

    (dJ1, ad1), (dJ2, ad2) = D(4e-3), D(2e-3)
    rich = (4 * (dJ2 / ad2) - dJ1 / ad1) / 3
    assert abs(rich - 1) <= 1e-5, (dJ1 / ad1, dJ2 / ad2, rich)


def test_adjoint_conductivity_torch_function(api):
    """dJ/dTC per medium through heat_damage_function (torch) against central
    differences of J in TC, and dJ/dQ, dJ/dT0 equal the plain adjoint's."""
    torch = pytest.importorskip("torch")
    from mcmatlab_amd.media import heat_coefficients

    T0, Om0, P, segs = case(1, seed=13)
    n = 16
    d = 0.1 / n
    VHC = np.array([1.2e-3, 3.7, 3.9, 3.8])
    TC = np.array([2.6e-4, 0.0037, 0.0052, 0.0045])
    setup = {"P": P, "VHC": VHC, "d": (d, d, d), "segments": segs, "checkpointEvery": 5}
    HD = api.heat_damage_function()
    q = torch.tensor(np.asarray(P["dTdt_abs"], np.float64), requires_grad=True)
    t0 = torch.tensor(T0.astype(np.float64), requires_grad=True)
    tc = torch.tensor(TC, requires_grad=True)
    J = HD.apply(q, t0, tc, setup)
    J.backward()
    ref = api.heat_adjoint(T0, Om0, dict(P, dTdtperdeltaT=heat_coefficients(VHC, TC, d, d, d)), segments=segs, checkpointEvery=5,
                           T_f64=T0.astype(np.float64))
    assert abs(float(J) - ref["J"]) <= 1e-12 * ref["J"]
    np.testing.assert_allclose(q.grad.numpy(), ref["dJ_dQ"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(t0.grad.numpy(), ref["dJ_dT0"], rtol=1e-12, atol=0)
    for m in range(4):
        h = 1e-3 * TC[m]
        Js = []
        for sgn in (1, -1):
            tcm = TC.copy()
            tcm[m] += sgn * h
            Js.append(api.heat_adjoint(T0, Om0, dict(P, dTdtperdeltaT=heat_coefficients(VHC, tcm, d, d, d)), segments=segs, grads=False,
                                       T_f64=T0.astype(np.float64))["J"])
        fd = (Js[0] - Js[1]) / (2 * h)
        gm = float(tc.grad[m])
        assert abs(fd - gm) <= 1e-3 * abs(gm) + 1e-6 * float(np.max(np.abs(tc.grad.numpy()))), (m, fd, gm)
