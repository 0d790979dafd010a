"""Parity on the benchmark workloads themselves, at their own sizes.

Each test builds one of the five benchmark configurations (layered skin with
a blood vessel; tests/../mcmatlab_amd/workloads.py) and checks the GPU
against the FP64 CPU oracle (OpenMP build, on the cores this process may
use) or against the plain GPU launch an item must reproduce:

  layered100   100^3 matched layers, pencil beam, 1e6 packets: 10 seeds per
               side; every scalar summary within 3 sigma (Welch), per-voxel
               NFR within its standard error.
  layered200   200^3 with Fresnel interfaces, top-hat beam with a 5 degree
               Gaussian spread, 1e6 packets x 10 seeds per side (the
               headline workload at 1/100 of its packets): the same bars.
  heat200      2000 FP64 explicit steps (667 lights on, 1333 off) on the
               200^3 grid, heated by the GPU fluence of layered200: T, Omega
               and the sensors within 1e-9 relative of the oracle; the
               reference's own DG-ADI scheme on the same grid for 40 steps.
  spectral150  64 wavelengths on 150^3: four wavelengths as batched items
               against the oracle (10 seeds per side), and all 64 at their
               full budgets (1.4e8 packets) in one batched launch against 64
               plain launches with the same packet streams.
  blob128      absorbed-power gradients on the 128^3 blob map, 1e7 packets:
               three single voxels and one direction against central
               differences at h = 1e-3 relative with common random numbers
               (2 %); the 3000-step heat adjoint on 128^3 (checkpoints every
               100) against Richardson-extrapolated central differences.
  sweep32      the 32 grids of 64^3..247^3 at their full budgets in one
               multi-grid launch, every item against the plain launch of its
               plan.

Oracle runs with several threads use dSFMT streams seeded seed+thread, so
the seeds of independent runs are 1000 apart.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_lib as OL
from mcstats import check_scalars_two_stage, compare_voxels

pytestmark = pytest.mark.gpu

SEEDS = 10
NTH = OL.host_threads()
_SHARED = {}  # the layered200 GPU fluence heats the heat200 runs
OUT_KEYS = ("NFR", "farField", "NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg", "nPhotons")


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def gpu_run(api, m, seed):
    mc = api.MCmatlab(m, 1, seed=seed).MC
    d = {k: getattr(mc, k) for k in OUT_KEYS}
    return {k: v for k, v in d.items() if not (isinstance(v, float) and math.isnan(v))}


def oracle_run(m, s):
    return OL.oracle_mc(m, seed=1000 * s, nthreads=NTH)


def mc_parity(api, m, name, gpu_seed0):
    print(f"[gpu-config] {name}: {SEEDS} GPU runs, {SEEDS} oracle runs on {NTH} threads", flush=True)
    G = [gpu_run(api, m, gpu_seed0 + s) for s in range(SEEDS)]
    O = [oracle_run(m, s) for s in range(SEEDS)]
    nph = int(m.MC.nPhotonsRequested)
    assert all(g["nPhotons"] == nph for g in G) and all(o["nPhotons"] == nph for o in O)
    bad = check_scalars_two_stage(lambda s: gpu_run(api, m, gpu_seed0 + s), lambda s: oracle_run(m, s), m, seeds=SEEDS,
                                  extra=SEEDS, runs=(G, O))
    assert not bad, f"{name}: scalar summaries outside 3 sigma: {bad}"
    frac, mt2, n = compare_voxels(G, O)
    print(f"[gpu-config] {name}: {n} voxels compared, {frac:.4f} beyond 3 sigma, mean t^2 {mt2:.3f}", flush=True)
    assert n > 1000
    assert frac < 0.02 and mt2 < 1.5, (frac, mt2, n)
    return G


def test_layered100_pencil_matches_oracle(api):
    from mcmatlab_amd.workloads import layered_tissue_mc

    m = layered_tissue_mc(n=100, matched=True, beam="pencil", nPhotons=1e6)
    mc_parity(api, m, "layered100", 5000)


@pytest.fixture(scope="module")
def layered200(api):
    from mcmatlab_amd.workloads import layered_tissue_mc

    m = layered_tissue_mc(n=200, matched=False, beam="tophat", nPhotons=1e6)
    return m


def test_layered200_fresnel_tophat_matches_oracle(api, layered200):
    G = mc_parity(api, layered200, "layered200", 6000)
    _SHARED["layered200_nfr"] = np.mean([np.asarray(g["NFR"], np.float64) for g in G], axis=0).astype(np.float32)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def heat200_inputs(layered200):
    from mcmatlab_amd.workloads import layered_tissue_heat

    nfr = _SHARED.get("layered200_nfr")
    return layered_tissue_heat(n=200, NFR=nfr)


def test_heat200_explicit_2000_steps_matches_oracle(api, layered200):
    T0, Om0, Pon, Poff = heat200_inputs(layered200)
    print(f"[gpu-config] heat200: 667 + 1333 FP64 explicit steps, oracle on {NTH} threads", flush=True)
    api.finiteElementHeatPropagator(T0, Om0, Pon, scheme="explicit")
    st = api.finiteElementHeatPropagator.last_state
    T, Om, sens, mx = api.finiteElementHeatPropagator(T0, Om0, Poff, scheme="explicit", state_f64=st)
    T64, Om64 = api.finiteElementHeatPropagator.last_state
    r1 = OL.oracle_heat(T0, Om0, Pon, variant=OL.HEAT_FP64_EXPLICIT, nthreads=NTH)
    r2 = OL.oracle_heat(T0, Om0, Poff, variant=OL.HEAT_FP64_EXPLICIT, T_f64=r1["T_f64"], Omega_f64=r1["Omega_f64"], nthreads=NTH)
    eT, eO = rel(T64, r2["T_f64"]), rel(Om64, r2["Omega_f64"])
    print(f"[gpu-config] heat200: max rel err T {eT:.2e} Omega {eO:.2e}", flush=True)
    assert eT <= 1e-9 and eO <= 1e-9
    assert float(np.max(Om64)) > 0 and float(np.max(T64)) > 37.5  # the run heats and damages
    np.testing.assert_allclose(sens, r2["sensorTemps"], rtol=1e-9)
    np.testing.assert_allclose(mx, r2["maxMediaTemps"], rtol=1e-7)


def test_heat200_dgadi_matches_oracle(api, layered200):
    T0, Om0, Pon, _ = heat200_inputs(layered200)
    P = dict(Pon, steps=40)
    api.finiteElementHeatPropagator(T0, Om0, P, scheme="dgadi")
    T64, Om64 = api.finiteElementHeatPropagator.last_state
    r = OL.oracle_heat(T0, Om0, P, variant=OL.HEAT_FP64_DGADI, nthreads=NTH)
    assert rel(T64, r["T_f64"]) <= 1e-9 and rel(Om64, r["Omega_f64"]) <= 1e-9


def single_wavelength(m, iL):
    from mcmatlab_amd.model import media_table

    mo = m.copy()
    mp = m.MC.mediaProperties
    col = lambda k: np.asarray(mp[k], np.float64)[:, iL]
    mo.MC.mediaProperties = media_table(col("mua"), col("mus"), col("g"), col("n"))
    mo.MC.spectrum = np.array([1.0])
    return mo


SPECTRAL_PICK = (0, 21, 42, 63)
SPECTRAL_N = 200_000


def test_spectral150_batched_wavelengths_match_oracle(api):
    from mcmatlab_amd.workloads import layered_tissue_spectral_mc

    m, _ = layered_tissue_spectral_mc(n=150, nL=64)
    plan = api.MCPlan(m)
    try:
        items = [(iL, SPECTRAL_N, 7000 + s, iL) for iL in SPECTRAL_PICK for s in range(SEEDS)]
        got = plan.batch(items)

        def more(iL, s):
            return plan.batch([(iL, SPECTRAL_N, 7000 + s, iL)])[0]

        for k, iL in enumerate(SPECTRAL_PICK):
            print(f"[gpu-config] spectral150: wavelength {iL}", flush=True)
            mo = single_wavelength(m, iL)
            mo.MC.nPhotonsRequested = SPECTRAL_N
            G = got[k * SEEDS:(k + 1) * SEEDS]
            O = [oracle_run(mo, s) for s in range(SEEDS)]
            assert all(g["nPhotons"] == SPECTRAL_N for g in G)
            bad = check_scalars_two_stage(lambda s, iL=iL: more(iL, s), lambda s, mo=mo: oracle_run(mo, s), mo, seeds=SEEDS,
                                          extra=SEEDS, runs=(G, O))
            assert not bad, (iL, bad)
            frac, mt2, n = compare_voxels(G, O)
            assert n > 500 and frac < 0.02 and mt2 < 1.5, (iL, frac, mt2, n)
    finally:
        plan.close()


def test_spectral150_full_batch_equals_plain_launches(api):
    """All 64 wavelengths at their budgets (round(10^U(5,7)), rng 1): the
    batched launch reproduces the 64 plain launches packet for packet."""
    from mcmatlab_amd.workloads import layered_tissue_spectral_mc
    from test_gpu_batch import check_equal

    m, budgets = layered_tissue_spectral_mc(n=150, nL=64)
    items = [(iL, n, 11, iL) for iL, n in enumerate(budgets)]
    plan = api.MCPlan(m)
    try:
        got = plan.batch(items)
        for (iL, n, seed, _), g in zip(items, got):
            assert g["nPhotons"] == n
            plan.simulate(n, seed=seed, iL=iL, firstPhoton=0)
            ref = plan.normalize(iL)
            check_equal(g, (ref, None), iL)
    finally:
        plan.close()


@pytest.fixture(scope="module")
def blob128(api):
    from mcmatlab_amd.workloads import blob_absorber_mc

    m, mua, box = blob_absorber_mc(n=128)
    plan = api.MCPlan(m)
    yield plan, mua, box
    plan.close()


BLOB_N = 10_000_000
BLOB_SEEDS = (31, 32, 33, 34, 35, 36)
H_REL = 1e-3


def test_blob128_gradients_match_central_differences(blob128):
    """dJ/dmua of single voxels and along a direction against
    (J(mua + h) - J(mua - h)) / 2h with h = 1e-3 mua and the same packet
    streams on both sides (common random numbers), 1e7 packets per seed.

    With common random numbers the difference quotient and the pathwise
    gradient see the same packets and differ only where a packet's roulette
    decision flips between mua + h and mua - h, so the per-seed differences
    measure that noise: the three voxels J is most sensitive to, and the
    direction of the map itself, agree within 2 % on the mean over the
    seeds.  The most sensitive voxel outside the target box (ten times
    weaker, only the path-length term) sits on the beam's ballistic path:
    there h = 1e-3 mua changes a packet's FP32 weight by ~2e-6 relative,
    which single precision resolves to only a few per cent -- the same
    rounding for every packet on that deterministic path, so it does not
    average out -- and its difference quotient takes h = 2e-2 mua; it agrees
    within 2 % or four standard errors of the per-seed differences."""
    plan, mua, box = blob128
    runs = [plan.absorbed_power(box, BLOB_N, seed=s, muaMap=mua) for s in BLOB_SEEDS]
    gs = np.array([r["dJ_dmua"] for r in runs])
    g = gs.mean(axis=0)
    assert all(r["J"] > 0 for r in runs) and np.isfinite(g).all()

    def fd(direction, h):
        out = []
        for s in BLOB_SEEDS:
            Jp = plan.absorbed_power(box, BLOB_N, seed=s, muaMap=mua + h * direction, grad=False)["J"]
            Jm = plan.absorbed_power(box, BLOB_N, seed=s, muaMap=mua - h * direction, grad=False)["J"]
            out.append((Jp - Jm) / (2 * h))
        return np.array(out)

    def check(name, per_seed_grad, per_seed_fd, allow_se):
        diff = per_seed_fd - per_seed_grad
        gm = float(per_seed_grad.mean())
        se = float(diff.std(ddof=1) / math.sqrt(len(diff)))
        rel = float(diff.mean()) / gm
        print(f"[gpu-config] blob128 {name}: grad {gm:.6e} fd {per_seed_fd.mean():.6e} ({rel:+.2%}, se {se / abs(gm):.2%})", flush=True)
        assert abs(diff.mean()) <= max(0.02 * abs(gm), 4 * se if allow_se else 0.0), (name, rel, se / abs(gm))

    sens = np.abs(g * mua)
    (x0, x1), (y0, y1), (z0, z1) = box
    inbox = np.zeros(mua.shape, bool)
    inbox[x0:x1, y0:y1, z0:z1] = True
    order = np.argsort(sens.ravel())[::-1][:3]
    picks = [tuple(int(i) for i in np.unravel_index(f, g.shape)) for f in order]
    for v in picks:
        e = np.zeros(mua.shape, np.float32)
        e[v] = 1.0
        check(f"voxel {v}", gs[(slice(None),) + v], fd(e, H_REL * float(mua[v])), False)
    vo = tuple(int(i) for i in np.unravel_index(np.argmax(np.where(inbox, 0, sens)), g.shape))
    e = np.zeros(mua.shape, np.float32)
    e[vo] = 1.0
    check(f"voxel {vo} (outside the box, h = 2e-2)", gs[(slice(None),) + vo], fd(e, 2e-2 * float(mua[vo])), True)
    # along the map itself: d/dt J(mua (1 + t)) at t = 0
    check("direction mua", (gs * mua).sum(axis=(1, 2, 3)), fd(mua, H_REL), False)


def test_blob128_heat_adjoint_3000_steps_matches_central_differences(api):
    """d(final blood damage)/d(heating rate) and /d(initial temperature) of
    3000 FP64 DG-ADI steps on 128^3, checkpoints every 100 steps."""
    from mcmatlab_amd.workloads import layered_tissue_heat

    steps, dt = 3000, 5e-3 / 667
    T0, Om0, P, _ = layered_tissue_heat(n=128, steps_on=steps, steps_off=1)
    segs = [(steps, dt, True)]
    r = api.heat_adjoint(T0, Om0, P, segments=segs, checkpointEvery=100)
    assert r["J"] > 0
    Q = np.asarray(P["dTdt_abs"], np.float64)

    def D(h):
        out = []
        for f in (1 + h, 1 - h):
            P2 = dict(P, dTdt_abs=(Q * f).astype(np.float32))
            out.append((api.heat_adjoint(T0, Om0, P2, segments=segs, grads=False)["J"], np.asarray(P2["dTdt_abs"], np.float64)))
        (jp, qp), (jm, qm) = out
        return jp - jm, float(np.sum(r["dJ_dQ"] * (qp - qm)))

    (dJ1, ad1), (dJ2, ad2) = D(4e-3), D(2e-3)
    rich = (4 * (dJ2 / ad2) - dJ1 / ad1) / 3
    print(f"[gpu-config] adjoint128 dQ: fd/adjoint {dJ1 / ad1:.8f} {dJ2 / ad2:.8f} Richardson {rich:.8f}", flush=True)
    assert abs(rich - 1) <= 1e-4
    rng = np.random.default_rng(8)
    d = rng.standard_normal(T0.shape)
    h = 1e-3
    base = T0.astype(np.float64)
    Jp = api.heat_adjoint(T0, Om0, P, segments=segs, grads=False, T_f64=base + h * d)["J"]
    Jm = api.heat_adjoint(T0, Om0, P, segments=segs, grads=False, T_f64=base - h * d)["J"]
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(r["dJ_dT0"] * d))
    print(f"[gpu-config] adjoint128 dT0: fd {fd:.8e} adjoint {ad:.8e}", flush=True)
    assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-12 * r["J"], (fd, ad)


def test_sweep32_multi_grid_items_equal_plain_launches(api):
    from mcmatlab_amd.workloads import multi_grid_sweep
    from test_gpu_batch import check_equal

    sweep = multi_grid_sweep(32, seed=3)
    plans = [api.MCPlan(s["model"]) for s in sweep]
    try:
        items = [(k, 0, s["nPhotons"], 40 + k, 0) for k, s in enumerate(sweep)]
        print(f"[gpu-config] sweep32: {sum(s['nPhotons'] for s in sweep)} packets on 32 grids", flush=True)
        got = api.multi_run(plans, items)
        for (k, iL, n, seed, _), g in zip(items, got):
            assert g["nPhotons"] == n
            assert g["NFR"].shape[:3] == sweep[k]["model"].MC.M.shape
            plans[k].simulate(n, seed=seed, iL=iL, firstPhoton=0)
            check_equal(g, (plans[k].normalize(iL), None), iL)
    finally:
        for p in plans:
            p.close()
