"""Absorbed-power gradients (mcm_mc_grad_run).  No reference counterpart:
pinned by a closed form and by central finite differences with common
random numbers (the same seed, hence the same packet streams, at mua +- h).

- Beer-Lambert: no scattering, pencil beam through a uniform absorber of
  thickness L: J = 1 - exp(-mua L) over the whole grid and
  dJ/dmua_v = dz exp(-mua L) on every voxel of the beam column (0 elsewhere);
  the paths are deterministic, so this holds to FP32 rounding.
- Blob absorber (per-voxel mua map): dJ/dmua of the strongest voxel and a
  directional derivative over a block of voxels against central differences
  of J with common random numbers (mua changes weights only, so the two
  differ only by roulette decisions and O(h^2)): 2 %.
- dJ/dlog(mus) (likelihood ratio) against central differences of J over a
  common mus scale, within 4 combined standard errors.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from mcmatlab_amd.model import Geometry, Model, media_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def slab(n=20, mua=2.0):
    d = 0.1 / n
    m = Model()
    m.G = Geometry(dx=d, dy=d, dz=d)
    m.MC.M = np.ones((n, n, n), np.uint8)
    m.MC.mediaProperties = media_table([mua], [1e-8], [1.0], [1.0])
    m.MC.matchedInterfaces = True
    m.MC.interfaceNormals = np.float32(np.nan)
    m.MC.boundaryType = 1
    m.MC.spectrum = np.array([1.0])
    m.MC.LS.sourceType = 0
    m.MC.nPhotonsRequested = 1000
    return m


def test_grad_beer_lambert_closed_form(api):
    n, mua = 20, 2.0
    m = slab(n, mua)
    plan = api.MCPlan(m)
    try:
        r = plan.absorbed_power(((0, n), (0, n), (0, n)), 2000, seed=1)
        rm = plan.absorbed_power(((0, n), (0, n), (0, n)), 2000, seed=1, muaMap=np.full((n, n, n), mua, np.float32))
    finally:
        plan.close()
    L = 0.1
    assert math.isclose(r["J"], 1 - math.exp(-mua * L), rel_tol=1e-5)
    assert math.isclose(rm["J"], r["J"], rel_tol=1e-6)
    g = r["dJ_dmua"]
    col = g[n // 2, n // 2, :]
    np.testing.assert_allclose(col, (L / n) * math.exp(-mua * L), rtol=1e-4)
    off = g.copy()
    off[n // 2, n // 2, :] = 0
    assert np.max(np.abs(off)) <= 1e-12
    assert math.isclose(g.sum(), L * math.exp(-mua * L), rel_tol=1e-4)


@pytest.fixture(scope="module")
def blobs(api):
    from mcmatlab_amd.workloads import blob_absorber_mc

    m, mua, box = blob_absorber_mc(n=48)
    plan = api.MCPlan(m)
    yield plan, mua, box
    plan.close()


N_FD = 300000


def test_grad_mua_matches_central_differences(blobs):
    plan, mua, box = blobs
    r = plan.absorbed_power(box, N_FD, seed=5, muaMap=mua)
    g = r["dJ_dmua"]
    assert r["J"] > 0 and np.isfinite(g).all()
    # strongest single voxel
    v = np.unravel_index(np.argmax(np.abs(g)), g.shape)
    h = 0.02 * float(mua[v])
    up, dn = mua.copy(), mua.copy()
    up[v] += h
    dn[v] -= h
    fd = (plan.absorbed_power(box, N_FD, seed=5, muaMap=up, grad=False)["J"] -
          plan.absorbed_power(box, N_FD, seed=5, muaMap=dn, grad=False)["J"]) / (2 * h)
    assert abs(fd - g[v]) <= 0.02 * abs(g[v]), (fd, g[v])
    # directional derivative over an 8^3 block around the box
    rng = np.random.default_rng(3)
    d = np.zeros_like(mua)
    (x0, x1), (y0, y1), (z0, z1) = box
    d[x0 - 2:x0 + 6, y0 - 2:y0 + 6, max(z0 - 4, 0):z0 + 4] = rng.choice([-1.0, 1.0], size=(8, 8, 8))
    eps = 0.02
    fd = (plan.absorbed_power(box, N_FD, seed=5, muaMap=mua + eps * d, grad=False)["J"] -
          plan.absorbed_power(box, N_FD, seed=5, muaMap=mua - eps * d, grad=False)["J"]) / (2 * eps)
    dd = float(np.sum(g * d))
    assert abs(fd - dd) <= 0.02 * abs(dd) + 1e-3 * np.sum(np.abs(g * d)), (fd, dd)


def test_grad_logmus_matches_central_differences(api):
    from mcmatlab_amd.workloads import blob_absorber_mc

    m, mua, box = blob_absorber_mc(n=48)
    base = m.MC.mediaProperties["mus"].copy()
    h = 0.05
    fds = []
    lrs, lr_se = [], []
    for seed in range(4):
        plan = api.MCPlan(m)
        try:
            r = plan.absorbed_power(box, N_FD, seed=40 + seed, muaMap=mua, grad=False)
        finally:
            plan.close()
        lrs.append(r["dJ_dlogmus"])
        lr_se.append(r["dJ_dlogmus_se"])
        J = []
        for sgn in (1, -1):
            m.MC.mediaProperties["mus"] = base * (1 + sgn * h)
            plan = api.MCPlan(m)
            try:
                J.append(plan.absorbed_power(box, N_FD, seed=40 + seed, muaMap=mua, grad=False)["J"])
            finally:
                plan.close()
        m.MC.mediaProperties["mus"] = base
        fds.append((J[0] - J[1]) / (2 * h))
    lr = float(np.mean(lrs))
    se_lr = float(np.sqrt(np.sum(np.square(lr_se)))) / len(lrs)
    fd = float(np.mean(fds))
    se_fd = float(np.std(fds, ddof=1) / math.sqrt(len(fds)))
    assert abs(lr - fd) <= 4 * math.hypot(se_lr, se_fd) + 0.01 * abs(fd), (lr, se_lr, fd, se_fd)


def test_grad_torch_autograd(api, blobs):
    torch = pytest.importorskip("torch")
    plan, mua, box = blobs
    AP = api.absorbed_power_function()
    t = torch.tensor(mua, dtype=torch.float32, requires_grad=True)
    J = AP.apply(t, plan, box, 20000, 9)
    J.backward()
    r = plan.absorbed_power(box, 20000, seed=9, muaMap=mua)
    assert math.isclose(float(J), r["J"], rel_tol=1e-9)
    np.testing.assert_allclose(t.grad.numpy(), r["dJ_dmua"], rtol=1e-5, atol=1e-12)


def test_gradients_in_threads_equal_one_at_a_time(api):
    """absorbed_powers runs independent plans from several host threads; each
    result equals the same call made alone (up to the order of FP64 atomic
    sums)."""
    from mcmatlab_amd.workloads import blob_absorber_mc

    jobs, plans = [], []
    for n in (24, 32, 28):
        m, mua, box = blob_absorber_mc(n=n, seed=n)
        p = api.MCPlan(m)
        plans.append(p)
        jobs.append((p, {"box": box, "nPhotons": 20000, "seed": 9, "muaMap": mua}))
    try:
        alone = [p.absorbed_power(**kw) for p, kw in jobs]
        many = api.absorbed_powers(jobs, workers=3)
    finally:
        for p in plans:
            p.close()
    for a, b in zip(alone, many):
        assert math.isclose(a["J"], b["J"], rel_tol=1e-12)
        assert math.isclose(a["dJ_dlogmus"], b["dJ_dlogmus"], rel_tol=1e-12, abs_tol=1e-15)
        np.testing.assert_allclose(b["dJ_dmua"], a["dJ_dmua"], rtol=1e-9, atol=1e-12 * float(np.max(np.abs(a["dJ_dmua"]))))


def test_forward_only_skips_replay_with_same_J(api):
    """Without a requested mua gradient the replay pass is skipped; J and the
    likelihood-ratio dJ/dlog mus come from the first pass and are unchanged."""
    from mcmatlab_amd.workloads import blob_absorber_mc

    m, mua, box = blob_absorber_mc(n=32, seed=4)
    p = api.MCPlan(m)
    try:
        a = p.absorbed_power(box, 30000, seed=2, muaMap=mua)
        b = p.absorbed_power(box, 30000, seed=2, muaMap=mua, grad=False)
    finally:
        p.close()
    assert b["dJ_dmua"] is None
    assert math.isclose(a["J"], b["J"], rel_tol=1e-12)
    assert math.isclose(a["dJ_dlogmus"], b["dJ_dlogmus"], rel_tol=1e-12, abs_tol=1e-15)
