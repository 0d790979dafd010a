"""The named workloads: layered skin with a blood vessel on a voxel cube.

Geometry (all sizes in cm, voxel centres as in geometry.m:130-138):
  air        z < 0.01                      mua 1e-8  mus 1e-8  g 1    n 1
  epidermis  0.01 <= z < 0.02              mua 20    mus 200   g 0.9  n 1 or 1.4
  dermis     z >= 0.02                     mua 0.5   mus 200   g 0.9  n 1 or 1.4
  blood      y^2 + (z - 0.05)^2 < 0.01^2   mua 230   mus 470   g 0.9  n 1 or 1.38
             (a cylinder along x; overrides the layers)

- layered_tissue_mc(n=100, matched=True):  100^3 pencil-beam run, all n = 1
  (the CPU-runnable case) -- 1e6 packets by default.
- layered_tissue_mc(n=200, matched=False, beam="tophat"): the 200^3 case with
  Fresnel refraction/reflection, a 0.03 cm top-hat focal-plane profile and a
  Gaussian angular profile of 5 degrees 1/e^2 half-angle -- 1e8 packets.
- layered_tissue_spectral_mc(n=150, nL=64): the matched pencil-beam case with
  64 wavelengths and per-wavelength packet budgets round(10^U(5,7)) (rng 1),
  for batched launches.
- blob_absorber_mc(n=128): a scatterer with a per-voxel absorption map of
  Gaussian blobs and a 16^3 target box, for absorbed-power gradients.
- layered_tissue_heat(n=200): the heat run on the same grid: every medium
  kappa 0.0037 W/(cm K), VHC 3.7 J/(cm^3 K); blood Arrhenius E 422.5 kJ/mol,
  A 7.6e66 /s; 1 W for 5 ms then 10 ms off as 667 + 1333 steps; 37 C start;
  fixed-temperature boundary voxels.
"""
from __future__ import annotations

import math

import numpy as np

from .media import heat_coefficients, interface_normals, interface_normals_fast, sensor_weights
from .model import Geometry, Model, SourceIntensityDistribution, media_table

SIDE = 0.1
MUA = [1e-8, 20.0, 0.5, 230.0]
MUS = [1e-8, 200.0, 200.0, 470.0]
G_ANISO = [1.0, 0.9, 0.9, 0.9]


def layered_media_map(n, side=SIDE):
    d = side / n
    z = (np.arange(n) + 0.5) * d
    y = (np.arange(n) - (n - 1) / 2) * d
    M = np.full((n, n, n), 3, np.uint8)
    M[:, :, z < 0.02] = 2
    M[:, :, z < 0.01] = 1
    Y, Z = np.meshgrid(y, z, indexing="ij")
    M[:, (Y ** 2 + (Z - 0.05) ** 2) < 1e-4] = 4
    return M


def layered_tissue_mc(n=100, matched=True, beam="pencil", nPhotons=1e6, boundaryType=1, side=SIDE):
    d = side / n
    m = Model()
    m.G = Geometry(dx=d, dy=d, dz=d)
    M = layered_media_map(n, side)
    ri = [1.0, 1.0, 1.0, 1.0] if matched else [1.0, 1.4, 1.4, 1.38]
    mc = m.MC
    mc.M = M
    mc.mediaProperties = media_table(MUA, MUS, G_ANISO, ri)
    mc.matchedInterfaces = matched
    mc.interfaceNormals = np.float32(np.nan) if matched else interface_normals(M, ri, (d, d, d))
    mc.nPhotonsRequested = float(nPhotons)
    mc.boundaryType = boundaryType
    mc.spectrum = np.array([1.0])
    ls = mc.LS
    if beam == "pencil":
        ls.sourceType = 0
    elif beam == "tophat":
        ls.sourceType = 4
        ls.FPID = SourceIntensityDistribution(radialDistr=0, radialWidth=0.03)
        ls.AID = SourceIntensityDistribution(radialDistr=1, radialWidth=math.radians(5.0))
    else:
        raise ValueError(beam)
    return m


SPECTRAL_A = [1e-8, 20.0, 0.5, 230.0]     # mua at 500 nm [1/cm]
SPECTRAL_B = [1e-8, 200.0, 200.0, 470.0]  # mus at 500 nm [1/cm]


def spectral_budgets(nL=64, seed=1):
    """Packets per wavelength of the spectral workload: round(10^U(5, 7))
    drawn with numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return [int(round(10 ** u)) for u in rng.uniform(5, 7, nL)]


def layered_tissue_spectral_mc(n=150, nL=64, side=SIDE):
    """The layered skin of layered_tissue_mc (matched, pencil beam) on n^3
    with nL wavelengths linspace(450, 950): mua(l) = a exp(-(l - 500)/150),
    mus(l) = b (l/500)^-1.3, g = 0.9 (air kept at 1e-8).  Returns (model,
    budgets)."""
    m = layered_tissue_mc(n=n, matched=True, beam="pencil", side=side)
    lam = np.linspace(450.0, 950.0, nL)
    a = np.asarray(SPECTRAL_A)[:, None]
    b = np.asarray(SPECTRAL_B)[:, None]
    mua = a * np.exp(-(lam[None, :] - 500.0) / 150.0)
    mus = b * (lam[None, :] / 500.0) ** -1.3
    mua[0, :] = 1e-8
    mus[0, :] = 1e-8
    g = np.tile(np.asarray(G_ANISO)[:, None], (1, nL))
    m.MC.mediaProperties = media_table(mua, mus, g, np.ones((4, nL)))
    m.MC.spectrum = np.ones(nL)
    budgets = spectral_budgets(nL)
    m.MC.nPhotonsRequested = float(sum(budgets))
    return m, budgets


def blob_absorber_mc(n=128, side=SIDE, seed=2, nBlobs=8):
    """A homogeneous scatterer (mus 100/cm, g 0.85, n 1.37 everywhere, so
    refraction happens only at the cuboid surface) with a per-voxel mua map
    1 + sum of 8 Gaussian blobs (peak U[5, 50] /cm, sigma U[4, 12] voxels,
    centres uniform; numpy.random.default_rng(seed)), a pencil beam along z
    and the target box voxels [56,72) x [56,72) x [24,40) scaled to n.
    Returns (model, muaMap float32 [n,n,n], box)."""
    rng = np.random.default_rng(seed)
    d = side / n
    m = Model()
    m.G = Geometry(dx=d, dy=d, dz=d)
    mc = m.MC
    mc.M = np.ones((n, n, n), np.uint8)
    mc.mediaProperties = media_table([1.0], [100.0], [0.85], [1.37])
    mc.matchedInterfaces = True
    mc.interfaceNormals = np.float32(np.nan)
    mc.boundaryType = 1
    mc.spectrum = np.array([1.0])
    mc.LS.sourceType = 0
    mc.nPhotonsRequested = 1e7
    X, Y, Z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    mua = np.ones((n, n, n))
    for _ in range(nBlobs):
        peak, sig = rng.uniform(5, 50), rng.uniform(4, 12) * n / 128
        c = rng.uniform(0, n, 3)
        mua += peak * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * sig ** 2))
    s = n / 128
    box = ((int(56 * s), int(72 * s)), (int(56 * s), int(72 * s)), (int(24 * s), int(40 * s)))
    return m, mua.astype(np.float32), box


def layered_tissue_heat(n=200, side=SIDE, power=1.0, NFR=None, steps_on=667, steps_off=1333,
                        duration_on=5e-3, duration_off=10e-3):
    """(T0, Omega0, P_on, P_off) MEX arguments for the heat run.  NFR is the
    normalized fluence rate of the MC run on the same grid (single
    [n,n,n]); None uses a synthetic exponential-decay fluence."""
    d = side / n
    M = layered_media_map(n, side)
    nM = 4
    VHC = np.full(nM, 3.7, np.float32)
    TC = np.full(nM, 0.0037, np.float32)
    c = heat_coefficients(VHC, TC, d, d, d)
    if NFR is None:
        z = (np.arange(n) + 0.5) * d
        x = (np.arange(n) - (n - 1) / 2) * d
        X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
        NFR = (np.exp(-(X ** 2 + Y ** 2) / (2 * 0.02 ** 2)) * np.exp(-Z / 0.03) / (2 * math.pi * 0.02 ** 2)).astype(np.float32)
    mua = np.asarray(MUA, np.float32)[M - 1]
    NA = NFR * mua                                   # normalized absorption [W/cm^3 per W]
    dTdt_abs = (NA * np.float32(power) / VHC[M - 1]).astype(np.float32)
    A = np.array([0, 0, 0, 7.6e66])
    E = np.array([0, 0, 0, 422.5e3])
    xs = (np.arange(n) - (n - 1) / 2) * d
    zs = (np.arange(n) + 0.5) * d
    li, w = sensor_weights([[0, 0, 0.015], [0, 0, 0.05], [0.01, 0.0, 0.03]], xs, xs, zs)
    base = {"M": (M - 1).astype(np.uint8), "A": A, "E": E, "dTdtperdeltaT": c, "dTdt_abs": dTdt_abs,
            "useAllCPUs": True, "heatBoundaryType": 1, "tempSensorCornerIdxs": li, "tempSensorInterpWeights": w}
    P_on = dict(base, lightsOn=True, steps=steps_on, dt=duration_on / steps_on)
    P_off = dict(base, lightsOn=False, steps=steps_off, dt=duration_off / steps_off)
    T0 = np.full((n, n, n), 37.0, np.float32)
    Om0 = np.zeros((n, n, n), np.float32)
    return T0, Om0, P_on, P_off


def multi_grid_sides(n_models=32, seed=3, lo=64, hi=256):
    """Grid sides log-uniform in [lo, hi] from numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), n_models))).astype(int)


def multi_grid_sweep(n_models=32, seed=3, side=SIDE):
    """Independent heterogeneous models on uneven grids (sides from
    multi_grid_sides(n_models, seed)): a scattering background (mua 1/cm,
    mus 100/cm, g 0.9, n 1.37) with 1-6 ellipsoid inclusions (each its own
    medium: mua U[0.5, 100] /cm, mus U[50, 500] /cm, g U[0.7, 0.95],
    n U[1.33, 1.45]; centres and semi-axes uniform in the cube), cyclic x/y
    boundaries, a pencil beam along z at the centre and 1e6-1e7 packets
    (log-uniform).  Returns a list of dicts: model, nPhotons, box (the
    gradient's target box, 0-based half-open voxel ranges), heat (a function
    NFR -> (T0, Omega0, P) for 1000 steps of 5 us with the light on,
    inclusions carrying the Arrhenius parameters of blood)."""
    rng = np.random.default_rng(seed + 1000)
    out = []
    for n in multi_grid_sides(n_models, seed):
        n = int(n)
        d = side / n
        c = (np.arange(n) + 0.5) * d
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        M = np.ones((n, n, n), np.uint8)
        nInc = int(rng.integers(1, 7))
        mua, mus, g, ri = [1.0], [100.0], [0.9], [1.37]
        for k in range(nInc):
            ctr = rng.uniform(0.15, 0.85, 3) * side
            ax = rng.uniform(0.05, 0.25, 3) * side
            inside = ((X - ctr[0]) / ax[0]) ** 2 + ((Y - ctr[1]) / ax[1]) ** 2 + ((Z - ctr[2]) / ax[2]) ** 2 <= 1
            M[inside] = k + 2
            mua.append(rng.uniform(0.5, 100))
            mus.append(rng.uniform(50, 500))
            g.append(rng.uniform(0.7, 0.95))
            ri.append(rng.uniform(1.33, 1.45))
        m = Model()
        m.G = Geometry(dx=d, dy=d, dz=d)
        mc = m.MC
        mc.M = M
        mc.mediaProperties = media_table(mua, mus, g, ri)
        mc.interfaceNormals = interface_normals_fast(M, ri, (d, d, d))
        mc.boundaryType = 3
        mc.spectrum = np.array([1.0])
        mc.LS.sourceType = 0
        mc.LS.xFocus = mc.LS.yFocus = 0.0
        mc.LS.zFocus = 0.0
        nph = int(round(math.exp(rng.uniform(math.log(1e6), math.log(1e7)))))
        mc.nPhotonsRequested = float(nph)
        lo, hi = n // 2 - max(n // 16, 1), n // 2 + max(n // 16, 1)
        box = ((lo, hi), (lo, hi), (n // 8, n // 4))
        out.append({"model": m, "nPhotons": nph, "box": box, "side": n,
                    "heat": (lambda NFR, M=M, mua=np.asarray(mua), nM=len(mua), d=d: _sweep_heat(NFR, M, mua, nM, d))})
    return out


def _sweep_heat(NFR, M, mua, nM, d, power=1.0, steps=1000, dt=5e-6):
    VHC = np.full(nM, 3.7, np.float32)
    TC = np.full(nM, 0.0037, np.float32)
    NFR = np.asarray(NFR, np.float32).reshape(M.shape, order="F") if np.ndim(NFR) != 3 else np.asarray(NFR, np.float32)
    dTdt_abs = (NFR * mua.astype(np.float32)[M - 1] * np.float32(power) / VHC[M - 1]).astype(np.float32)
    A = np.zeros(nM)
    E = np.zeros(nM)
    A[1:], E[1:] = 7.6e66, 422.5e3
    P = {"M": (M - 1).astype(np.uint8), "A": A, "E": E, "dTdtperdeltaT": heat_coefficients(VHC, TC, d, d, d),
         "dTdt_abs": dTdt_abs, "useAllCPUs": True, "heatBoundaryType": 1, "tempSensorCornerIdxs": np.zeros(0),
         "tempSensorInterpWeights": np.zeros((0, 3)), "lightsOn": True, "steps": steps, "dt": dt}
    T0 = np.full(M.shape, 37.0, np.float32)
    return T0, np.zeros(M.shape, np.float32), P


def thermochromic_params(n=200, seed=4):
    """Per-voxel optical parameter sets (mua, mus, g) of one temperature-
    dependent medium filling an n^3 grid, as getOpticalMediaProperties.m
    collects them for binning (3 x n^3 single, one column per voxel, x
    fastest): a heated spot (Gaussian, 37 -> 77 degC, with 0.2 degC noise)
    drives a thermochromic absorber, mua = 0.5 + 9.5 * s and mus = 200 -
    60 * s with s = (T - 37) / 40, and g = 0.9 throughout (a constant row)."""
    rng = np.random.default_rng(seed)
    ax = (np.arange(n) - (n - 1) / 2) / n
    X, Y, Z = np.meshgrid(ax, ax, (np.arange(n) + 0.5) / n, indexing="ij")
    T = 37 + 40 * np.exp(-(X ** 2 + Y ** 2) / 0.02 - Z ** 2 / 0.08) + 0.2 * rng.standard_normal(X.shape)
    s = ((T - 37) / 40).ravel(order="F")
    # column-major like the MATLAB array the MEX file receives (one set per column)
    return np.asfortranarray(np.vstack([0.5 + 9.5 * s, 200 - 60 * s, np.full(s.size, 0.9)]).astype(np.float32))

