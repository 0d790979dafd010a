"""Closed-form checks of the heat propagator (runner-agnostic: the oracle or
the GPU).  `run(T0, Omega0, P, T_f64=None) -> dict(T_f64, Omega_f64,
sensorTemps, maxMediaTemps)` in FP64; T_f64 optionally gives the initial state
in FP64.

  uniform           uniform T with no source stays uniform (to rounding).
  conservation      insulating boundaries: sum(VHC * T) is conserved, plus
                    dt * sum(VHC * Q) per lit step (c[a,b]*VHC[a] is symmetric).
  eigenmode         homogeneous medium, insulating boundaries: a product of
                    cosine modes cos(pi k (i + 1/2) / n) is an eigenvector of
                    each 1-D operator with eigenvalue -4 c sin^2(pi k / 2n);
                    one forward-Euler step multiplies it by
                    1 + dt (lx + ly + lz), one Douglas-Gunn step by
                    ((1 + dt (lx/2 + ly + lz)) / (1 - dt lx/2) - dt ly/2) / (1 - dt ly/2)
                    then (. - dt lz/2) / (1 - dt lz/2).
  frozen            fixed-temperature boundary voxels never change.
  arrhenius         constant T: Omega = n dt A exp(-E / (R (T + 273.15))).
  arrhenius_midpoint one step of a decaying eigenmode on a 60 degC offset:
                    Omega = dt A exp(-E / (R ((T^n + T^(n+1)) / 2 + 273.15)))
                    per voxel, with T^(n+1) from the eigenmode's closed form
                    (the damage integrand is taken at the step's mean
                    temperature, not at either end).
  steady_series     three media in slabs along x between two fixed-temperature
                    ends: the series-conductance profile
                    T_k = T_0 + (T_N - T_0) sum_{j<k} 1/G_j / sum_j 1/G_j,
                    G_j = 2 TC_a TC_b / (TC_a + TC_b) of the face between
                    voxels j and j+1, is a steady state of the heterogeneous
                    operator (flux continuity), for any VHC of the media;
                    a coefficient read with its media swapped breaks it.
  sensors_linear    a field linear in x, y and z on a homogeneous grid with
                    fixed-temperature boundaries is steady, and trilinear
                    sensors read it exactly at their positions.
"""
from __future__ import annotations

import math

import numpy as np

from mcmatlab_amd.media import heat_coefficients


def _params(n, nM=1, VHC=(3.7,), TC=(0.0037,), d=0.005, steps=10, dt=1e-3, lights=False, bt=0, M=None, Q=None,
            A=(0.0,), E=(0.0,)):
    nx, ny, nz = n
    c = heat_coefficients(VHC, TC, d, d, d)
    return {"M": np.zeros(n, np.uint8) if M is None else M, "A": np.array(A, float), "E": np.array(E, float),
            "dTdtperdeltaT": c, "dTdt_abs": np.zeros(n, np.float32) if Q is None else Q, "useAllCPUs": True,
            "heatBoundaryType": bt, "tempSensorCornerIdxs": np.zeros(0), "tempSensorInterpWeights": np.zeros((0, 3)),
            "lightsOn": lights, "steps": steps, "dt": dt}


def uniform(run, scheme):
    n = (9, 8, 7)
    P = _params(n, nM=3, VHC=(3.7, 1.0, 2.0), TC=(0.0037, 0.001, 0.01), M=(np.arange(9 * 8 * 7).reshape(n) % 3).astype(np.uint8),
                dt=2e-4 if scheme == "explicit" else 1e-3)
    T0 = np.full(n, 42.5, np.float32)
    r = run(T0, np.zeros(n, np.float32), P)
    np.testing.assert_allclose(r["T_f64"], 42.5, rtol=1e-13)


def conservation(run, scheme, lights):
    rng = np.random.default_rng(1)
    n = (10, 9, 8)
    VHC = np.array([3.7, 1.2, 4.1])
    TC = np.array([0.0037, 0.0011, 0.0052])
    M = rng.integers(0, 3, n).astype(np.uint8)
    Q = (rng.random(n) * 50).astype(np.float32)
    steps = 9
    dt = 1.5e-4 if scheme == "explicit" else 2e-3
    P = _params(n, VHC=VHC, TC=TC, M=M, Q=Q, lights=lights, steps=steps, dt=dt)
    T0 = (30 + 20 * rng.random(n)).astype(np.float32)
    r = run(T0, np.zeros(n, np.float32), P)
    vhc = VHC.astype(np.float32).astype(np.float64)[M]
    e0 = np.sum(vhc * T0)
    e1 = np.sum(vhc * r["T_f64"])
    inj = steps * dt * np.sum(vhc * Q.astype(np.float64)) if lights else 0.0
    # c[a,b] VHC[a] is symmetric up to the FP32 rounding of c
    assert abs(e1 - e0 - inj) <= 1e-6 * e0, (e1 - e0, inj)


def eigenmode(run, scheme):
    n = (12, 10, 8)
    d = 0.005
    k = (1, 2, 1)
    VHC, TC = 3.7, 0.0037
    dt = 1e-3 if scheme == "explicit" else 5e-2
    P = _params(n, d=d, steps=1, dt=dt)
    c = float(np.asarray(P["dTdtperdeltaT"], dtype=np.float64)[0, 0, 0])
    modes = [np.cos(math.pi * k[a] * (np.arange(n[a]) + 0.5) / n[a]) for a in range(3)]
    v = modes[0][:, None, None] * modes[1][None, :, None] * modes[2][None, None, :]
    lam = [-4 * c * math.sin(math.pi * k[a] / (2 * n[a])) ** 2 for a in range(3)]
    T0 = v.astype(np.float32)
    v = np.asfortranarray(v)
    if scheme == "explicit":
        g = 1 + dt * sum(lam)
    else:
        lx, ly, lz = lam
        g = (1 + dt * (lx / 2 + ly + lz)) / (1 - dt * lx / 2)
        g = (g - dt * ly / 2) / (1 - dt * ly / 2)
        g = (g - dt * lz / 2) / (1 - dt * lz / 2)
    r = run(T0, np.zeros(n, np.float32), P, T_f64=v)
    np.testing.assert_allclose(r["T_f64"], g * v, rtol=0, atol=1e-12)


def frozen(run):
    rng = np.random.default_rng(2)
    n = (8, 9, 10)
    Q = (rng.random(n) * 100).astype(np.float32)
    P = _params(n, Q=Q, lights=True, bt=1, steps=5)
    T0 = np.full(n, 37.0, np.float32)
    T0[0] += rng.random(n[1:]).astype(np.float32)  # uneven fixed boundary values
    r = run(T0, np.zeros(n, np.float32), P)
    T = r["T_f64"]
    B = np.zeros(n, bool)
    B[0], B[-1], B[:, 0], B[:, -1], B[:, :, 0], B[:, :, -1] = True, True, True, True, True, True
    assert np.array_equal(T[B], T0.astype(np.float64)[B])
    assert np.all(T[~B] > 37.0)


def arrhenius(run):
    n = (6, 6, 6)
    A, E = 7.6e66, 422.5e3
    steps, dt = 25, 1e-3
    P = _params(n, A=(A,), E=(E,), steps=steps, dt=dt)
    T0 = np.full(n, 70.0, np.float32)
    r = run(T0, np.zeros(n, np.float32), P)
    want = steps * dt * A * math.exp(-E / (8.3144598 * (70.0 + 273.15)))
    np.testing.assert_allclose(r["Omega_f64"], want, rtol=1e-12)
    np.testing.assert_allclose(r["maxMediaTemps"], 70.0, rtol=1e-6)


def arrhenius_midpoint(run, scheme):
    n = (12, 10, 8)
    d = 0.005
    kk = (1, 2, 1)
    A, E = 7.6e66, 422.5e3
    dt = 1e-3 if scheme == "explicit" else 5e-2
    P = _params(n, d=d, steps=1, dt=dt, A=(A,), E=(E,))
    c = float(np.asarray(P["dTdtperdeltaT"], dtype=np.float64)[0, 0, 0])
    modes = [np.cos(math.pi * kk[a] * (np.arange(n[a]) + 0.5) / n[a]) for a in range(3)]
    v = 4.0 * modes[0][:, None, None] * modes[1][None, :, None] * modes[2][None, None, :]
    lam = [-4 * c * math.sin(math.pi * kk[a] / (2 * n[a])) ** 2 for a in range(3)]
    if scheme == "explicit":
        g = 1 + dt * sum(lam)
    else:
        lx, ly, lz = lam
        g = (1 + dt * (lx / 2 + ly + lz)) / (1 - dt * lx / 2)
        g = (g - dt * ly / 2) / (1 - dt * ly / 2)
        g = (g - dt * lz / 2) / (1 - dt * lz / 2)
    T0 = np.asfortranarray(60.0 + v)
    r = run(T0.astype(np.float32), np.zeros(n, np.float32), P, T_f64=T0)
    T1 = 60.0 + g * v
    np.testing.assert_allclose(r["T_f64"], T1, rtol=0, atol=1e-11)
    want = dt * A * np.exp(-E / (8.3144598 * ((T0 + T1) / 2 + 273.15)))
    np.testing.assert_allclose(r["Omega_f64"], want, rtol=1e-9)
    # the damage at either end of the step differs by far more than that
    assert np.max(np.abs(dt * A * np.exp(-E / (8.3144598 * (T1 + 273.15))) / want - 1)) > 1e-4


def steady_series(run, scheme):
    nx = 14
    n = (nx, 3, 3)
    d = 0.005
    VHC = np.array([3.7, 1.2, 4.1])
    TC = np.array([0.0037, 0.0011, 0.0052])
    media = np.array([0] * 4 + [1] * 5 + [2] * 5, np.uint8)
    M = np.repeat(np.repeat(media[:, None, None], 3, axis=1), 3, axis=2).copy(order="F")
    dt = 1e-4 if scheme == "explicit" else 1e-3
    P = _params(n, VHC=VHC, TC=TC, M=M, bt=1, steps=20, dt=dt)
    a, b = media[:-1], media[1:]
    G = 2 * TC[a] * TC[b] / (TC[a] + TC[b])
    R = np.concatenate([[0.0], np.cumsum(1 / G)])
    prof = 20.0 + 30.0 * R / R[-1]
    T0 = np.asfortranarray(np.repeat(np.repeat(prof[:, None, None], 3, axis=1), 3, axis=2))
    r = run(T0.astype(np.float32), np.zeros(n, np.float32), P, T_f64=T0)
    # the coefficient table is single precision: c[a,b] VHC[a] is symmetric to ~1e-7
    np.testing.assert_allclose(r["T_f64"], T0, rtol=0, atol=2e-6 * 30)


def sensors_linear(run, scheme):
    from mcmatlab_amd.media import sensor_weights

    n = (9, 8, 7)
    d = 0.005
    dt = 2e-4 if scheme == "explicit" else 1e-3
    P = _params(n, d=d, bt=1, steps=4, dt=dt)
    x = (np.arange(n[0]) - (n[0] - 1) / 2) * d
    y = (np.arange(n[1]) - (n[1] - 1) / 2) * d
    z = (np.arange(n[2]) + 0.5) * d
    pos = np.array([[0.0031, -0.0047, 0.0122], [-0.011, 0.006, 0.021], [0.0, 0.0, 0.0125]])
    li, w = sensor_weights(pos, x, y, z)
    P["tempSensorCornerIdxs"], P["tempSensorInterpWeights"] = li, w
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    f = lambda px, py, pz: 40.0 + 300.0 * px - 200.0 * py + 500.0 * pz
    T0 = np.asfortranarray(f(X, Y, Z))
    r = run(T0.astype(np.float32), np.zeros(n, np.float32), P, T_f64=T0)
    np.testing.assert_allclose(r["T_f64"], T0, rtol=0, atol=1e-11)
    want = f(pos[:, 0], pos[:, 1], pos[:, 2])
    S = np.asarray(r["sensorTemps"], np.float64).reshape(len(pos), -1, order="F")
    assert S.shape[1] == 5
    np.testing.assert_allclose(S, np.repeat(want[:, None], 5, axis=1), rtol=0, atol=1e-9)
