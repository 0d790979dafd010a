"""Heat simulation with a temperature-dependent medium (the update loop of
simulateHeatDistribution.m with heatSimParamsRecalc, :515-541) around the
GPU operators.  Host logic over GPU runs; no reference output exists to
compare with (MATLAB), so the checks are: with temperature-independent
properties the split run equals one plain run bit for bit, and with a
temperature-dependent heat capacity the hottest voxels land in the
highest-capacity sub-medium."""
from __future__ import annotations

import numpy as np
import pytest

from mcmatlab_amd.media import heat_coefficients
from mcmatlab_amd.workloads import layered_media_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def setup(n=24):
    side = 0.1
    d = side / n
    M = layered_media_map(n, side)
    VHC = np.array([3.7, 3.5, 3.9, 3.8])
    TC = np.array([0.0037, 0.004, 0.0052, 0.0045])
    z = (np.arange(n) + 0.5) * d
    x = (np.arange(n) - (n - 1) / 2) * d
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    Q = 4e3 * np.exp(-(X ** 2 + Y ** 2) / 2e-4) * np.exp(-Z / 0.03)
    P = {"A": np.array([0, 0, 0, 7.6e66]), "E": np.array([0, 0, 0, 422.5e3]), "useAllCPUs": True, "heatBoundaryType": 1,
         "tempSensorCornerIdxs": np.zeros(0), "tempSensorInterpWeights": np.zeros((0, 3))}
    T0 = np.full((n, n, n), 37.0, np.float32)
    Om0 = np.zeros((n, n, n), np.float32)
    return M, VHC, TC, (d, d, d), Q, P, T0, Om0


def test_temperature_independent_split_equals_plain_run(api):
    M, VHC, TC, d, Q, P, T0, Om0 = setup()
    updates = [(7, 2e-5, True), (5, 2e-5, True), (9, 3e-5, False)]
    T, Om, Ms = api.t_dependent_heat(T0, Om0, P, M, VHC, TC, d, 3, lambda T: (VHC[2], TC[2]), 4, updates, Q_abs=Q)
    assert np.all(Ms[M == 3] == 3) and np.all(Ms[M == 4] == 4 + 3)
    # the same schedule on the unsplit media
    Pp = dict(P, M=(M - 1).astype(np.uint8), dTdtperdeltaT=heat_coefficients(VHC, TC, *d),
              dTdt_abs=(Q / VHC.astype(np.float32)[M - 1]).astype(np.float32))
    state = None
    for steps, dt, on in updates:
        api.finiteElementHeatPropagator(T0, Om0, dict(Pp, steps=steps, dt=dt, lightsOn=on), state_f64=state)
        state = api.finiteElementHeatPropagator.last_state
    assert np.array_equal(T, state[0]) and np.array_equal(Om, state[1])


def test_heat_capacity_bins_follow_the_temperature(api):
    M, VHC, TC, d, Q, P, T0, Om0 = setup()
    updates = [(20, 3e-5, True)] * 4
    nB = 5
    T, Om, Ms = api.t_dependent_heat(T0, Om0, P, M, VHC, TC, d, 3, lambda T: (3.9 * (1 + 0.02 * (T - 37)), TC[2]), nB, updates,
                                     Q_abs=10 * Q)
    assert T.max() > 40
    vox = M == 3
    lo, hi = (Ms == 3) & vox, (Ms == 3 + nB - 1) & vox
    assert lo.any() and hi.any()
    assert T[hi].mean() > T[lo].mean()  # sub-media sorted by VHC, which rises with T
    assert np.all(Ms[M == 4] == 4 + nB - 1)
