"""Host-side input builders (mcmatlab_amd/media.py) against hand-computed
values of the reference's MATLAB formulas."""
import math

import numpy as np

from mcmatlab_amd.media import heat_coefficients, heat_dtmax, interface_normals, sensor_weights


def test_interface_normals_planar_interface_points_into_each_medium():
    M = np.ones((9, 9, 12), np.uint8)
    M[:, :, 6:] = 2
    IN = interface_normals(M, [1.0, 1.5], (0.01, 0.01, 0.01))
    theta = IN[0].reshape(M.shape, order="F")
    # the n=1.5 class: gradient of its indicator points +z (theta = 0);
    # the n=1 class: -z (theta = pi)
    assert np.allclose(theta[4, 4, 6], 0.0, atol=1e-6)
    assert np.allclose(theta[4, 4, 5], math.pi, atol=1e-6)


def test_heat_coefficients_harmonic_mean_over_vhc():
    VHC, TC = [3.7, 1.2], [0.0037, 0.0011]
    c = heat_coefficients(VHC, TC, 0.01, 0.02, 0.04)
    eff = 2 * TC[0] * TC[1] / (TC[0] + TC[1])
    np.testing.assert_allclose(c[0, 1, 0], eff / VHC[0] / 0.01 ** 2, rtol=1e-6)
    np.testing.assert_allclose(c[1, 0, 2], eff / VHC[1] / 0.04 ** 2, rtol=1e-6)
    np.testing.assert_allclose(c[0, 0, 1], TC[0] / VHC[0] / 0.02 ** 2, rtol=1e-6)
    assert c.dtype == np.float32 and c.shape == (2, 2, 3)


def test_heat_dtmax_homogeneous_interior():
    M = np.ones((6, 6, 6), np.uint8)
    d = 5e-4
    dtmax = heat_dtmax(M, [3.7], [0.0037], d, d, d)
    # an interior voxel has 6 neighbours: VHC / (6 TC / d^2)
    np.testing.assert_allclose(dtmax, 3.7 / (6 * 0.0037 / d ** 2), rtol=1e-6)


def test_sensor_weights_trilinear_corner():
    x = np.arange(5) * 0.1
    li, w = sensor_weights([[0.15, 0.2, 0.05], [-1, 9, 0.0]], x, x, x)
    assert li[0] == 1 + 2 * 5 + 0 * 25
    np.testing.assert_allclose(w[0], [0.5, 0.0, 0.5], atol=1e-12)
    assert li[1] == 0 + 3 * 5 + 0
    np.testing.assert_allclose(w[1], [0.0, 1.0, 0.0])


def test_spectral_budget_fixture_matches_generator():
    import json
    import os

    from mcmatlab_amd.workloads import layered_tissue_spectral_mc, spectral_budgets

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "spectral64_budgets.json")))
    assert fx["budgets"] == spectral_budgets(64)
    assert 1e5 <= min(fx["budgets"]) and max(fx["budgets"]) <= 1e7
    m, b = layered_tissue_spectral_mc(n=10, nL=64)
    mua = m.MC.mediaProperties["mua"]
    assert mua.shape == (4, 64)
    lam = np.linspace(450, 950, 64)
    assert np.allclose(mua[1], 20 * np.exp(-(lam - 500) / 150))
    assert np.allclose(m.MC.mediaProperties["mus"][3], 470 * (lam / 500) ** -1.3)


def test_conductivity_gradient_is_the_chain_rule_of_the_coefficients():
    """conductivity_gradient(G) is the gradient of sum(G * c(TC)) with
    c[a,b,d] = 2 TC_a TC_b / (TC_a + TC_b) / VHC_a / d^2 (central differences
    in double precision), including the a == b entries (TC_eff = TC_a)."""
    import numpy as np

    from mcmatlab_amd.media import conductivity_gradient

    rng = np.random.default_rng(3)
    nM = 4
    VHC = rng.uniform(1, 4, nM)
    TC = rng.uniform(1e-3, 6e-3, nM)
    dx, dy, dz = 1e-3, 2e-3, 1.5e-3
    G = rng.standard_normal((nM, nM, 3))

    def F(tc):
        te = 2 * np.outer(tc, tc) / (tc[:, None] + tc[None, :])
        c = np.stack([te / VHC[:, None] / dd ** 2 for dd in (dx, dy, dz)], axis=2)
        return float(np.sum(G * c))

    g = conductivity_gradient(G, VHC, TC, dx, dy, dz)
    for m in range(nM):
        h = 1e-6 * TC[m]
        e = np.zeros(nM)
        e[m] = h
        fd = (F(TC + e) - F(TC - e)) / (2 * h)
        assert abs(fd - g[m]) <= 1e-7 * abs(fd) + 1e-9 * np.abs(g).max(), (m, fd, g[m])
