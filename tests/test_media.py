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


def test_sobel_gradient_of_a_ramp_and_of_a_constant():
    """The reference's Sobel kernel (difference of the two neighbours weighted
    by [[1,3,1],[3,6,3],[1,3,1]] over the orthogonal neighbourhood, weights
    summing to 22): a unit ramp along one axis has gradient 44 there and 0
    along the others in the interior; replicated edges give half of it on the
    end faces; a constant has none."""
    import numpy as np

    from mcmatlab_amd.media import _sobel_gradient

    x = np.arange(7.0)
    for ax in range(3):
        sh = [1, 1, 1]
        sh[ax] = 7
        ramp = np.broadcast_to(x.reshape(sh), (7, 7, 7)).copy()
        g = _sobel_gradient(ramp)
        inner = (slice(1, -1),) * 3
        assert np.allclose(g[ax][inner], 44.0)
        end = [slice(None)] * 3
        end[ax] = 0
        assert np.allclose(g[ax][tuple(end)], 22.0)
        for other in range(3):
            if other != ax:
                assert np.allclose(g[other], 0.0)
    assert all(np.all(c == 0) for c in _sobel_gradient(np.full((5, 6, 7), 3.0)))


def test_smoothn_scales_each_cosine_mode_by_its_filter_factor():
    """smoothn with a given smoothness is the DCT-II filter
    1 / (1 + s Lambda^2): a cosine mode cos(pi k (i + 1/2) / n) along one axis
    is multiplied by 1 / (1 + s (2 - 2 cos(pi k / n))^2 / dI^4), a constant is
    kept (Lambda = 0)."""
    import numpy as np

    from mcmatlab_amd.media import _smoothn

    n = (12, 10, 8)
    s = 7.0
    spacing = (1.0, 2.0, 1.0)  # dI = spacing / max(spacing) = (0.5, 1, 0.5)
    dI = np.array(spacing) / max(spacing)
    for ax, k in ((0, 3), (1, 2), (2, 5)):
        sh = [1, 1, 1]
        sh[ax] = n[ax]
        mode = np.broadcast_to(np.cos(np.pi * k * (np.arange(n[ax]) + 0.5) / n[ax]).reshape(sh), n).copy()
        (out,) = _smoothn([mode], s, spacing)
        lam = (2 - 2 * np.cos(np.pi * k / n[ax])) / dI[ax] ** 2
        np.testing.assert_allclose(out, mode / (1 + s * lam ** 2), atol=1e-12)
    (c,) = _smoothn([np.full(n, 2.5)], s, spacing)
    np.testing.assert_allclose(c, 2.5, rtol=1e-12)


def test_interface_normals_of_a_sphere_point_into_its_own_class():
    """A sphere of n = 1.4 in n = 1: on the voxels either side of its surface
    the normal points into the voxel's own class, within 20 degrees of the
    radial direction (inward inside the sphere, outward outside)."""
    import numpy as np

    n, d = 32, 0.01
    c = (np.arange(n) + 0.5 - n / 2) * d
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    M = np.where(R < 0.1, 2, 1).astype(np.uint8)
    IN = interface_normals(M, [1.0, 1.4], (d, d, d), smoothingLengthScale=0.03)
    th = IN[0].reshape(M.shape, order="F").astype(np.float64)
    ph = IN[1].reshape(M.shape, order="F").astype(np.float64)
    nv = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    rhat = np.stack([X, Y, Z]) / np.maximum(R, 1e-12)
    shell_in = (R < 0.1) & (R > 0.1 - 1.5 * d)
    shell_out = (R >= 0.1) & (R < 0.1 + 1.5 * d)
    cos_in = np.sum(nv * -rhat, axis=0)[shell_in]
    cos_out = np.sum(nv * rhat, axis=0)[shell_out]
    lim = np.cos(np.radians(20))
    assert np.all(cos_in > lim) and np.all(cos_out > lim), (cos_in.min(), cos_out.min())


def test_phase_function_cdf_and_sharing():
    """The custom phase function's CDF (getOpticalMediaProperties.m:94-104)
    reproduces the Henyey-Greenstein polar CDF to the midpoint rule's
    accuracy, and identical CDFs share one column."""
    import math

    import numpy as np

    from mcmatlab_amd.media import phase_function_cdf, unique_cdfs

    g = 0.5
    hg = lambda lam, t: (1 - g * g) / (1 + g * g - 2 * g * math.cos(t)) ** 1.5
    cdf = phase_function_cdf(hg)
    assert cdf.shape == (201,) and cdf[0] == 0 and abs(cdf[-1] - 1) < 1e-15 and np.all(np.diff(cdf) > 0)
    edges = math.pi / 200 * np.arange(201)
    # P(theta <= t) = 1 - P(cos theta <= cos t), the HG cosine CDF
    exact = 1 - (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * np.cos(edges)) - 1 / (1 + g))
    np.testing.assert_allclose(cdf, exact, atol=2e-4)
    iso = phase_function_cdf(lambda lam, t: 1.0)
    table, idx = unique_cdfs([None, cdf, iso, cdf.copy()])
    assert table.shape == (201, 2) and np.isnan(idx[0]) and list(idx[1:]) == [0.0, 1.0, 0.0]
