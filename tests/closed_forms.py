"""Closed-form checks of Monte Carlo transport, independent of any reference
run.  Each check builds a model whose answer the mathematics fixes, runs it
through `run(model, seed) -> result dict` (the oracle, the host build of the
GPU engine, or the GPU), and asserts.  A dropped term, a wrong sign or index,
or a swapped axis in launch/step/absorb/scatter/Fresnel/tallies fails one of
them.

  beer_lambert      non-scattering absorber, pencil beam: every packet
                    deposits w_k (1 - e^{-mua dz}) in voxel k, so
                    NFR_k = e^{-mua z_k} (1 - e^{-mua dz}) / (dx dy dz mua) exactly,
                    and the transmitted fraction is e^{-mua L}.
  fresnel_normal    normal incidence on an n = 1.5 half-space from n = 1: the
                    specular reflectance is ((n1 - n2) / (n1 + n2))^2 = 0.04.
  tir               oblique incidence from n = 1.5 beyond asin(1/1.5): nothing
                    leaves through the top face.
  hg_single         a slab thin enough that packets scatter at most once: the
                    far-field polar histogram of escaped light follows the
                    Henyey-Greenstein distribution, mean cosine g.
  energy            absorbed + everything leaving through the faces = 1
                    (roulette is unbiased), to statistical precision.
  point_source      isotropic point source in a non-scattering absorber:
                    fluence e^{-mua r} / (4 pi r^2) per unit power.
"""
from __future__ import annotations

import math

import numpy as np

from mcmatlab_amd.model import Geometry, Model, media_table


def _slab(n=(21, 21, 40), d=0.005, mua=10.0, mus=1e-8, g=1.0, nPh=4000, bt=1):
    m = Model()
    m.G = Geometry(dx=d, dy=d, dz=d)
    m.MC.M = np.ones(n, np.uint8)
    m.MC.mediaProperties = media_table([mua], [mus], [g], [1.0])
    m.MC.nPhotonsRequested = float(nPh)
    m.MC.boundaryType = bt
    m.MC.spectrum = np.array([1.0])
    return m


def beer_lambert(run, rtol=2e-6):
    mua, d, nz = 10.0, 0.005, 40
    m = _slab(mua=mua, d=d, n=(21, 21, nz))
    r = run(m, 1)
    assert r["nPhotons"] == 4000
    nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(21, 21, nz, order="F")
    z = np.arange(nz) * d
    want = np.exp(-mua * z) * (1 - np.exp(-mua * d)) / (d ** 3 * mua)
    np.testing.assert_allclose(nfr[10, 10, :], want, rtol=rtol)
    off = nfr.copy()
    off[10, 10, :] = 0
    assert np.all(off == 0), "deposit outside the beam column"
    trans = float(np.sum(r["NI_zpos"]) * d * d)
    np.testing.assert_allclose(trans, math.exp(-mua * nz * d), rtol=rtol)
    assert float(np.sum(r["NI_zneg"])) == 0.0


def _two_layer(n2=1.5, nPh=20000, theta=0.0, top=8, nz=24):
    from mcmatlab_amd.media import interface_normals

    d = 0.005
    m = _slab(n=(21, 21, nz), d=d, mua=1e-8, mus=1e-8, nPh=nPh)
    M = np.ones((21, 21, nz), np.uint8)
    M[:, :, top:] = 2
    m.MC.M = M
    ri = [1.0, n2]
    m.MC.mediaProperties = media_table([1e-8, 1e-8], [1e-8, 1e-8], [1, 1], ri)
    m.MC.interfaceNormals = interface_normals(M, ri, (d, d, d))
    m.MC.LS.theta = theta
    return m


def fresnel_normal(run, seeds=(1, 2)):
    R = ((1 - 1.5) / (1 + 1.5)) ** 2
    fr = []
    for s in seeds:
        m = _two_layer()
        r = run(m, s)
        fr.append(float(np.sum(r["NI_zneg"])) * m.G.dx * m.G.dy)
    N = 20000 * len(seeds)
    got = np.mean(fr)
    se = math.sqrt(R * (1 - R) / N)
    assert abs(got - R) < 4 * se, (got, R, se)


def tir(run):
    # start inside the n = 1.5 layer heading up at 60 degrees from the normal
    # (critical angle 41.8 degrees): the interface reflects every packet
    from mcmatlab_amd.media import interface_normals

    d = 0.005
    nz = 24
    m = _slab(n=(41, 41, nz), d=d, mua=1e-8, mus=1e-8, nPh=2000)
    M = np.full((41, 41, nz), 2, np.uint8)
    M[:, :, :4] = 1
    m.MC.M = M
    ri = [1.0, 1.5]
    m.MC.mediaProperties = media_table([1e-8, 1e-8], [1e-8, 1e-8], [1, 1], ri)
    m.MC.interfaceNormals = interface_normals(M, ri, (d, d, d))
    m.MC.LS.sourceType = 1  # isotropic point source deep in the layer...
    m.MC.LS.zFocus = 0.08
    r = run(m, 3)
    # ...so some packets reach the top interface at every angle; those beyond
    # the critical angle cannot escape through the top face
    esc = float(np.sum(r["NI_zneg"])) * d * d
    # fraction of isotropic emission within the escape cone from 0.06 cm below
    # the interface is below (1 - cos 41.8deg)/2 = 0.127 (and Fresnel losses)
    assert 0.0 < esc < 0.127


def hg_single(run, g=0.8, seeds=(1, 2, 3), nbins=20):
    """Packets scatter at most once in a thin slab (P ~ 2%); the escaped
    polar-angle histogram beyond the first bin follows HG."""
    mus = 2.0
    d = 0.01
    res = []
    for s in seeds:
        m = _slab(n=(201, 201, 1), d=d, mua=1e-8, mus=mus, g=g, nPh=100000)
        m.MC.farFieldRes = nbins
        r = run(m, s)
        res.append(np.asarray(r["farField"], dtype=np.float64).reshape(nbins, nbins, order="F"))
    ff = np.mean(res, axis=0).sum(axis=1)  # sum over phi -> theta histogram
    edges = np.linspace(0, math.pi, nbins + 1)
    mu = np.cos(edges)

    def hg_cdf(c):  # P(cos theta <= c) for HG
        return (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * c) - 1 / (1 + g))

    p_bin = hg_cdf(mu[:-1]) - hg_cdf(mu[1:])  # theta in [edge_k, edge_k+1]
    # packets scattered once leave through the slab faces; with mus*d = 0.02
    # the scattered fraction is 1 - e^{-0.02} (times ~1 for escape)
    scat = 1 - math.exp(-mus * d)
    want = scat * p_bin
    got = ff.copy()
    got[0] -= math.exp(-mus * d)  # unscattered beam lands in the first theta bin
    N = 100000 * len(seeds)
    se = np.sqrt(np.maximum(want, 1e-12) / N)
    # the path to the face is short (at most the slab) so absorption is nil;
    # compare bins away from grazing angles where the slab edge interferes
    sel = slice(1, nbins // 2 - 2)
    z = (got[sel] - want[sel]) / se[sel]
    assert np.all(np.abs(z) < 5), z
    # mean cosine of the scattered light over the forward hemisphere bins
    c_mid = np.cos(0.5 * (edges[:-1] + edges[1:]))
    assert abs(np.sum(got[1:] * c_mid[1:]) / np.sum(got[1:]) - np.sum(want[1:] * c_mid[1:]) / np.sum(want[1:])) < 0.03


def energy(run, seeds=(1, 2, 3, 4)):
    """absorbed + all six faces = 1 for the layered tissue with Fresnel."""
    import models

    vals = []
    for s in seeds:
        m = models.cases()["fresnel_bt1"]()
        m.MC.nPhotonsRequested = 20000
        r = run(m, s)
        mua = np.asarray(m.MC.mediaProperties["mua"])[:, 0]
        V = m.G.dx * m.G.dy * m.G.dz
        tot = float(np.sum(np.asarray(r["NFR"], dtype=np.float64).reshape(m.MC.M.shape, order="F") * mua[m.MC.M - 1]) * V)
        for k, a in (("NI_xpos", m.G.dy * m.G.dz), ("NI_xneg", m.G.dy * m.G.dz), ("NI_ypos", m.G.dx * m.G.dz),
                     ("NI_yneg", m.G.dx * m.G.dz), ("NI_zpos", m.G.dx * m.G.dy), ("NI_zneg", m.G.dx * m.G.dy)):
            tot += float(np.sum(np.asarray(r[k], dtype=np.float64))) * a
        vals.append(tot)
    v = np.array(vals)
    se = v.std(ddof=1) / math.sqrt(len(v))
    assert abs(v.mean() - 1) < max(5 * se, 2e-4), (v, se)


def point_source(run):
    mua, d, n = 2.0, 0.004, 41
    m = _slab(n=(n, n, n), d=d, mua=mua, nPh=200000)
    m.MC.LS.sourceType = 1
    m.MC.LS.zFocus = (n // 2 + 0.5) * d
    r = run(m, 5)
    nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(n, n, n, order="F")
    c = n // 2
    for k in (4, 6, 9):
        # average over the six on-axis voxels k voxels from the source voxel
        vals = [nfr[c + k, c, c], nfr[c - k, c, c], nfr[c, c + k, c], nfr[c, c - k, c], nfr[c, c, c + k], nfr[c, c, c - k]]
        rr = k * d
        want = math.exp(-mua * rr) / (4 * math.pi * rr * rr)
        assert abs(np.mean(vals) / want - 1) < 0.08, (k, np.mean(vals), want)


# ---------------------------------------------------------------------------
# Launch, boundary, collector and criteria checks.  "Vacuum" means
# mua = mus = 1e-8 /cm and n = 1: a packet flies straight from its launch
# point to the face it leaves through, so exit images are projections of the
# launch distributions.

C_LIGHT = 29979245800.0  # cm/s (MCmatlab.c:159)
KILLRANGE = 5             # MCmatlab.c:165


def _vacuum(n=(41, 41, 20), d=0.005, nPh=20000, bt=1):
    return _slab(n=n, d=d, mua=1e-8, mus=1e-8, g=1.0, nPh=nPh, bt=bt)


def _face_image(r, key, m):
    return np.asarray(r[key], dtype=np.float64).reshape(m.MC.M.shape[0], m.MC.M.shape[1], -1, order="F")[..., 0]


def _zpos_moments(r, m):
    """Power-weighted <x^2>, <y^2>, <r^2>, <r^4> of the packets leaving
    through the bottom face, at pixel centres, and the total power there."""
    nx, ny, _ = m.MC.M.shape
    d = m.G.dx
    img = _face_image(r, "NI_zpos", m) * d * d
    x = (np.arange(nx) + 0.5 - nx / 2) * d
    y = (np.arange(ny) + 0.5 - ny / 2) * m.G.dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    tot = img.sum()
    R2 = X ** 2 + Y ** 2
    return {"tot": tot, "x2": np.sum(img * X ** 2) / tot, "y2": np.sum(img * Y ** 2) / tot, "r2": np.sum(img * R2) / tot,
            "r4": np.sum(img * R2 ** 2) / tot, "xm": np.sum(img * X) / tot, "ym": np.sum(img * Y) / tot}


def planewave(run):
    """Plane wave (sourceType 2) straight down into a non-scattering absorber.
    Boundary type 1: launches cover the top face; every packet deposits
    e^{-mua z_k}(1 - e^{-mua dz}) in layer k, so each layer's summed NFR is
    fixed exactly, and the packets spread evenly over the columns.
    Boundary type 2: launches cover KILLRANGE^2 = 25 times the face and the
    normalisation divides by 25 (MCmatlablib.c:1195-1200), so the layer sums
    carry the fraction of packets that hit the cuboid times 25: the same
    ratio to the closed form in every layer, 1 within its binomial error."""
    mua, d, nz, nxy = 10.0, 0.005, 24, 21
    want = np.exp(-mua * np.arange(nz) * d) * (1 - np.exp(-mua * d)) / (d * mua)
    m = _slab(n=(nxy, nxy, nz), d=d, mua=mua, nPh=8820)
    m.MC.LS.sourceType = 2
    r = run(m, 1)
    nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(nxy, nxy, nz, order="F")
    layers = nfr.sum(axis=(0, 1)) * d * d
    np.testing.assert_allclose(layers, want, rtol=5e-5)
    cnt = nfr[:, :, 0] / nfr[:, :, 0].sum() * 8820  # packets per column
    chi2 = np.sum((cnt - 20.0) ** 2 / 20.0)
    assert chi2 / (nxy * nxy - 1) < 1.5, chi2
    N = 40000
    m = _slab(n=(nxy, nxy, nz), d=d, mua=mua, nPh=N, bt=2)
    m.MC.LS.sourceType = 2
    r = run(m, 2)
    nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(nxy, nxy, nz, order="F")
    ratio = nfr.sum(axis=(0, 1)) * d * d / want
    np.testing.assert_allclose(ratio, ratio[0], rtol=5e-5)
    se = math.sqrt(25 * (1 - 1 / 25) / N)
    assert abs(ratio[0] - 1) < 4 * se, (ratio[0], se)


def radial_focal_profiles(run, N=20000):
    """Radial beam (sourceType 4) with parallel rays (zero angular width):
    the bottom-face image is the focal-plane distribution.  Top-hat of radius
    R: <r^2> = R^2/2; Gaussian of 1/e^2 radius w: <r^2> = w^2/2; pixel
    centres add d^2/6."""
    from mcmatlab_amd.model import SourceIntensityDistribution as SID

    for distr, width, want, var in ((0, 0.06, 0.06 ** 2 / 2, 0.06 ** 4 / 12), (1, 0.03, 0.03 ** 2 / 2, 0.03 ** 4 / 4)):
        m = _vacuum(nPh=N)
        m.MC.LS.sourceType = 4
        m.MC.LS.FPID = SID(radialDistr=distr, radialWidth=width)
        m.MC.LS.AID = SID(radialDistr=1, radialWidth=0.0)
        r = run(m, 3)
        mo = _zpos_moments(r, m)
        assert abs(mo["tot"] - 1) < 1e-6
        got = mo["r2"] - m.G.dx ** 2 / 6
        assert abs(got - want) < 5 * math.sqrt(var / N), (distr, got, want)
        assert abs(mo["xm"]) < 5 * math.sqrt(want / 2 / N) and abs(mo["ym"]) < 5 * math.sqrt(want / 2 / N)


def _farfield_theta(r, res):
    ff = np.asarray(r["farField"], dtype=np.float64).reshape(res, res, order="F")
    return ff.sum(axis=1)  # theta bins over [0, pi]


def radial_angular_profiles(run, N=20000, res=20):
    """Radial beam from a point focus: the escaped light's polar histogram
    (far field) is the angular distribution.  Top-hat of half-angle t0:
    P(theta < t) = tan^2 t / tan^2 t0; Gaussian: 1 - exp(-2 tan^2 t / tan^2 t0);
    Lambertian: sin^2 t."""
    from mcmatlab_amd.model import SourceIntensityDistribution as SID

    t0 = 0.5
    cdfs = {0: lambda t: np.minimum(np.tan(t) ** 2 / math.tan(t0) ** 2, 1.0),
            1: lambda t: 1 - np.exp(-2 * np.tan(t) ** 2 / math.tan(t0) ** 2),
            2: lambda t: np.sin(t) ** 2}
    edges = np.linspace(0, math.pi / 2, res // 2 + 1)
    for distr, cdf in cdfs.items():
        m = _vacuum(n=(81, 81, 20), nPh=N)
        m.MC.farFieldRes = res
        m.MC.LS.sourceType = 4
        m.MC.LS.FPID = SID(radialDistr=0, radialWidth=0.0)
        m.MC.LS.AID = SID(radialDistr=distr, radialWidth=t0 if distr != 2 else 0.0)
        r = run(m, 4)
        got = _farfield_theta(r, res)
        assert abs(got.sum() - 1) < 1e-6 and np.all(got[res // 2:] == 0)
        p = np.diff(cdf(edges))
        z = (got[:res // 2] - p) / np.sqrt(np.maximum(p * (1 - p), 1e-12) / N)
        assert np.all(np.abs(z) < 5), (distr, z)


def xy_profiles(run, N=20000):
    """X/Y beam (sourceType 5).  Parallel rays: the bottom image has
    <x^2> = w_x^2/3 for a top-hat X of half-width w_x and <y^2> = w_y^2/4 for
    a Gaussian Y of 1/e^2 half-width w_y; psi = pi/2 turns X onto y.  From a
    point focus at the top centre: slopes u_x/u_z = tan(phi_X) (top-hat in
    tan, half-width tan t_x) and |u_y/u_z| = |tan(phi_Y)| (Gaussian), so the
    exit point Lz below has <x^2> = Lz^2 tan^2 t_x / 3, <y^2> = Lz^2 tan^2 t_y / 4."""
    from mcmatlab_amd.model import SourceIntensityDistribution as SID

    d = 0.005
    for psi, (kx, ky) in ((0.0, ("x2", "y2")), (math.pi / 2, ("y2", "x2"))):
        m = _vacuum(nPh=N)
        m.MC.LS.sourceType = 5
        m.MC.LS.psi = psi
        m.MC.LS.FPID = SID(XDistr=0, XWidth=0.04, YDistr=1, YWidth=0.02)
        m.MC.LS.AID = SID(XDistr=0, XWidth=0.0, YDistr=0, YWidth=0.0)
        mo = _zpos_moments(run(m, 5), m)
        wx, wy = 0.04 ** 2 / 3, 0.02 ** 2 / 4
        assert abs(mo[kx] - d * d / 12 - wx) < 5 * math.sqrt(0.04 ** 4 * (1 / 5 - 1 / 9) / N), (psi, mo[kx], wx)
        assert abs(mo[ky] - d * d / 12 - wy) < 5 * math.sqrt(wy * wy * 2 / N), (psi, mo[ky], wy)
    m = _vacuum(nPh=N)
    Lz = m.MC.M.shape[2] * d
    m.MC.LS.sourceType = 5
    m.MC.LS.FPID = SID(XDistr=0, XWidth=0.0, YDistr=0, YWidth=0.0)
    m.MC.LS.AID = SID(XDistr=0, XWidth=0.2, YDistr=1, YWidth=0.3)
    mo = _zpos_moments(run(m, 6), m)
    wx, wy = Lz ** 2 * math.tan(0.2) ** 2 / 3, Lz ** 2 * math.tan(0.3) ** 2 / 4
    assert abs(mo["x2"] - d * d / 12 - wx) < 5 * math.sqrt(wx * wx * 0.8 / N) + 1e-7, (mo["x2"], wx)
    assert abs(mo["y2"] - d * d / 12 - wy) < 5 * math.sqrt(wy * wy * 2 / N) + 1e-7, (mo["y2"], wy)


def lg01(run, N=40000):
    """Laguerre-Gaussian LG01 (sourceType 3): r = w sqrt((W_-1(-U/e) + 1)/-2)
    / 1.50087 (MCmatlablib.c:586).  With s = -(W + 1): (1 + s) e^{-s} = U, so s
    is Gamma(2, 1) distributed and the intensity is the LG01 ring
    r^2 exp(-2 r^2 / c^2), c = w / 1.50087: <r^2> = c^2 and <r^4>/<r^2>^2 = 3/2
    (a Gaussian spot gives 2, a top-hat 4/3).  The angular half of the law
    tilts each ray by phi with <phi^2> = (t0 / 1.50087)^2."""
    from mcmatlab_amd.model import SourceIntensityDistribution as SID

    d = 0.005
    m = _vacuum(n=(61, 61, 20), nPh=N)
    m.MC.LS.sourceType = 3
    m.MC.LS.FPID = SID(radialWidth=0.04)
    m.MC.LS.AID = SID(radialWidth=0.0)
    mo = _zpos_moments(run(m, 7), m)
    c2 = (0.04 / 1.50087) ** 2
    r2 = mo["r2"] - d * d / 6
    assert abs(r2 - c2) < 5 * math.sqrt(c2 * c2 / 2 / N), (r2, c2)  # var(r^2) = c^4 var(s) / 4, var(s) = 2
    # pixel centres: <r_c^4> = <r^4> + (2/3) d^2 <r^2> + O(d^4)
    k = (mo["r4"] - 2 / 3 * d * d * r2) / r2 ** 2
    assert abs(k - 1.5) < 0.08, k
    m = _vacuum(n=(61, 61, 20), nPh=N)
    Lz = 20 * d
    m.MC.LS.sourceType = 3
    m.MC.LS.FPID = SID(radialWidth=0.0)
    m.MC.LS.AID = SID(radialWidth=0.2)
    mo = _zpos_moments(run(m, 8), m)
    c = 0.2 / 1.50087
    want = Lz ** 2 * (c ** 2 + c ** 4 + 17 / 15 * c ** 6)  # E[tan^2 phi] for phi^2 = c^2 s / 2
    got = mo["r2"] - d * d / 6
    assert abs(got - want) < 5 * math.sqrt(2 * want * want / N), (got, want)


def cyclic_wrap(run):
    """Boundary type 3 wraps x and y: an oblique ray through a non-scattering
    absorber crosses the full depth Lz along a path Lz / cos(theta) and leaves
    through the bottom at x = Lz tan(theta) wrapped into the cuboid; with
    boundary type 1 it leaves through the +x face after (Lx / 2) / sin(theta)."""
    mua, d, n = 4.0, 0.005, (20, 20, 20)
    th = 1.0
    L = 20 * d
    for bt in (3, 1):
        m = _slab(n=n, d=d, mua=mua, nPh=500, bt=bt)
        m.MC.LS.theta = th
        r = run(m, 9)
        nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(n, order="F")
        absorbed = float(np.sum(nfr) * mua * d ** 3)
        path = L / math.cos(th) if bt == 3 else (L / 2) / math.sin(th)
        assert abs(absorbed - (1 - math.exp(-mua * path))) < 2e-5 * (1 + absorbed), (bt, absorbed)
        if bt == 3:
            img = _face_image(r, "NI_zpos", m) * d * d
            assert abs(img.sum() - math.exp(-mua * path)) < 2e-5
            xe = ((L * math.tan(th) + L / 2) % L)  # exit x from the cuboid's -x face
            ix, iy = np.unravel_index(np.argmax(img), img.shape)
            assert ix == int(xe / d) and iy == 10, (ix, iy, xe / d)


def too_old(run, N=4000):
    """Boundary type 3 kills packets older than 1000 diagonal light-crossing
    times (MCmatlablib.c:844).  An isotropic source in the middle of an n = 1.5
    non-scattering slab between n = 1 layers: the 3/4 of packets steeper than
    the critical angle (cos theta_c = sqrt(1 - 1/n^2)) are trapped by total
    reflection at both interfaces, absorb 1 - exp(-mua * 1000 diag / n) by the
    time they are killed, and leave nothing through the faces; the others
    escape after absorbing little."""
    from mcmatlab_amd.media import interface_normals

    d, nn = 0.01, 10
    L = nn * d
    M = np.full((nn, nn, nn), 2, np.uint8)
    M[:, :, :2] = 1
    M[:, :, -2:] = 1
    ri = [1.0, 1.5]
    mua = 0.0106
    m = _slab(n=(nn, nn, nn), d=d, mua=mua, nPh=N, bt=3)
    m.MC.M = M
    m.MC.mediaProperties = media_table([mua, mua], [1e-8, 1e-8], [1, 1], ri)
    m.MC.interfaceNormals = interface_normals(M, ri, (d, d, d))
    m.MC.LS.sourceType = 1
    m.MC.LS.zFocus = L / 2
    r = run(m, 10)
    nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(M.shape, order="F")
    absorbed = float(np.sum(nfr) * mua * d ** 3)
    escaped = float((np.sum(r["NI_zpos"]) + np.sum(r["NI_zneg"])) * d * d)
    trapped = math.sqrt(1 - 1 / 1.5 ** 2)
    diag = math.hypot(L, L)
    a_trap = 1 - math.exp(-mua * 1000 * diag / 1.5)
    se = math.sqrt(trapped * (1 - trapped) / N)
    want = trapped * a_trap
    assert want - 4 * se < absorbed < want + 4 * se + 0.006, (absorbed, want)
    lost = 1 - absorbed - escaped  # the weight the killed packets carried
    assert abs(lost - trapped * (1 - a_trap)) < 4 * se + 0.006, (lost, trapped * (1 - a_trap))


def fresnel_oblique(run, N=20000):
    """60 degrees incidence from n = 1 onto a plane n = 1.5 interface:
    R = (rs^2 + rp^2)/2 = 0.0892 leaves through the top face; the refracted
    rays travel at asin(sin 60 / 1.5) = 35.26 degrees (Snell) and leave
    through the bottom at the pixel that angle predicts."""
    from mcmatlab_amd.media import interface_normals

    d, nx, nz, top = 0.005, 81, 24, 4
    th = math.radians(60.0)
    M = np.ones((nx, nx, nz), np.uint8)
    M[:, :, top:] = 2
    ri = [1.0, 1.5]
    m = _slab(n=(nx, nx, nz), d=d, mua=1e-8, mus=1e-8, nPh=N)
    m.MC.M = M
    m.MC.mediaProperties = media_table([1e-8, 1e-8], [1e-8, 1e-8], [1, 1], ri)
    m.MC.interfaceNormals = interface_normals(M, ri, (d, d, d))
    m.MC.LS.theta = th
    r = run(m, 11)
    ci, st = math.cos(th), math.sin(th) / 1.5
    ct = math.sqrt(1 - st * st)
    rs = (ci - 1.5 * ct) / (ci + 1.5 * ct)
    rp = (1.5 * ci - ct) / (1.5 * ci + ct)
    R = (rs * rs + rp * rp) / 2
    refl = float(np.sum(r["NI_zneg"]) * d * d)
    assert abs(refl - R) < 5 * math.sqrt(R * (1 - R) / N), (refl, R)
    img = _face_image(r, "NI_zpos", m) * d * d
    assert abs(img.sum() - (1 - R)) < 5 * math.sqrt(R * (1 - R) / N)
    x_exit = top * d * math.tan(th) + (nz - top) * d * st / ct
    ix, iy = np.unravel_index(np.argmax(img), img.shape)
    assert ix == int(x_exit / d + nx / 2) and iy == nx // 2, (ix, iy, x_exit / d + nx / 2)


def collectors(run, N=40000):
    """Fiber and lens collectors on an isotropic point source in vacuum, a
    height H above (here: below in z) the collector plane.  A ray at polar
    angle a reaches the plane at radius H tan a: the fiber (diameter D,
    numerical aperture NA) collects a < min(atan(D/2H), asin NA), i.e. the
    fraction (1 - cos a_max)/2, all of it ballistic, arriving at H / c.
    A lens whose object plane passes through the source (focal length H)
    images it onto the centre pixel and dates every ray to that plane,
    time 0."""
    d, nn = 0.005, 41
    L = nn * d
    H_face = 0.05
    H = L / 2 + H_face
    for D, NA, nt in ((0.4, 0.2, 5), (0.05, 0.9, 1)):
        m = _vacuum(n=(nn, nn, nn), nPh=N)
        m.MC.LS.sourceType = 1
        m.MC.LS.zFocus = L / 2
        m.MC.useLightCollector = True
        lc = m.MC.LC
        lc.x, lc.y, lc.z, lc.theta, lc.phi = 0.0, 0.0, -H_face, 0.0, math.pi / 2
        lc.f, lc.diam, lc.NA, lc.res = math.inf, D, NA, 1
        t_arr = H / C_LIGHT
        lc.nTimeBins = nt if nt > 1 else 0
        lc.tStart, lc.tEnd = 0.0, 3 * t_arr  # bins 1..nt of width 3 H/c/nt; bin 0 and nt+1 catch the rest
        r = run(m, 12)
        a = min(math.atan(D / 2 / H), math.asin(NA))
        want = (1 - math.cos(a)) / 2
        img = np.asarray(r["image"], dtype=np.float64).ravel()
        assert abs(img.sum() - want) < 5 * math.sqrt(want * (1 - want) / N), (D, NA, img.sum(), want)
        if nt > 1:
            k = 1 + int(nt * t_arr / (3 * t_arr))
            assert img[k] == img.sum(), (img, k)
    m = _vacuum(n=(nn, nn, nn), nPh=N)
    m.MC.LS.sourceType = 1
    m.MC.LS.zFocus = L / 2
    m.MC.useLightCollector = True
    lc = m.MC.LC
    # a lens's (x, y, z) is the centre of its object plane; the lens sits f
    # further along the viewing axis (MCmatlab.c: LC->r = position - f * axis)
    lc.x, lc.y, lc.z, lc.theta, lc.phi = 0.0, 0.0, L / 2, 0.0, math.pi / 2
    lc.f, lc.diam, lc.fieldSize, lc.res = H, 0.2, 0.05, 5
    lc.nTimeBins, lc.tStart, lc.tEnd = 4, -1.5e-12, 2.5e-12
    r = run(m, 13)
    img = np.asarray(r["image"], dtype=np.float64).reshape(5, 5, -1, order="F")
    want = (1 - math.cos(math.atan(0.1 / H))) / 2
    pix = (0.05 / 5) ** 2
    tot = img.sum() * pix
    assert abs(tot - want) < 5 * math.sqrt(want * (1 - want) / N), (tot, want)
    k = 1 + int(4 * 1.5e-12 / 4e-12)  # time 0 in bin 1 + floor(4 * 1.5 / 4) of [-1.5 ps, 2.5 ps]
    assert img[2, 2, k] * pix == tot, (img[..., k], tot)


def unscattered_criteria(run, N=20000):
    """maxScatterings = 0.  Immediate evaluation: only deposits made before a
    packet's first scattering count; on the beam axis
    E[deposit_k] = mua/(mua+mus) (e^{-(mua+mus) z_k} - e^{-(mua+mus)(z_k+dz)})
    and nothing lands off the axis.  End-of-life evaluation: only packets that
    never scattered count (they cross the slab, probability e^{-mus L}), and
    the reference weights each recorded deposit by the packet's final weight
    (MCmatlab.c:262-264): NFR_k V mua = e^{-mus L} e^{-mua L} e^{-mua z_k}(1 - e^{-mua dz})."""
    mua, d, nxy, nz = 10.0, 0.005, 11, 20
    L = nz * d
    z = np.arange(nz) * d
    for at_end, mus in ((False, 60.0), (True, 20.0)):
        m = _slab(n=(nxy, nxy, nz), d=d, mua=mua, mus=mus, g=0.9, nPh=N)
        dc = m.MC.depositionCriteria
        dc.maxScatterings = 0
        dc.evaluateOnlyAtEndOfLife = at_end
        r = run(m, 14)
        nfr = np.asarray(r["NFR"], dtype=np.float64).reshape(nxy, nxy, nz, order="F")
        axis = nfr[nxy // 2, nxy // 2, :] * d ** 3 * mua
        off = nfr.copy()
        off[nxy // 2, nxy // 2, :] = 0
        assert np.all(off == 0), "deposit off the beam axis"
        dmax = np.exp(-mua * z) * (1 - math.exp(-mua * d))  # a packet's largest deposit in layer k
        if not at_end:
            k = mua + mus
            want = mua / k * (np.exp(-k * z) - np.exp(-k * (z + d)))
            se = np.sqrt(want * (dmax - want) / N)  # deposits lie in [0, dmax]
        else:
            p = math.exp(-mus * L)
            want = p * math.exp(-mua * L) * dmax
            se = want * math.sqrt((1 - p) / (p * N))  # a fixed deposit times a Bernoulli(p) count
        sel = slice(0, 6)
        assert np.all(np.abs(axis[sel] - want[sel]) < 5 * se[sel] + 1e-4 * want[sel]), (at_end, axis[sel], want[sel])


def tabulated_phase(run, seeds=(1, 2), nbins=20):
    """A tabulated phase function (g = NaN, 201-point CDF of theta): theta
    is drawn piecewise uniform over the 200 intervals with the CDF's
    probabilities (MCmatlablib.c:1140-1144), so in a thin slab where packets
    scatter at most once the escaped polar histogram's bins hold
    CDF[10(k+1)] - CDF[10k] of the scattered light."""
    mus, d = 2.0, 0.01
    th = math.pi / 200 * (np.arange(1, 201) - 0.5)
    pdf = np.sin(th) * (1 + 3 * np.cos(th) ** 2 + np.cos(th))
    cdf = np.concatenate([[0.0], np.cumsum(pdf) / pdf.sum()])
    res = []
    for s in seeds:
        m = _slab(n=(201, 201, 1), d=d, mua=1e-8, mus=mus, g=1.0, nPh=100000)
        mp = m.MC.mediaProperties
        mp["g"][:, :] = np.nan
        mp["CDFidx"][:, :] = 0
        m.MC.CDFs = cdf.reshape(201, 1)
        m.MC.farFieldRes = nbins
        r = run(m, s)
        res.append(_farfield_theta(r, nbins))
    got = np.mean(res, axis=0)
    got[0] -= math.exp(-mus * d)  # the unscattered beam
    scat = 1 - math.exp(-mus * d)
    want = scat * (cdf[10 * np.arange(1, nbins + 1)] - cdf[10 * np.arange(nbins)])
    N = 100000 * len(seeds)
    se = np.sqrt(np.maximum(want, 1e-12) / N)
    sel = np.r_[1:nbins // 2 - 2, nbins // 2 + 2:nbins - 1]  # away from grazing angles, where the slab edge interferes
    z = (got[sel] - want[sel]) / se[sel]
    assert np.all(np.abs(z) < 5), z


def example_paths(run):
    """Example paths of a pencil beam through a non-scattering absorber: each
    path starts after two NaN columns, every point lies on the beam axis, z
    increases by whole voxels from 0 and the weight is e^{-mua z}
    (addToPhotonPath at launch and after every sub-step, MCmatlablib.c:478-502,
    1117)."""
    mua, d, nz = 10.0, 0.005, 12
    m = _slab(n=(9, 9, nz), d=d, mua=mua, nPh=50)
    m.MC.nExamplePaths = 3
    m.MC.LS.xFocus = 0.001
    r = run(m, 15)
    P = np.asarray(r["examplePaths"], dtype=np.float64).reshape(4, -1, order="F")
    nan = np.isnan(P[0])
    starts = [k for k in range(2, P.shape[1]) if nan[k - 1] and nan[k - 2] and not nan[k]]
    assert nan[0] and nan[1] and len(starts) == 3, starts
    pts = P[:, ~nan]
    assert pts.shape[1] == 3 * (nz + 1), pts.shape
    np.testing.assert_allclose(pts[0], 0.001, rtol=0, atol=2e-7)  # FP32 positions on the GPU
    np.testing.assert_allclose(pts[1], 0.0, rtol=0, atol=2e-7)
    np.testing.assert_allclose(pts[2], np.tile(np.arange(nz + 1) * d, 3), rtol=0, atol=2e-7)
    np.testing.assert_allclose(pts[3], np.exp(-mua * pts[2]), rtol=5e-6)


def _face_fraction(points, L):
    """Mean fraction of isotropic emission from `points` [k, 3] (cube
    [0,L]^3 coordinates) leaving through each face, from the solid angle of an
    axis-aligned rectangle: {'xneg','xpos','yneg','ypos','zneg','zpos'}."""
    def rect(h, a1, a2, b1, b2):
        F = lambda a, b: np.arctan(a * b / (h * np.sqrt(a * a + b * b + h * h)))
        return (F(a2, b2) - F(a1, b2) - F(a2, b1) + F(a1, b1)) / (4 * math.pi)

    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    out = {}
    for ax, (p, q, r_) in enumerate(((x, y, z), (y, x, z), (z, x, y))):
        nm = "xyz"[ax]
        out[nm + "neg"] = float(np.mean(rect(p, -q, L - q, -r_, L - r_)))
        out[nm + "pos"] = float(np.mean(rect(L - p, -q, L - q, -r_, L - r_)))
    return out


def _gauss_voxel(ix, iy, iz, d, k=6):
    g, w = np.polynomial.legendre.leggauss(k)
    u = (g + 1) / 2
    X, Y, Z = np.meshgrid((ix + u) * d, (iy + u) * d, (iz + u) * d, indexing="ij")
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel() / 8
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), W


def source3d(run, N=40000):
    """3D source in vacuum with emission in two voxels, weights 3:1: each face
    receives the weighted mean of the solid-angle fractions of the two
    voxels (exact rectangle solid angles, Gauss-Legendre over each voxel).
    The reference's normalisation divides by the source power twice
    (normfactor = nPhotons / power / power, MCmatlab.c:628, 752 and
    MCmatlablib.c:1193-1194), so the face totals carry power^2."""
    nn, d = 21, 0.005
    L = nn * d
    m = _vacuum(n=(nn, nn, nn), nPh=N)
    V = d ** 3
    S = np.zeros((nn, nn, nn), np.float32)
    S[4, 10, 10] = 3.0 / V
    S[10, 15, 6] = 1.0 / V
    power = float(S.sum(dtype=np.float64) * V)
    m.MC.sourceDistribution = S[..., None]
    r = run(m, 16)
    want = {}
    for (ix, iy, iz), wt in (((4, 10, 10), 0.75), ((10, 15, 6), 0.25)):
        pts, W = _gauss_voxel(ix, iy, iz, d)
        def rect_avg(pts=pts, W=W):
            res = {}
            for k in range(len(W)):
                f = _face_fraction(pts[k:k + 1], L)
                for kk, v in f.items():
                    res[kk] = res.get(kk, 0.0) + W[k] * v
            return res
        for k, v in rect_avg().items():
            want[k] = want.get(k, 0.0) + wt * v
    areas = {"x": d * d, "y": d * d, "z": d * d}
    for face, w_ in want.items():
        got = float(np.sum(r["NI_" + face])) * areas[face[0]] / power ** 2
        assert abs(got - w_) < 5 * math.sqrt(w_ * (1 - w_) / N), (face, got, w_)


def line_source(run, N=40000):
    """Isotropic line emitter of length Le along x (sourceType 1,
    emitterLength, theta = pi/2): positions uniform on the segment; in vacuum
    the face totals are the solid-angle fractions averaged along it."""
    nn, d = 21, 0.005
    L = nn * d
    Le = 0.06
    m = _vacuum(n=(nn, nn, nn), nPh=N)
    m.MC.LS.sourceType = 1
    m.MC.LS.zFocus = L / 2
    m.MC.LS.theta = math.pi / 2
    m.MC.LS.phi = 0.0
    m.MC.LS.emitterLength = Le
    r = run(m, 17)
    g, w = np.polynomial.legendre.leggauss(40)
    xs = L / 2 + (g / 2) * Le
    pts = np.stack([xs, np.full_like(xs, L / 2), np.full_like(xs, L / 2)], 1)
    want = {k: 0.0 for k in ("xneg", "xpos", "yneg", "ypos", "zneg", "zpos")}
    for k in range(len(w)):
        for kk, v in _face_fraction(pts[k:k + 1], L).items():
            want[kk] += w[k] / 2 * v
    for face, w_ in want.items():
        got = float(np.sum(r["NI_" + face])) * d * d
        assert abs(got - w_) < 5 * math.sqrt(w_ * (1 - w_) / N), (face, got, w_)
