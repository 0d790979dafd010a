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
