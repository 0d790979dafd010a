"""Small model factories covering the MEX feature matrix (sources 0-5 and 3D,
boundary types 0-3, Fresnel, light collectors, far field, deposition
criteria, example paths, custom phase functions, several wavelengths).
Sizes are chosen so the single-threaded reference build finishes each case in
well under a second."""
from __future__ import annotations

import math

import numpy as np

from mcmatlab_amd.media import interface_normals
from mcmatlab_amd.model import Geometry, Model, SourceIntensityDistribution as SID, media_table
from mcmatlab_amd.workloads import layered_media_map


def base(n=20, bt=1, nph=2e3, matched=True, nL=1):
    side = 0.1
    d = side / n
    m = Model()
    m.G = Geometry(dx=d, dy=d, dz=d)
    M = layered_media_map(n, side)
    ri = [1.0, 1.0, 1.0, 1.0] if matched else [1.0, 1.4, 1.4, 1.38]
    mua = np.array([1e-8, 20.0, 0.5, 230.0])
    mus = np.array([1e-8, 200.0, 200.0, 470.0])
    if nL > 1:
        scale = np.linspace(0.5, 1.5, nL)
        mua = mua[:, None] * scale[None, :]
        mus = mus[:, None] * scale[None, ::-1]
    m.MC.M = M
    m.MC.mediaProperties = media_table(mua, mus, [1, .9, .9, .9], ri)
    m.MC.interfaceNormals = np.float32(np.nan) if matched else interface_normals(M, ri, (d, d, d))
    m.MC.nPhotonsRequested = float(nph)
    m.MC.boundaryType = bt
    m.MC.spectrum = np.ones(nL) / nL
    return m


def _custom_cdf_media(m):
    """Medium 3 (dermis) uses a tabulated phase function p(theta) ~ 1 + cos^2."""
    th = math.pi / 200 * (np.arange(1, 201) - 0.5)
    pdf = np.sin(th) * (1 + np.cos(th) ** 2)
    cdf = np.concatenate([[0.0], np.cumsum(pdf) / pdf.sum()])
    mp = m.MC.mediaProperties
    mp["g"][2, :] = np.nan
    mp["CDFidx"][:, :] = np.nan
    mp["CDFidx"][2, :] = 0
    m.MC.CDFs = cdf.reshape(201, 1)
    return m


def cases():
    """name -> zero-arg model factory"""
    C = {}
    for bt in (0, 1, 2, 3):
        C[f"pencil_bt{bt}"] = (lambda bt=bt: base(bt=bt))
    C["pencil_oblique"] = lambda: _set(base(), LS_theta=0.4, LS_phi=0.7, xFocus=0.01, zFocus=0.02)
    C["fresnel_bt1"] = lambda: _set(base(matched=False), LS_theta=0.3, LS_phi=0.5)
    C["fresnel_bt3"] = lambda: _set(base(matched=False, bt=3), sourceType=4, FPID=SID(radialDistr=0, radialWidth=0.02),
                                    AID=SID(radialDistr=1, radialWidth=0.2))
    C["isotropic_point"] = lambda: _set(base(), sourceType=1, zFocus=0.05)
    C["isotropic_line"] = lambda: _set(base(), sourceType=1, zFocus=0.05, emitterLength=0.04, LS_theta=0.5, LS_phi=0.2)
    C["planewave_bt1"] = lambda: _set(base(), sourceType=2, LS_theta=0.2)
    C["planewave_bt2"] = lambda: _set(base(bt=2), sourceType=2)
    C["radial_tophat_gauss"] = lambda: _set(base(), sourceType=4, FPID=SID(radialDistr=0, radialWidth=0.03),
                                            AID=SID(radialDistr=1, radialWidth=0.1))
    C["radial_gauss_lambert"] = lambda: _set(base(), sourceType=4, FPID=SID(radialDistr=1, radialWidth=0.02),
                                             AID=SID(radialDistr=2, radialWidth=0.0), zFocus=0.03)
    C["radial_custom"] = lambda: _set(base(), sourceType=4, FPID=SID(radialDistr=np.array([1, 2, 3, 2, 1, 0.5]), radialWidth=0.03),
                                      AID=SID(radialDistr=np.array([3, 2, 1, 0.0]), radialWidth=0.3))
    C["xy_tophat_gauss"] = lambda: _set(base(), sourceType=5, LS_psi=0.3,
                                        FPID=SID(XDistr=0, XWidth=0.02, YDistr=1, YWidth=0.01),
                                        AID=SID(XDistr=1, XWidth=0.1, YDistr=0, YWidth=0.2))
    C["xy_custom_lambert"] = lambda: _set(base(), sourceType=5,
                                          FPID=SID(XDistr=np.array([1, 3, 1.0]), XWidth=0.02, YDistr=np.array([0, 1, 2, 1, 0.0]), YWidth=0.03),
                                          AID=SID(XDistr=2, XWidth=0.0, YDistr=2, YWidth=0.0))
    C["xy_custom_angular"] = lambda: _set(base(), sourceType=5, FPID=SID(XDistr=0, XWidth=0.01, YDistr=0, YWidth=0.01),
                                          AID=SID(XDistr=np.array([1, 2, 1.0]), XWidth=0.2, YDistr=np.array([2, 1.0]), YWidth=0.1))
    C["lens_timebins"] = lambda: _lc(base(), f=0.5, diam=0.5, fieldSize=0.1, res=8, nTimeBins=6, tEnd=2e-11, z=-0.3)
    C["fiber"] = lambda: _lc(base(), f=math.inf, diam=0.05, NA=0.5, res=1, z=-0.05, tEnd=1e-11, nTimeBins=3)
    C["farfield"] = lambda: _set(base(), farFieldRes=12)
    C["criteria_scatter"] = lambda: _dc(base(), minScatterings=3, maxScatterings=40)
    C["criteria_collected"] = lambda: _dc(_lc(base(), f=math.inf, diam=0.1, NA=0.9, res=1, z=-0.02), onlyCollected=True)
    C["criteria_now"] = lambda: _dc(base(), minScatterings=2, evaluateOnlyAtEndOfLife=False, minSubmediaIdx=2, maxSubmediaIdx=3)
    C["paths"] = lambda: _set(base(), nExamplePaths=4)
    C["paths_cyclic"] = lambda: _set(base(bt=3, matched=False), nExamplePaths=3)
    C["custom_phase"] = lambda: _custom_cdf_media(base())
    C["wavelengths3"] = lambda: base(nL=3, nph=3001)
    C["source3d"] = lambda: _source3d(base())
    C["lg01"] = lambda: _set(base(), sourceType=3, FPID=SID(radialWidth=0.02), AID=SID(radialWidth=0.1), LS_theta=0.2, LS_phi=0.4)
    C["fmc_source3d"] = lambda: _fluorescence(_source3d(base()))
    return C


def _fluorescence(m):
    """The fluorescence simulation (simType 2) reads model.FMC: the same
    inputs moved there, the emitter distribution as its source
    (runMonteCarlo.m:116-175 calls the MEX with simType 2)."""
    m.FMC = m.MC.copy()
    return m


def _set(m, **kw):
    ls = m.MC.LS
    for k, v in kw.items():
        if k.startswith("LS_"):
            setattr(ls, k[3:], v)
        elif hasattr(ls, k) or k in ("FPID", "AID", "sourceType", "emitterLength", "xFocus", "yFocus", "zFocus"):
            setattr(ls, k, v)
        else:
            setattr(m.MC, k, v)
    return m


def _lc(m, **kw):
    m.MC.useLightCollector = True
    for k, v in kw.items():
        setattr(m.MC.LC, k, v)
    return m


def _dc(m, **kw):
    for k, v in kw.items():
        setattr(m.MC.depositionCriteria, k, v)
    return m


def _source3d(m):
    """3D emitter distribution (the fluorescence path's source): emission
    proportional to the local absorption coefficient inside the vessel."""
    M = m.MC.M
    mua = np.asarray(m.MC.mediaProperties["mua"])[:, 0]
    S = (mua[M - 1] * (M == 4)).astype(np.float32)
    S[S == 0] = 1e-3
    m.MC.sourceDistribution = S[..., None]
    return m
