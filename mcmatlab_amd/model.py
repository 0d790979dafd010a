"""Python mirror of the MCmatlab.model object tree that the two MEX entry
points read and write.

Field names, meanings and defaults follow the reference's MATLAB value
classes (+MCmatlab/@monteCarloSimulation/monteCarloSimulation.m:5-85,
@lightSource/lightSource.m:5-22, @lightCollector/lightCollector.m:7-31,
@depositionCriteria/depositionCriteria.m:31-51,
@sourceIntensityDistribution/sourceIntensityDistribution.m:5-14,
@geometry/geometry.m:8-41, @model/model.m:7-13).  The hidden short aliases
the MEX files use (LS, LC, FPID, AID, NFR, calcNFR, NI_*) are provided as
properties.  Only the fields on the MEX boundary are modelled; the MATLAB
geometry/media authoring layer (geomFunc, mediaPropertiesFunc, plotting) is
not part of this package.

Arrays are numpy arrays in MATLAB (column-major) index order: M[ix, iy, iz].
"""
from __future__ import annotations

import copy
import math

import numpy as np

NaN = float("nan")
Inf = float("inf")


class _Props:
    _aliases: dict = {}

    def __getattr__(self, name):
        al = type(self)._aliases.get(name)
        if al is not None:
            return getattr(self, al)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        al = type(self)._aliases.get(name)
        object.__setattr__(self, al if al is not None else name, value)

    def copy(self):
        return copy.deepcopy(self)


class SourceIntensityDistribution(_Props):
    """MCmatlab.sourceIntensityDistribution (focal-plane or angular)."""

    def __init__(self, **kw):
        self.radialDistr = NaN   # 0 top-hat, 1 Gaussian, 2 Lambertian, or a custom array
        self.radialWidth = NaN
        self.XDistr = NaN
        self.XWidth = NaN
        self.YDistr = NaN
        self.YWidth = NaN
        for k, v in kw.items():
            setattr(self, k, v)


class LightSource(_Props):
    """MCmatlab.lightSource."""
    _aliases = {"FPID": "focalPlaneIntensityDistribution", "AID": "angularIntensityDistribution"}

    def __init__(self, **kw):
        self.sourceType = 0
        self.xFocus = 0.0
        self.yFocus = 0.0
        self.zFocus = 0.0
        self.theta = 0.0
        self.phi = 0.0
        self.psi = 0.0
        self.emitterLength = 0.0
        self.focalPlaneIntensityDistribution = SourceIntensityDistribution()
        self.angularIntensityDistribution = SourceIntensityDistribution()
        for k, v in kw.items():
            setattr(self, k, v)


class LightCollector(_Props):
    """MCmatlab.lightCollector."""

    def __init__(self, **kw):
        self.x = 0.0
        self.y = 0.0
        self.z = 0.0
        self.theta = 0.0
        self.phi = math.pi / 2
        self.f = Inf
        self.diam = 1.0
        self.fieldSize = 1.0
        self.NA = 0.22
        self.res = 0
        self.tStart = 0.0
        self.tEnd = 1e-12
        self.nTimeBins = 0
        self.image = NaN
        for k, v in kw.items():
            setattr(self, k, v)


class DepositionCriteria(_Props):
    """MCmatlab.depositionCriteria.  minSubmediaIdx/maxSubmediaIdx are the
    hidden 1-based (sub-)medium bounds the MEX reads; they default to the
    whole media list."""

    def __init__(self, **kw):
        self.minScatterings = 0
        self.maxScatterings = Inf
        self.minRefractions = 0
        self.maxRefractions = Inf
        self.minReflections = 0
        self.maxReflections = Inf
        self.minInterfaceTransitions = 0
        self.maxInterfaceTransitions = Inf
        self.minSubmediaIdx = 1
        self.maxSubmediaIdx = Inf
        self.onlyCollected = False
        self.evaluateOnlyAtEndOfLife = True
        for k, v in kw.items():
            setattr(self, k, v)


class MonteCarloSimulation(_Props):
    """MCmatlab.monteCarloSimulation (model.MC) / fluorescence (model.FMC)."""
    _aliases = {
        "LS": "lightSource", "LC": "lightCollector", "NFR": "normalizedFluenceRate",
        "calcNFR": "calcNormalizedFluenceRate",
        "NI_xpos": "normalizedIrradiance_xpos", "NI_xneg": "normalizedIrradiance_xneg",
        "NI_ypos": "normalizedIrradiance_ypos", "NI_yneg": "normalizedIrradiance_yneg",
        "NI_zpos": "normalizedIrradiance_zpos", "NI_zneg": "normalizedIrradiance_zneg",
    }

    def __init__(self, **kw):
        self.silentMode = True
        self.useAllCPUs = False
        self.useGPU = True
        self.GPUdevice = 0
        self.simulationTimeRequested = 0.1
        self.nPhotonsRequested = NaN
        self.requestCollectedPhotons = False
        self.calcNormalizedFluenceRate = True
        self.nExamplePaths = 0
        self.farFieldRes = 0
        self.matchedInterfaces = True
        self.boundaryType = 1
        self.wavelength = 800.0
        self.P = 1.0
        self.lightSource = LightSource()
        self.useLightCollector = False
        self.lightCollector = LightCollector()
        self.depositionCriteria = DepositionCriteria()
        self.sourceDistribution = np.float32(NaN)
        # inputs normally produced by getOpticalMediaProperties / runMonteCarlo
        self.mediaProperties = None   # dict mua, mus, g, n, CDFidx: [nM, nL]
        self.CDFs = np.zeros((0,))
        self.M = None                 # uint8 [nx, ny, nz], 1-based
        self.interfaceNormals = np.float32(NaN)
        self.spectrum = None          # [nL]
        # outputs
        self.simulationTime = NaN
        self.nPhotons = NaN
        self.nPhotonsCollected = NaN
        self.nThreads = NaN
        self.examplePaths = NaN
        self.normalizedFluenceRate = 0
        self.farField = NaN
        self.normalizedIrradiance_xpos = NaN
        self.normalizedIrradiance_xneg = NaN
        self.normalizedIrradiance_ypos = NaN
        self.normalizedIrradiance_yneg = NaN
        self.normalizedIrradiance_zpos = NaN
        self.normalizedIrradiance_zneg = NaN
        for k, v in kw.items():
            setattr(self, k, v)


class Geometry(_Props):
    """MCmatlab.geometry: only the voxel pitch reaches the MEX files."""

    def __init__(self, dx=0.01, dy=0.01, dz=0.01, **kw):
        self.dx = dx
        self.dy = dy
        self.dz = dz
        for k, v in kw.items():
            setattr(self, k, v)


class Model(_Props):
    """MCmatlab.model (G, MC, FMC; the heat simulation is driven through
    finiteElementHeatPropagator's (T, Omega, P) arguments)."""

    def __init__(self, **kw):
        self.G = Geometry()
        self.MC = MonteCarloSimulation()
        self.FMC = MonteCarloSimulation()
        for k, v in kw.items():
            setattr(self, k, v)


def media_table(mua, mus, g, n, CDFidx=None):
    """mediaProperties struct from per-medium lists (one wavelength) or
    [nM, nL] arrays."""
    mua = np.asarray(mua, dtype=np.float64)
    if mua.ndim == 1:
        mua = mua[:, None]
    shp = mua.shape

    def col(v):
        v = np.asarray(v, dtype=np.float64)
        return np.broadcast_to(v[:, None] if v.ndim == 1 else v, shp).copy()

    return {"mua": mua, "mus": col(mus), "g": col(g), "n": col(n),
            "CDFidx": col(CDFidx if CDFidx is not None else np.ones(shp[0]))}
