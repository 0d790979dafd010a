"""ctypes mirror of include/mcm_abi.h and marshalling from the Python model.

The struct layouts here must match the header field for field; the test
`tests/test_abi.py` checks sizes and symbol exports against the header.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

MCM_OK = 0
MCM_ERR_INVALID_ARGUMENT = -1
MCM_ERR_OUT_OF_MEMORY = -2
MCM_ERR_DEVICE = -3
MCM_ERR_NO_DEVICE = -4
MCM_ERR_UNSUPPORTED = -5

MCM_HEAT_FP64 = 0
MCM_HEAT_FP32 = 1
MCM_HEAT_DGADI = 0
MCM_HEAT_EXPLICIT = 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
INTERRUPT_FN = C.CFUNCTYPE(C.c_int)


class mcm_beam_distr(C.Structure):
    _fields_ = [
        ("radialDistr", _dp), ("n_radialDistr", C.c_int64), ("radialWidth", C.c_double),
        ("XDistr", _dp), ("n_XDistr", C.c_int64), ("XWidth", C.c_double),
        ("YDistr", _dp), ("n_YDistr", C.c_int64), ("YWidth", C.c_double),
    ]


class mcm_mc_input(C.Structure):
    _fields_ = [
        ("simType", C.c_int32), ("silentMode", C.c_int32), ("calcNFR", C.c_int32),
        ("sourceDistribution", _fp),
        ("nPhotonsRequested", C.c_double), ("requestCollectedPhotons", C.c_int32),
        ("simulationTimeRequested", C.c_double),
        ("nM", C.c_int64), ("nL", C.c_int64),
        ("mua", _dp), ("mus", _dp), ("g", _dp), ("n", _dp), ("CDFidx", _dp),
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("farFieldRes", C.c_int64), ("boundaryType", C.c_int32),
        ("interfaceNormals", _fp), ("M", _u8p),
        ("CDFs", _dp), ("nCDFs", C.c_int64),
        ("minScatterings", C.c_double), ("maxScatterings", C.c_double),
        ("minRefractions", C.c_double), ("maxRefractions", C.c_double),
        ("minReflections", C.c_double), ("maxReflections", C.c_double),
        ("minInterfaceTransitions", C.c_double), ("maxInterfaceTransitions", C.c_double),
        ("minSubmediaIdx", C.c_double), ("maxSubmediaIdx", C.c_double),
        ("onlyCollected", C.c_int32), ("evaluateOnlyAtEndOfLife", C.c_int32),
        ("useLightCollector", C.c_int32),
        ("LC_x", C.c_double), ("LC_y", C.c_double), ("LC_z", C.c_double),
        ("LC_theta", C.c_double), ("LC_phi", C.c_double), ("LC_f", C.c_double), ("LC_diam", C.c_double),
        ("LC_fieldSize", C.c_double), ("LC_NA", C.c_double),
        ("LC_res", C.c_double), ("LC_nTimeBins", C.c_double), ("LC_tStart", C.c_double), ("LC_tEnd", C.c_double),
        ("nExamplePaths", C.c_int64),
        ("spectrum", _dp),
        ("useAllCPUs", C.c_int32), ("GPUdevice", C.c_int32),
        ("sourceType", C.c_double), ("emitterLength", C.c_double),
        ("LS_theta", C.c_double), ("LS_phi", C.c_double), ("LS_psi", C.c_double),
        ("xFocus", C.c_double), ("yFocus", C.c_double), ("zFocus", C.c_double),
        ("FPID", mcm_beam_distr), ("AID", mcm_beam_distr),
        ("seed", C.c_uint64),
        ("isInterruptPending", INTERRUPT_FN),
    ]


class mcm_mc_output(C.Structure):
    _fields_ = [
        ("NFR", _fp), ("image", _fp), ("farField", _fp),
        ("NI_xpos", _fp), ("NI_xneg", _fp), ("NI_ypos", _fp), ("NI_yneg", _fp), ("NI_zpos", _fp), ("NI_zneg", _fp),
        ("examplePaths", _dp), ("examplePathsElems", C.c_int64),
        ("nPhotons", C.c_double), ("nPhotonsCollected", C.c_double), ("nThreads", C.c_double),
        ("simulationTime", C.c_double),
    ]


class mcm_mc_sizes(C.Structure):
    _fields_ = [
        ("NFR", C.c_int64), ("image", C.c_int64), ("farField", C.c_int64),
        ("NI_xpos", C.c_int64), ("NI_xneg", C.c_int64), ("NI_ypos", C.c_int64), ("NI_yneg", C.c_int64),
        ("NI_zpos", C.c_int64), ("NI_zneg", C.c_int64),
        ("image_dims", C.c_int64 * 4), ("NI_dims", (C.c_int64 * 3) * 6),
    ]


class mcm_heat_input(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64), ("nM", C.c_int64),
        ("T", _fp), ("Omega", _fp),
        ("steps", C.c_int64), ("dt", C.c_double),
        ("M", _u8p), ("dTdtperdeltaT", _fp), ("dTdt_abs", _fp),
        ("A", _dp), ("E", _dp),
        ("lightsOn", C.c_int32), ("heatBoundaryType", C.c_int32),
        ("tempSensorCornerIdxs", _dp), ("nSensors", C.c_int64), ("tempSensorInterpWeights", _dp),
        ("useAllCPUs", C.c_int32),
        ("precision", C.c_int32), ("scheme", C.c_int32),
        ("T_f64", _dp), ("Omega_f64", _dp),
        ("GPUdevice", C.c_int32),
        ("isInterruptPending", INTERRUPT_FN),
    ]


class mcm_heat_output(C.Structure):
    _fields_ = [
        ("T", _fp), ("Omega", _fp), ("sensorTemps", _dp), ("maxMediaTemps", _fp),
        ("T_f64", _dp), ("Omega_f64", _dp), ("stepsDone", C.c_int64),
    ]


# ----------------------------------------------------------------- marshalling

def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F"))


def _ptr(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def _is_nan_scalar(a):
    if a is None:
        return True
    arr = np.asarray(a)
    return arr.size == 1 and bool(np.isnan(arr.astype(np.float64)).all())


def _bool(v):
    return 1 if bool(v) else 0


class MCInputs:
    """Owns the numpy buffers behind one mcm_mc_input (keep it alive while
    the struct is in use)."""

    def __init__(self, model, simType=1, seed=0, interrupt=None):
        mc = model.FMC if simType == 2 else model.MC
        G = model.G
        keep = []

        def hold(a):
            keep.append(a)
            return a

        s = mcm_mc_input()
        s.simType = simType
        s.silentMode = _bool(mc.silentMode)
        s.calcNFR = _bool(mc.calcNFR)
        if _is_nan_scalar(mc.sourceDistribution):
            s.sourceDistribution = None
        else:
            s.sourceDistribution = _ptr(hold(_f32(mc.sourceDistribution)), C.c_float)
        s.nPhotonsRequested = float(mc.nPhotonsRequested)
        s.requestCollectedPhotons = _bool(mc.requestCollectedPhotons)
        s.simulationTimeRequested = float(mc.simulationTimeRequested)
        mp = mc.mediaProperties
        mua = np.atleast_2d(np.asarray(mp["mua"], dtype=np.float64))
        if mua.shape[0] == 1 and np.asarray(mp["mua"]).ndim == 1:
            mua = mua.T
        nM, nL = mua.shape
        s.nM, s.nL = nM, nL
        for name in ("mua", "mus", "g", "n", "CDFidx"):
            arr = np.asarray(mp[name], dtype=np.float64).reshape(nM, nL, order="F") if np.asarray(mp[name]).size == nM * nL \
                else np.full((nM, nL), np.nan)
            setattr(s, name, _ptr(hold(_f64(arr)), C.c_double))
        M = np.asarray(mc.M)
        if M.ndim != 3:
            M = M.reshape(M.shape + (1,) * (3 - M.ndim))
        s.nx, s.ny, s.nz = M.shape
        s.dx, s.dy, s.dz = float(G.dx), float(G.dy), float(G.dz)
        s.farFieldRes = int(mc.farFieldRes)
        s.boundaryType = int(mc.boundaryType)
        if _is_nan_scalar(mc.interfaceNormals):
            s.interfaceNormals = None
        else:
            s.interfaceNormals = _ptr(hold(_f32(mc.interfaceNormals)), C.c_float)
        s.M = _ptr(hold(np.ascontiguousarray(M.astype(np.uint8).ravel(order="F"))), C.c_uint8)
        cdfs = np.asarray(mc.CDFs, dtype=np.float64) if mc.CDFs is not None else np.zeros(0)
        cdfs = hold(_f64(cdfs)) if cdfs.size else hold(np.zeros(1))
        s.CDFs = _ptr(cdfs, C.c_double)
        s.nCDFs = int(np.asarray(mc.CDFs).size) if mc.CDFs is not None else 0
        dc = mc.depositionCriteria
        for name in ("minScatterings", "maxScatterings", "minRefractions", "maxRefractions",
                     "minReflections", "maxReflections", "minInterfaceTransitions", "maxInterfaceTransitions",
                     "minSubmediaIdx", "maxSubmediaIdx"):
            setattr(s, name, float(getattr(dc, name)))
        s.onlyCollected = _bool(dc.onlyCollected)
        s.evaluateOnlyAtEndOfLife = _bool(dc.evaluateOnlyAtEndOfLife)
        lc = mc.LC
        s.useLightCollector = _bool(mc.useLightCollector)
        for name in ("x", "y", "z", "theta", "phi", "f", "diam", "fieldSize", "NA", "res", "nTimeBins", "tStart", "tEnd"):
            setattr(s, "LC_" + name, float(getattr(lc, name)))
        s.nExamplePaths = int(mc.nExamplePaths)
        s.spectrum = _ptr(hold(_f64(mc.spectrum if mc.spectrum is not None else np.ones(nL))), C.c_double)
        s.useAllCPUs = _bool(mc.useAllCPUs)
        s.GPUdevice = int(mc.GPUdevice)
        ls = mc.LS
        s.sourceType = float(ls.sourceType)
        s.emitterLength = float(ls.emitterLength)
        s.LS_theta, s.LS_phi, s.LS_psi = float(ls.theta), float(ls.phi), float(ls.psi)
        s.xFocus, s.yFocus, s.zFocus = float(ls.xFocus), float(ls.yFocus), float(ls.zFocus)
        for dname in ("FPID", "AID"):
            d = getattr(ls, dname)
            bd = mcm_beam_distr()
            for f in ("radialDistr", "XDistr", "YDistr"):
                arr = hold(_f64(np.atleast_1d(getattr(d, f))))
                setattr(bd, f, _ptr(arr, C.c_double))
                setattr(bd, "n_" + f, int(arr.size))
            bd.radialWidth, bd.XWidth, bd.YWidth = float(d.radialWidth), float(d.XWidth), float(d.YWidth)
            setattr(s, dname, bd)
        s.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        if interrupt is not None:
            cb = INTERRUPT_FN(lambda: 1 if interrupt() else 0)
            hold(cb)
            s.isInterruptPending = cb
        self.struct = s
        self.keep = keep
        self.shape = (int(s.nx), int(s.ny), int(s.nz))
        self.nL = nL
        self.nM = nM


def mc_output_sizes(shape, nL, boundaryType, calcNFR, farFieldRes, useLightCollector, LC_res, LC_nTimeBins, sourceDistributionGiven):
    """Element counts and MATLAB dimensions of every MEX output
    (MCmatlab.c:466-493)."""
    nx, ny, nz = shape
    out = {}
    if calcNFR:
        out["NFR"] = (nx, ny, nz, nL)
    if useLightCollector:
        nTB = 0 if sourceDistributionGiven else int(LC_nTimeBins)
        r = int(LC_res)
        out["image"] = (r, r, nTB + 2 if nTB else 1, nL)
    if farFieldRes:
        out["farField"] = (int(farFieldRes), int(farFieldRes), nL)
    if boundaryType == 1:
        out["NI_xpos"] = out["NI_xneg"] = (ny, nz, nL)
        out["NI_ypos"] = out["NI_yneg"] = (nx, nz, nL)
        out["NI_zpos"] = out["NI_zneg"] = (nx, ny, nL)
    elif boundaryType == 2:
        out["NI_zneg"] = (5 * nx, 5 * ny, nL)
    elif boundaryType == 3:
        out["NI_zpos"] = out["NI_zneg"] = (nx, ny, nL)
    return out


class mcm_mc_grad_input(C.Structure):
    _fields_ = [("muaMap", C.POINTER(C.c_float)), ("box", C.c_int64 * 6), ("iL", C.c_int64), ("nPhotons", C.c_uint64),
                ("seed", C.c_uint64)]


class mcm_mc_grad_output(C.Structure):
    _fields_ = [("J", C.c_double), ("J_se", C.c_double), ("dJ_dmua", C.POINTER(C.c_double)), ("dJ_dlogmus", C.c_double),
                ("dJ_dlogmus_se", C.c_double)]


class mcm_heat_segment(C.Structure):
    _fields_ = [("steps", C.c_int64), ("dt", C.c_double), ("lightsOn", C.c_int32), ("reserved", C.c_int32)]


class mcm_heat_adjoint_output(C.Structure):
    _fields_ = [("J", C.c_double), ("dJ_dQ", C.POINTER(C.c_double)), ("dJ_dT0", C.POINTER(C.c_double)),
                ("Omega_f64", C.POINTER(C.c_double)), ("dJ_dc", C.POINTER(C.c_double))]


class mcm_mc_item(C.Structure):
    _fields_ = [("iL", C.c_int64), ("nPhotons", C.c_uint64), ("seed", C.c_uint64), ("stream", C.c_uint32),
                ("reserved", C.c_uint32)]


class mcm_mc_multi_item(C.Structure):
    _fields_ = [("plan", C.c_int64), ("iL", C.c_int64), ("nPhotons", C.c_uint64), ("seed", C.c_uint64),
                ("stream", C.c_uint32), ("reserved", C.c_uint32)]


class MCOutputs:
    """Caller-owned output buffers for one mcm_mc_output (for `nL`
    wavelengths; default: the input's)."""

    def __init__(self, inp: MCInputs, nL=None):
        s = inp.struct
        dims = mc_output_sizes(inp.shape, inp.nL if nL is None else nL, int(s.boundaryType), bool(s.calcNFR), int(s.farFieldRes),
                               bool(s.useLightCollector), s.LC_res if math.isfinite(s.LC_res) else 0,
                               s.LC_nTimeBins if math.isfinite(s.LC_nTimeBins) else 0,
                               bool(s.sourceDistribution))
        self.dims = dims
        self.arrays = {k: np.zeros(int(np.prod(v)), dtype=np.float32) for k, v in dims.items()}
        o = mcm_mc_output()
        for k, a in self.arrays.items():
            setattr(o, k, _ptr(a, C.c_float))
        self.struct = o

    def result(self, free_fn=None):
        """Dict of MATLAB-shaped (Fortran-order) arrays plus scalars."""
        res = {k: a.reshape(self.dims[k], order="F") for k, a in self.arrays.items()}
        o = self.struct
        n = int(o.examplePathsElems)
        if n and bool(o.examplePaths):
            res["examplePaths"] = np.ctypeslib.as_array(o.examplePaths, shape=(4 * n,)).copy().reshape((4, n), order="F")
            if free_fn is not None:
                free_fn(o.examplePaths)
        for k in ("nPhotons", "nPhotonsCollected", "nThreads", "simulationTime"):
            res[k] = float(getattr(o, k))
        return res


class HeatInputs:
    """Owns the buffers behind one mcm_heat_input built from the MEX
    arguments (T, Omega, P) of finiteElementHeatPropagator."""

    def __init__(self, T, Omega, P, precision=MCM_HEAT_FP64, scheme=MCM_HEAT_DGADI, T_f64=None, Omega_f64=None,
                 interrupt=None):
        keep = []

        def hold(a):
            keep.append(a)
            return a

        T = np.asarray(T)
        if T.ndim != 3:
            T = T.reshape(T.shape + (1,) * (3 - T.ndim))
        s = mcm_heat_input()
        s.nx, s.ny, s.nz = T.shape
        c = np.asarray(P["dTdtperdeltaT"], dtype=np.float32)
        s.nM = c.shape[0]
        s.T = _ptr(hold(_f32(T)), C.c_float)
        if _is_nan_scalar(Omega):
            s.Omega = None
        else:
            s.Omega = _ptr(hold(_f32(Omega)), C.c_float)
        s.steps = int(P["steps"])
        s.dt = float(P["dt"])
        s.M = _ptr(hold(np.ascontiguousarray(np.asarray(P["M"]).astype(np.uint8).ravel(order="F"))), C.c_uint8)
        s.dTdtperdeltaT = _ptr(hold(_f32(c)), C.c_float)
        s.dTdt_abs = _ptr(hold(_f32(P["dTdt_abs"])), C.c_float)
        s.A = _ptr(hold(_f64(P["A"])), C.c_double)
        s.E = _ptr(hold(_f64(P["E"])), C.c_double)
        s.lightsOn = _bool(P["lightsOn"])
        s.heatBoundaryType = int(P["heatBoundaryType"])
        idx = np.atleast_1d(np.asarray(P.get("tempSensorCornerIdxs", np.zeros(0)), dtype=np.float64))
        s.nSensors = int(idx.size)
        s.tempSensorCornerIdxs = _ptr(hold(_f64(idx) if idx.size else np.zeros(1)), C.c_double)
        w = np.asarray(P.get("tempSensorInterpWeights", np.zeros((0, 3))), dtype=np.float64)
        s.tempSensorInterpWeights = _ptr(hold(_f64(w) if w.size else np.zeros(3)), C.c_double)
        s.useAllCPUs = _bool(P.get("useAllCPUs", False))
        s.precision = precision
        s.scheme = scheme
        s.T_f64 = _ptr(hold(_f64(T_f64)), C.c_double) if T_f64 is not None else None
        s.Omega_f64 = _ptr(hold(_f64(Omega_f64)), C.c_double) if Omega_f64 is not None else None
        s.GPUdevice = int(P.get("GPUdevice", 0))
        if interrupt is not None:
            cb = INTERRUPT_FN(lambda: 1 if interrupt() else 0)
            hold(cb)
            s.isInterruptPending = cb
        self.struct = s
        self.keep = keep
        self.shape = (int(s.nx), int(s.ny), int(s.nz))
        self.calcDamage = bool(s.Omega)


class HeatOutputs:
    def __init__(self, inp: HeatInputs, want_f64=False, steps=None):
        s = inp.struct
        L = int(np.prod(inp.shape))
        self.shape = inp.shape
        self.T = np.zeros(L, np.float32)
        self.Omega = np.zeros(L if inp.calcDamage else 1, np.float32)
        self.nS = int(s.nSensors)
        self.steps = int(s.steps) if steps is None else int(steps)
        self.sensor = np.zeros(max(self.nS * (self.steps + 1), 1), np.float64)
        self.maxT = np.zeros(int(s.nM), np.float32)
        o = mcm_heat_output()
        o.T = _ptr(self.T, C.c_float)
        o.Omega = _ptr(self.Omega, C.c_float)
        o.sensorTemps = _ptr(self.sensor, C.c_double)
        o.maxMediaTemps = _ptr(self.maxT, C.c_float)
        self.T64 = self.Om64 = None
        if want_f64:
            self.T64 = np.zeros(L, np.float64)
            o.T_f64 = _ptr(self.T64, C.c_double)
            if inp.calcDamage:
                self.Om64 = np.zeros(L, np.float64)
                o.Omega_f64 = _ptr(self.Om64, C.c_double)
        self.calcDamage = inp.calcDamage
        self.struct = o

    def result(self):
        T = self.T.reshape(self.shape, order="F")
        Om = self.Omega.reshape(self.shape, order="F") if self.calcDamage else self.Omega.reshape(1, 1)
        sens = self.sensor[: self.nS * (self.steps + 1)].reshape((self.nS, self.steps + 1), order="F")
        res = {"T": T, "Omega": Om, "sensorTemps": sens, "maxMediaTemps": self.maxT.reshape(1, -1),
               "stepsDone": int(self.struct.stepsDone)}
        if self.T64 is not None:
            res["T_f64"] = self.T64.reshape(self.shape, order="F")
            if self.Om64 is not None:
                res["Omega_f64"] = self.Om64.reshape(self.shape, order="F")
        return res


def repo_root():
    return os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
