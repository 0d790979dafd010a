"""Drive this project's MEX layer (the gateways under mcmatlab_amd/mex, built
against the project's mx-API shim in tests/mexshim) from Python:
model -> mxArray object tree -> mexFunction -> outputs.

Test infrastructure only.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from mcmatlab_amd.model import Model

mxCELL, mxSTRUCT, mxLOGICAL, mxCHAR, mxVOID, mxDOUBLE, mxSINGLE, mxINT8, mxUINT8 = 1, 2, 3, 4, 5, 6, 7, 8, 9
mxINT32 = 12
mxOBJECT = 18

_CLASS_DTYPE = {mxDOUBLE: np.float64, mxSINGLE: np.float32, mxUINT8: np.uint8, mxLOGICAL: np.uint8,
                mxINT32: np.int32, mxINT8: np.int8, mxCHAR: np.uint16}


class MexLib:
    def __init__(self, path):
        self.lib = C.CDLL(path)
        L = self.lib
        vp = C.c_void_p
        L.mxs_create_numeric.restype = vp
        L.mxs_create_numeric.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int64), vp]
        L.mxs_create_object.restype = vp
        L.mxs_create_struct.restype = vp
        L.mxs_set.argtypes = [vp, C.c_char_p, vp]
        L.mxs_get.restype = vp
        L.mxs_get.argtypes = [vp, C.c_char_p]
        L.mxs_class.argtypes = [vp]
        L.mxs_ndim.argtypes = [vp]
        L.mxs_dim.restype = C.c_int64
        L.mxs_dim.argtypes = [vp, C.c_int]
        L.mxs_data.restype = vp
        L.mxs_data.argtypes = [vp]
        L.mxs_numel.restype = C.c_int64
        L.mxs_numel.argtypes = [vp]
        L.mxs_destroy.argtypes = [vp]
        L.mxs_call.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp)]
        L.mxs_last_error.restype = C.c_char_p
        L.mxs_last_error_id.restype = C.c_char_p
        L.mxs_set_interrupt_after.argtypes = [C.c_int64]
        self.mexFunction = C.cast(L.mexFunction, vp)

    # ---- construction
    def num(self, value, cls=mxDOUBLE, shape=None):
        dt = _CLASS_DTYPE[cls]
        a = np.asarray(value, dtype=dt)
        if shape is None:
            shape = a.shape if a.ndim >= 2 else ((1, a.size) if a.ndim == 1 else (1, 1))
        a = np.asfortranarray(a.reshape(shape, order="F"))
        dims = (C.c_int64 * len(shape))(*shape)
        buf = np.ascontiguousarray(a.ravel(order="F"))
        return self.lib.mxs_create_numeric(cls, len(shape), dims, buf.ctypes.data_as(C.c_void_p))

    def logical(self, v):
        return self.num(1 if v else 0, mxLOGICAL)

    def obj(self, fields, struct=False):
        o = self.lib.mxs_create_struct() if struct else self.lib.mxs_create_object()
        for k, v in fields.items():
            self.lib.mxs_set(o, k.encode(), v)
        return o

    # ---- reading
    def get(self, arr, name):
        return self.lib.mxs_get(arr, name.encode())

    def array(self, arr):
        if not arr:
            return None
        cls = self.lib.mxs_class(arr)
        nd = self.lib.mxs_ndim(arr)
        shape = tuple(self.lib.mxs_dim(arr, k) for k in range(nd))
        n = self.lib.mxs_numel(arr)
        dt = _CLASS_DTYPE[cls]
        p = self.lib.mxs_data(arr)
        flat = np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(max(n, 1),))[:n].copy()
        return flat.reshape(shape, order="F")

    def call(self, nlhs, prhs):
        plhs = (C.c_void_p * max(nlhs, 1))()
        pr = (C.c_void_p * len(prhs))(*prhs)
        rc = self.lib.mxs_call(self.mexFunction, nlhs, plhs, len(prhs), pr)
        if rc:
            raise RuntimeError("%s: %s" % (self.lib.mxs_last_error_id().decode(), self.lib.mxs_last_error().decode()))
        return list(plhs)


def _distr(m: MexLib, d):
    return m.obj({
        "radialDistr": m.num(np.atleast_1d(d.radialDistr)), "radialWidth": m.num(d.radialWidth),
        "XDistr": m.num(np.atleast_1d(d.XDistr)), "XWidth": m.num(d.XWidth),
        "YDistr": m.num(np.atleast_1d(d.YDistr)), "YWidth": m.num(d.YWidth),
    })


def _mc_object(m: MexLib, mc):
    ls, lc, dc = mc.LS, mc.LC, mc.depositionCriteria
    LS = m.obj({
        "sourceType": m.num(ls.sourceType), "emitterLength": m.num(ls.emitterLength),
        "FPID": _distr(m, ls.FPID), "AID": _distr(m, ls.AID),
        "theta": m.num(ls.theta), "phi": m.num(ls.phi), "psi": m.num(ls.psi),
        "xFocus": m.num(ls.xFocus), "yFocus": m.num(ls.yFocus), "zFocus": m.num(ls.zFocus),
    })
    LC = m.obj({k: m.num(getattr(lc, k)) for k in
                ("x", "y", "z", "theta", "phi", "f", "diam", "fieldSize", "NA", "res", "nTimeBins", "tStart", "tEnd")})
    m.lib.mxs_set(LC, b"image", m.num(np.nan, mxSINGLE))
    DC = m.obj({k: m.num(getattr(dc, k)) for k in
                ("minScatterings", "maxScatterings", "minRefractions", "maxRefractions", "minReflections",
                 "maxReflections", "minInterfaceTransitions", "maxInterfaceTransitions", "minSubmediaIdx",
                 "maxSubmediaIdx")})
    m.lib.mxs_set(DC, b"onlyCollected", m.logical(dc.onlyCollected))
    m.lib.mxs_set(DC, b"evaluateOnlyAtEndOfLife", m.logical(dc.evaluateOnlyAtEndOfLife))
    mp = mc.mediaProperties
    nM, nL = np.asarray(mp["mua"]).shape
    MP = m.obj({k: m.num(np.asarray(mp[k], dtype=np.float64), shape=(nM, nL)) for k in ("mua", "mus", "g", "n", "CDFidx")},
               struct=True)
    M = np.asarray(mc.M, dtype=np.uint8)
    sd = np.asarray(mc.sourceDistribution, dtype=np.float32)
    IN = np.asarray(mc.interfaceNormals, dtype=np.float32)
    cdfs = np.asarray(mc.CDFs, dtype=np.float64)
    spectrum = np.asarray(mc.spectrum if mc.spectrum is not None else np.ones(nL), dtype=np.float64)
    return m.obj({
        "silentMode": m.logical(mc.silentMode), "calcNFR": m.logical(mc.calcNFR), "LS": LS,
        "sourceDistribution": m.num(sd, mxSINGLE, shape=sd.shape if sd.ndim >= 2 else None),
        "nPhotonsRequested": m.num(mc.nPhotonsRequested),
        "requestCollectedPhotons": m.logical(mc.requestCollectedPhotons),
        "simulationTimeRequested": m.num(mc.simulationTimeRequested),
        "mediaProperties": MP, "M": m.num(M, mxUINT8, shape=M.shape),
        "farFieldRes": m.num(mc.farFieldRes), "boundaryType": m.num(mc.boundaryType),
        "interfaceNormals": m.num(IN, mxSINGLE, shape=IN.shape if IN.ndim >= 2 else None),
        "CDFs": m.num(cdfs, shape=cdfs.shape if cdfs.ndim >= 2 else (cdfs.size and 1, cdfs.size)),
        "depositionCriteria": DC, "LC": LC, "useLightCollector": m.logical(mc.useLightCollector),
        "nExamplePaths": m.num(mc.nExamplePaths), "spectrum": m.num(spectrum),
        "useAllCPUs": m.logical(mc.useAllCPUs), "GPUdevice": m.num(mc.GPUdevice, mxINT32),
        "NFR": m.num(0.0), "farField": m.num(np.nan),
        "nPhotons": m.num(np.nan), "nPhotonsCollected": m.num(np.nan), "nThreads": m.num(np.nan),
        "simulationTime": m.num(np.nan), "examplePaths": m.num(np.nan),
    })


def run_mc_mex(m: MexLib, model: Model, simType=1, interrupt_after=-1):
    """Call `model = MCmatlab(model, simType)` and return the output dict."""
    m.lib.mxs_set_interrupt_after(int(interrupt_after))
    G = m.obj({"dx": m.num(model.G.dx), "dy": m.num(model.G.dy), "dz": m.num(model.G.dz)})
    mc = model.FMC if simType == 2 else model.MC
    top = m.obj({"G": G, "FMC" if simType == 2 else "MC": _mc_object(m, mc)})
    st = m.num(float(simType))
    try:
        (out,) = m.call(1, [top, st])
        mco = m.get(out, "FMC" if simType == 2 else "MC")
        res = {}
        for k in ("NFR", "farField", "NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg",
                  "examplePaths", "nPhotons", "nPhotonsCollected", "nThreads", "simulationTime"):
            a = m.get(mco, k)
            if a:
                res[k] = m.array(a)
        img = m.get(m.get(mco, "LC"), "image")
        if img:
            res["image"] = m.array(img)
        for k in ("nPhotons", "nPhotonsCollected", "nThreads", "simulationTime"):
            if k in res:
                res[k] = float(np.asarray(res[k]).ravel()[0])
        m.lib.mxs_destroy(out)
        return res
    finally:
        m.lib.mxs_destroy(top)
        m.lib.mxs_destroy(st)


def run_heat_mex(m: MexLib, T, Omega, P, interrupt_after=-1):
    """Call `[T, Omega, sensorTemps, maxMediaTemps] = finiteElementHeatPropagator(T, Omega, P)`."""
    m.lib.mxs_set_interrupt_after(int(interrupt_after))
    T = np.asarray(T, dtype=np.float32)
    Om = np.asarray(Omega, dtype=np.float32)
    fields = {
        "M": m.num(np.asarray(P["M"], dtype=np.uint8), mxUINT8, shape=np.asarray(P["M"]).shape),
        "A": m.num(np.asarray(P["A"], dtype=np.float64)), "E": m.num(np.asarray(P["E"], dtype=np.float64)),
        "dTdtperdeltaT": m.num(np.asarray(P["dTdtperdeltaT"], dtype=np.float32), mxSINGLE,
                               shape=np.asarray(P["dTdtperdeltaT"]).shape),
        "dTdt_abs": m.num(np.asarray(P["dTdt_abs"], dtype=np.float32), mxSINGLE, shape=np.asarray(P["dTdt_abs"]).shape),
        "useAllCPUs": m.logical(P.get("useAllCPUs", False)),
        "heatBoundaryType": m.num(P["heatBoundaryType"]),
        "tempSensorCornerIdxs": m.num(np.asarray(P.get("tempSensorCornerIdxs", np.zeros((0, 1))), dtype=np.float64).reshape(-1, 1)),
        "tempSensorInterpWeights": m.num(np.asarray(P.get("tempSensorInterpWeights", np.zeros((0, 3))), dtype=np.float64).reshape(-1, 3)),
        "lightsOn": m.logical(P["lightsOn"]), "steps": m.num(P["steps"]), "dt": m.num(P["dt"]),
    }
    if "precision" in P:  # optional: 'single' selects the reference's FP32 arithmetic
        fields["precision"] = m.num(np.frombuffer(P["precision"].encode(), np.uint8).astype(np.uint16), mxCHAR)
    Ps = m.obj(fields, struct=True)
    Tm = m.num(T, mxSINGLE, shape=T.shape)
    Omm = m.num(Om, mxSINGLE, shape=Om.shape if Om.ndim >= 2 else None)
    try:
        outs = m.call(4, [Tm, Omm, Ps])
        res = {"T": m.array(outs[0]), "Omega": m.array(outs[1]), "sensorTemps": m.array(outs[2]),
               "maxMediaTemps": m.array(outs[3])}
        for o in outs:
            m.lib.mxs_destroy(o)
        return res
    finally:
        for a in (Tm, Omm, Ps):
            m.lib.mxs_destroy(a)
