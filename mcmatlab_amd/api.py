"""The two operator entry points, with the reference MEX files' names,
arguments and outputs:

  model = MCmatlab(model, simType)
      (+MCmatlab/src/MCmatlab.c:317; called by @model/runMonteCarlo.m:99-104,
      :168-175) -- voxel Monte Carlo on the GPU; returns a copy of the model
      with MC (simType 1) or FMC (simType 2) outputs filled in: NFR,
      farField, NI_xpos ... NI_zneg, LC.image, examplePaths, nPhotons,
      nPhotonsCollected, nThreads, simulationTime.

  T, Omega, sensorTemps, maxMediaTemps = finiteElementHeatPropagator(T, Omega, P)
      (+MCmatlab/src/finiteElementHeatPropagator.c:688; called by
      @model/simulateHeatDistribution.m:259-264) -- `steps` heat steps of
      size dt on the GPU.

  downsampledParamVals, binidx = getDownsampledParamVals(paramVals, N2)
      (+MCmatlab/src/getDownsampledParamVals.c:54; called by
      @model/private/getOpticalMediaProperties.m:157 and
      getThermalMediaProperties.m:80) -- farthest-point binning of a
      medium's per-voxel parameter sets on the GPU.

  model = fr_dependent_MCmatlab(model, medium, params_of_FR, nBins, nIterations)
      (runMonteCarlo.m:80-113) -- the fluence-rate-dependent iteration for
      one medium: re-binned sub-media before every run, packet ramp.

  T, Omega, M = t_dependent_heat(T, Omega, P, M, VHC, TC, d, medium, thermal_of_T, nBins, updates)
      (simulateHeatDistribution.m:515-541) -- heat updates with one
      temperature-dependent medium re-binned before each update.

Errors raise MCMError (the MEX files raise MATLAB errors via
mexErrMsgIdAndTxt).  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import abi
from ._lib import check, lib

_OUT_FIELDS = ("NFR", "farField", "NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg")


def _squeeze_trailing(a):
    """MATLAB drops trailing singleton dimensions beyond the second."""
    shape = list(a.shape)
    while len(shape) > 2 and shape[-1] == 1:
        shape.pop()
    return a.reshape(shape, order="F")


def MCmatlab(model, simType=1, seed=None, interrupt=None):
    """model = MCmatlab(model, simType).  `seed` keys the packet streams
    (the reference seeds from the clock; default: the clock)."""
    if simType not in (1, 2):
        raise ValueError("simType must be 1 (MC) or 2 (fluorescence MC)")
    if seed is None:
        seed = time.time_ns() & 0xFFFFFFFFFFFFFFFF
    out_model = model.copy()
    inp = abi.MCInputs(model, simType=simType, seed=seed, interrupt=interrupt)
    sizes = abi.mcm_mc_sizes()
    check(lib().mcm_mc_output_sizes(C.byref(inp.struct), C.byref(sizes)))
    outs = abi.MCOutputs(inp)
    check(lib().mcm_mc_run(C.byref(inp.struct), C.byref(outs.struct)))
    res = outs.result(free_fn=lambda p: lib().mcm_free(C.cast(p, C.c_void_p)))
    mc = out_model.FMC if simType == 2 else out_model.MC
    for k in _OUT_FIELDS:
        if k in res:
            setattr(mc, k, _squeeze_trailing(res[k]))
    if "image" in res:
        img = _squeeze_trailing(res["image"])
        mc.LC.image = img.reshape(()) if img.size == 1 else img
    if "examplePaths" in res:
        mc.examplePaths = res["examplePaths"]
    for k in ("nPhotons", "nPhotonsCollected", "nThreads", "simulationTime"):
        setattr(mc, k, res[k])
    return out_model


def finiteElementHeatPropagator(T, Omega, P, precision="fp64", scheme="dgadi", state_f64=None, interrupt=None):
    """[T, Omega, sensorTemps, maxMediaTemps] = finiteElementHeatPropagator(T, Omega, P).

    precision: 'fp64' (device state in FP64, default) or 'fp32' (the
    reference's FP32 arithmetic).  scheme: 'dgadi' (the reference's
    Douglas-Gunn ADI, default) or 'explicit' (forward Euler; stable only for
    P['dt'] <= dtmax).  state_f64=(T64, Omega64) continues an FP64 state from
    a previous call; the FP64 final state is returned in the dict
    `finiteElementHeatPropagator.last_state`."""
    prec = {"fp64": abi.MCM_HEAT_FP64, "fp32": abi.MCM_HEAT_FP32}[precision]
    sch = {"dgadi": abi.MCM_HEAT_DGADI, "explicit": abi.MCM_HEAT_EXPLICIT}[scheme]
    T64, Om64 = (state_f64 if state_f64 is not None else (None, None))
    inp = abi.HeatInputs(T, Omega, P, precision=prec, scheme=sch, T_f64=T64, Omega_f64=Om64, interrupt=interrupt)
    out = abi.HeatOutputs(inp, want_f64=True)
    check(lib().mcm_heat_run(C.byref(inp.struct), C.byref(out.struct)))
    r = out.result()
    finiteElementHeatPropagator.last_state = (r.get("T_f64"), r.get("Omega_f64"))
    return r["T"], r["Omega"], r["sensorTemps"], r["maxMediaTemps"]


finiteElementHeatPropagator.last_state = (None, None)


def multi_run(plans, items):
    """mcm_mc_multi_run: items = [(plan, iL, nPhotons, seed, stream), ...]
    over MCPlans of different grids, run in one sequence of persistent
    launches per interface class; returns one output dict per item (one
    wavelength each, normalised like MCmatlab)."""
    plans = list(plans)
    hs = (C.c_void_p * len(plans))(*[p.h.value for p in plans])
    n = len(items)
    arr = (abi.mcm_mc_multi_item * n)()
    for k, (pi, iL, nph, seed, stream) in enumerate(items):
        a = arr[k]
        a.plan, a.iL, a.nPhotons, a.seed, a.stream, a.reserved = int(pi), int(iL), int(nph), int(seed), int(stream), 0
    # (an out-of-range plan index is reported by the library)
    outs = [abi.MCOutputs(plans[int(it[0]) if 0 <= int(it[0]) < len(plans) else 0]._inp, nL=1) for it in items]
    structs = (abi.mcm_mc_output * n)(*[o.struct for o in outs])
    check(lib().mcm_mc_multi_run(hs, len(plans), arr, n, structs))
    res = []
    for o, st in zip(outs, structs):
        o.struct = st
        res.append(o.result())
    return res


class HeatPlan:
    """Device-resident heat propagation (mcm_heat_plan_*): upload T, Omega
    and the model once, advance any schedule of (steps, dt, lightsOn) runs,
    download when needed.  Runs of 2 * graph_steps steps or more are replayed
    as captured HIP graphs."""

    def __init__(self, T, Omega, P, precision="fp64", scheme="dgadi", graph_steps=None):
        prec = {"fp64": abi.MCM_HEAT_FP64, "fp32": abi.MCM_HEAT_FP32}[precision]
        sch = {"dgadi": abi.MCM_HEAT_DGADI, "explicit": abi.MCM_HEAT_EXPLICIT}[scheme]
        self._inp = abi.HeatInputs(T, Omega, P, precision=prec, scheme=sch)
        h = C.c_void_p()
        check(lib().mcm_heat_plan_create(C.byref(self._inp.struct), C.byref(h)))
        self.h = h
        self.steps = 0
        if graph_steps is not None:
            self.set_graph_steps(graph_steps)

    def set_graph_steps(self, k):
        check(lib().mcm_heat_plan_set_graph_steps(self.h, int(k)))

    def set_explicit_zrun(self, z):
        """Explicit scheme: z planes per workgroup of the stencil (8, 12, 16,
        24 or 32); results do not depend on it."""
        check(lib().mcm_heat_plan_set_explicit_zrun(self.h, int(z)))

    @property
    def explicit_zrun(self):
        return int(lib().mcm_heat_plan_explicit_zrun(self.h))

    def advance(self, steps, dt, lightsOn):
        check(lib().mcm_heat_plan_advance(self.h, int(steps), float(dt), int(bool(lightsOn))))
        self.steps += int(steps)

    def download(self, f64=True):
        out = abi.HeatOutputs(self._inp, want_f64=f64, steps=self.steps)
        check(lib().mcm_heat_plan_download(self.h, C.byref(out.struct)))
        return out.result()

    def close(self):
        if self.h:
            lib().mcm_heat_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HeatGroup:
    """HeatPlans advanced together (mcm_heat_group_*): every run of
    steps_per_graph steps of all plans is one graph launch whose per-plan
    branches run concurrently.  The plans stay owned by the caller."""

    def __init__(self, plans, steps_per_graph=32):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.h.value for p in self.plans])
        h = C.c_void_p()
        check(lib().mcm_heat_group_create(arr, len(self.plans), C.byref(h)))
        self.h = h
        self.k = int(steps_per_graph)

    def advance(self, steps, dts, lights):
        n = len(self.plans)
        d = (C.c_double * n)(*[float(x) for x in dts])
        l = (C.c_int32 * n)(*[int(bool(x)) for x in lights])
        check(lib().mcm_heat_group_advance(self.h, int(steps), d, l, self.k))
        for p in self.plans:
            p.steps += int(steps)

    def close(self):
        if self.h:
            lib().mcm_heat_group_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def getDownsampledParamVals(paramVals, N2, device=0):
    """[downsampledParamVals, binidx] = getDownsampledParamVals(paramVals, N2)
    on GPU `device`: paramVals is 3 x N1 (rows mua, mus, g, or VHC, TC, 1),
    single precision; returns the 3 x N2 single downsampled sets and the
    1-based uint8 label (length N1) of the set each input set is nearest to,
    bit for bit the reference's outputs."""
    pv = np.asarray(paramVals, dtype=np.float32)
    if pv.ndim != 2 or pv.shape[0] != 3:
        raise ValueError("paramVals must be 3 x N1")
    n1, n2 = int(pv.shape[1]), int(N2)
    flat = np.ascontiguousarray(pv.ravel(order="F"))
    out = np.zeros(3 * max(n2, 1), np.float32)
    bins = np.zeros(max(n1, 1), np.uint8)
    check(lib().mcm_downsample_params(flat.ctypes.data, n1, n2, out.ctypes.data, bins.ctypes.data, int(device)))
    return out[:3 * n2].reshape(3, n2, order="F"), bins[:n1]


def fr_dependent_MCmatlab(model, medium, params_of_FR, nBins, nIterations, seed=None):
    """The fluence-rate-dependent iteration of runMonteCarlo.m:80-113 for one
    FR-dependent medium (single wavelength, simType 1), around the GPU
    operators: run i of nIterations launches nPhotonsRequested / 2^(nIterations
    - i) packets; before each run the medium's voxels get their (mua, mus, g)
    from the current fluence rate FR = P * NFR (zero before the first run)
    through params_of_FR(FR) -> (mua, mus, g) arrays, are binned into nBins
    sub-media (getDownsampledParamVals on the GPU with the callers' sortrows
    relabelling, getOpticalMediaProperties.m:148-163), and the media table and
    media map are rebuilt with those sub-media in the medium's place (later
    media shift by nBins - 1; sub-media keep the medium's n and CDFidx).
    `medium` is 1-based as in model.MC.M.  Returns the last run's model, whose
    MC.M and MC.mediaProperties hold the split media."""
    from .media import bin_medium

    M0 = np.asarray(model.MC.M)
    mp0 = {k: np.asarray(v, dtype=np.float64).reshape(np.asarray(v).shape[0], -1)[:, :1]
           for k, v in model.MC.mediaProperties.items()}
    nM = mp0["mua"].shape[0]
    if not 1 <= medium <= nM:
        raise ValueError("medium out of range")
    if nM - 1 + nBins > 255:
        raise ValueError("more than 255 media after splitting")
    vox = (M0 == medium).ravel(order="F")
    if not vox.any():
        raise ValueError("the medium has no voxels")
    nPh = float(model.MC.nPhotonsRequested)
    FR = np.zeros(int(vox.sum()))
    res = None
    for i in range(1, int(nIterations) + 1):
        mua, mus, g = (np.broadcast_to(np.asarray(v, np.float64), FR.shape) for v in params_of_FR(FR))
        ds, lab = bin_medium(np.vstack([mua, mus, g]).astype(np.float32), nBins)
        mp = {}
        for k, col in mp0.items():
            sub = ds[{"mua": 0, "mus": 1, "g": 2}[k]].astype(np.float64)[:, None] if k in ("mua", "mus", "g") \
                else np.repeat(col[medium - 1:medium], nBins, axis=0)
            mp[k] = np.concatenate([col[:medium - 1], sub, col[medium:]], axis=0)
        Mf = M0.ravel(order="F").astype(np.int64)
        Mf = np.where(Mf > medium, Mf + nBins - 1, Mf)
        Mf[vox] = medium + lab - 1
        m = model.copy()
        m.MC.M = Mf.reshape(M0.shape, order="F").astype(np.uint8)
        m.MC.mediaProperties = mp
        # deposition-criteria medium bounds follow the split: a lower bound on
        # the medium starts at its first sub-medium, an upper bound ends at
        # its last, later media move up (getOpticalMediaProperties.m:126-133)
        dc = m.MC.depositionCriteria
        lo, hi = float(dc.minSubmediaIdx), float(dc.maxSubmediaIdx)
        dc.minSubmediaIdx = lo + nBins - 1 if lo > medium else lo
        dc.maxSubmediaIdx = hi + nBins - 1 if hi >= medium else hi
        m.MC.nPhotonsRequested = nPh * 0.5 ** (int(nIterations) - i)
        res = MCmatlab(m, 1, seed=None if seed is None else seed + i)
        if i < nIterations:
            FR = float(model.MC.P) * np.asarray(res.MC.NFR).ravel(order="F")[vox]
    return res


def t_dependent_heat(T, Omega, P, M, VHC, TC, d, medium, thermal_of_T, nBins, updates, Q_abs=None):
    """Heat simulation with a temperature-dependent medium, as
    simulateHeatDistribution.m runs it: the schedule is cut into updates
    (steps, dt, lightsOn), each one call of the heat operator continuing the
    FP64 state, and before every update the medium's voxels get (VHC, TC)
    from their current temperature through thermal_of_T(T) -> (VHC, TC)
    arrays and are binned into nBins sub-media (GPU binning with the
    callers' sortrows relabelling, getThermalMediaProperties.m:75-84); the
    media map, dTdtperdeltaT (harmonic-mean conductances, per sub-medium
    VHC) and dTdt_abs = Q_abs / VHC(voxel) are rebuilt from them
    (heatSimParamsRecalc, simulateHeatDistribution.m:515-541).  M is the
    1-based media map, VHC and TC the per-medium values of the other media,
    d = (dx, dy, dz), Q_abs the absorbed power density per voxel [W/cm^3]
    (None: no source); P supplies A, E (per medium), the boundary type and
    the sensors.  `medium` is 1-based.  Returns (T, Omega, M_split) of the
    final state (FP64 T and Omega) and the last media map (1-based)."""
    from .media import bin_medium, heat_coefficients

    M = np.asarray(M)
    nM = len(VHC)
    if not 1 <= medium <= nM:
        raise ValueError("medium out of range")
    if nM - 1 + nBins > 255:
        raise ValueError("more than 255 media after splitting")
    vox = (M == medium).ravel(order="F")
    if not vox.any():
        raise ValueError("the medium has no voxels")
    A = np.asarray(P["A"], np.float64)
    E = np.asarray(P["E"], np.float64)
    Q = None if Q_abs is None else np.asarray(Q_abs, np.float64).ravel(order="F")
    state = None
    T_cur = np.asarray(T, np.float64).ravel(order="F")
    Om_cur = Omega
    Ms = M
    for steps, dt, on in updates:
        vhc, tc = (np.broadcast_to(np.asarray(v, np.float64), (int(vox.sum()),)) for v in thermal_of_T(T_cur[vox]))
        ds, lab = bin_medium(np.vstack([vhc, tc, np.ones_like(vhc)]).astype(np.float32), nBins)
        rep = lambda a, sub: np.concatenate([np.asarray(a, np.float64)[:medium - 1], sub, np.asarray(a, np.float64)[medium:]])
        VHCs = rep(VHC, ds[0].astype(np.float64))
        TCs = rep(TC, ds[1].astype(np.float64))
        Mf = M.ravel(order="F").astype(np.int64)
        Mf = np.where(Mf > medium, Mf + nBins - 1, Mf)
        Mf[vox] = medium + lab - 1
        Ms = Mf.reshape(M.shape, order="F")
        Pu = dict(P)
        Pu["M"] = (Ms - 1).astype(np.uint8)
        Pu["A"] = rep(A, np.full(nBins, A[medium - 1]))
        Pu["E"] = rep(E, np.full(nBins, E[medium - 1]))
        Pu["dTdtperdeltaT"] = heat_coefficients(VHCs, TCs, *d)
        vhc_vox = VHCs.astype(np.float32)[Mf - 1]
        Pu["dTdt_abs"] = (np.zeros(Mf.size, np.float32) if Q is None else (Q / vhc_vox).astype(np.float32)).reshape(M.shape, order="F")
        Pu["steps"], Pu["dt"], Pu["lightsOn"] = int(steps), float(dt), bool(on)
        Tout, Om_cur, _, _ = finiteElementHeatPropagator(T, Om_cur, Pu, state_f64=state)
        state = finiteElementHeatPropagator.last_state
        T_cur = np.asarray(state[0], np.float64).ravel(order="F")
    return state[0], state[1], Ms


def heat_adjoint(T, Omega, P, segments=None, targetMask=None, checkpointEvery=0, grads=True, T_f64=None, Omega_f64=None,
                 coefGrad=False):
    """mcm_heat_adjoint_run: final damage J summed over the target voxels
    (targetMask, default: media with E != 0) after the schedule `segments`
    = [(steps, dt, lightsOn), ...] (default: P's steps, dt, lightsOn) of FP64
    DG-ADI steps, and its exact discrete gradients with respect to the
    heating rate dTdt_abs and the initial temperature (grads) and to the
    conduction coefficients dTdtperdeltaT (coefGrad; media.conductivity_gradient
    turns it into dJ/dTC).  Returns a dict J, dJ_dQ, dJ_dT0 ([nx,ny,nz]
    Fortran order, or None), dJ_dc ([nM,nM,3] or None), Omega (final FP64
    damage)."""
    inp = abi.HeatInputs(T, Omega, P, precision=abi.MCM_HEAT_FP64, scheme=abi.MCM_HEAT_DGADI, T_f64=T_f64, Omega_f64=Omega_f64)
    shape = (int(inp.struct.nx), int(inp.struct.ny), int(inp.struct.nz))
    L = int(np.prod(shape))
    segs = segments or []
    arr = (abi.mcm_heat_segment * max(len(segs), 1))()
    for k, (steps, dt, on) in enumerate(segs):
        arr[k].steps, arr[k].dt, arr[k].lightsOn, arr[k].reserved = int(steps), float(dt), int(bool(on)), 0
    mask = None
    if targetMask is not None:
        mask = np.ascontiguousarray(np.asarray(targetMask, dtype=np.uint8).reshape(shape).ravel(order="F"))
    out = abi.mcm_heat_adjoint_output()
    om = np.zeros(L)
    out.Omega_f64 = om.ctypes.data_as(C.POINTER(C.c_double))
    gq = gt = None
    if grads:
        gq, gt = np.zeros(L), np.zeros(L)
        out.dJ_dQ = gq.ctypes.data_as(C.POINTER(C.c_double))
        out.dJ_dT0 = gt.ctypes.data_as(C.POINTER(C.c_double))
    gc = None
    if coefGrad:
        nM = int(inp.struct.nM)
        gc = np.zeros(nM * nM * 3)
        out.dJ_dc = gc.ctypes.data_as(C.POINTER(C.c_double))
    check(lib().mcm_heat_adjoint_run(C.byref(inp.struct), arr, len(segs),
                                     None if mask is None else mask.ctypes.data_as(C.POINTER(C.c_uint8)),
                                     int(checkpointEvery), C.byref(out)))
    f = lambda a: None if a is None else a.reshape(shape, order="F")
    return {"J": out.J, "dJ_dQ": f(gq), "dJ_dT0": f(gt), "Omega": f(om),
            "dJ_dc": None if gc is None else gc.reshape(int(inp.struct.nM), int(inp.struct.nM), 3, order="F")}


def heat_damage_function():
    """A torch.autograd.Function over mcm_heat_adjoint_run: forward(Q, T0, TC,
    setup) returns the final damage J (a 0-d float64 tensor) of the FP64
    DG-ADI schedule; backward gives dJ/dQ (heating rate dTdt_abs [nx,ny,nz]),
    dJ/dT0 (initial temperature [nx,ny,nz]) and dJ/dTC (thermal conductivity
    per medium [nM]) times the incoming gradient.  `setup` is a dict with the
    heat MEX fields P (M, A, E, sensors, heatBoundaryType; its dTdt_abs and
    dTdtperdeltaT are replaced), VHC [nM], d = (dx, dy, dz), segments
    [(steps, dt, lightsOn), ...] and optionally targetMask and
    checkpointEvery.  Usage:
        HD = heat_damage_function()
        J = HD.apply(Q, T0, TC, setup)
        J.backward()   # Q.grad, T0.grad, TC.grad
    """
    import torch

    from .media import conductivity_gradient, heat_coefficients

    class HeatDamage(torch.autograd.Function):
        @staticmethod
        def forward(ctx, Q, T0, TC, setup):
            VHC = np.asarray(setup["VHC"], np.float64)
            dx, dy, dz = setup["d"]
            tc = TC.detach().cpu().numpy().astype(np.float64)
            P = dict(setup["P"])
            P["dTdt_abs"] = np.asfortranarray(Q.detach().cpu().numpy().astype(np.float32))
            P["dTdtperdeltaT"] = heat_coefficients(VHC, tc, dx, dy, dz)
            t0 = np.asfortranarray(T0.detach().cpu().numpy().astype(np.float64))
            om0 = np.zeros(t0.shape, np.float32)
            r = heat_adjoint(t0.astype(np.float32), om0, P, segments=setup["segments"], targetMask=setup.get("targetMask"),
                             checkpointEvery=setup.get("checkpointEvery", 0), T_f64=t0, coefGrad=True)
            gtc = conductivity_gradient(r["dJ_dc"], VHC, tc, dx, dy, dz)
            dev = Q.device
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(r["dJ_dQ"])).to(dev, Q.dtype),
                                  torch.from_numpy(np.ascontiguousarray(r["dJ_dT0"])).to(dev, T0.dtype),
                                  torch.from_numpy(gtc).to(dev, TC.dtype))
            return torch.tensor(r["J"], dtype=torch.float64, device=dev)

        @staticmethod
        def backward(ctx, gJ):
            gq, gt, gtc = ctx.saved_tensors
            return gJ.to(gq.dtype) * gq, gJ.to(gt.dtype) * gt, gJ.to(gtc.dtype) * gtc, None

    return HeatDamage


def heat_adjoints(cases, workers=4, **kw):
    """heat_adjoint of independent models, `workers` at a time: each call
    runs on its own plan and stream from its own host thread (the library
    releases the interpreter lock while it works), so small models overlap
    on the GPU.  cases = [(T, Omega, P, segments), ...]; keyword arguments
    go to every heat_adjoint call.  Returns the results in order."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        futs = [ex.submit(heat_adjoint, T, Om, P, segments=segs, **kw) for T, Om, P, segs in cases]
        return [f.result() for f in futs]


def absorbed_powers(jobs, workers=4):
    """MCPlan.absorbed_power of independent plans, `workers` at a time from
    host threads (each plan has its own stream, so the runs overlap on the
    GPU and one model's long-lived packets do not idle it).  jobs =
    [(plan, kwargs), ...]; returns the results in order."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        futs = [ex.submit(lambda pk: pk[0].absorbed_power(**pk[1]), job) for job in jobs]
        return [f.result() for f in futs]


class MCPlan:
    """Device-resident Monte Carlo plan (mcm_mc_plan_*): upload once, launch
    many times.  Used by the benchmark and by sweeps."""

    def __init__(self, model, simType=1):
        self._inp = abi.MCInputs(model, simType=simType, seed=0)
        h = C.c_void_p()
        check(lib().mcm_mc_plan_create(C.byref(self._inp.struct), C.byref(h)))
        self.h = h

    def simulate(self, nPhotons, seed=0, iL=0, firstPhoton=0, stream=None):
        check(lib().mcm_mc_plan_simulate(self.h, iL, int(nPhotons), int(seed), int(firstPhoton), stream))

    def normalize(self, iL=0):
        outs = abi.MCOutputs(self._inp)
        check(lib().mcm_mc_plan_normalize(self.h, iL, C.byref(outs.struct)))
        return outs.result()

    def tallies(self):
        p = C.c_void_p()
        n = C.c_uint64()
        check(lib().mcm_mc_plan_tallies(self.h, C.byref(p), C.byref(n)))
        return p.value, n.value

    def counters(self):
        """Event counters since the last normalize (mcm_mc_plan_counters)."""
        buf = (C.c_double * 8)()
        check(lib().mcm_mc_plan_counters(self.h, buf))
        keys = ("launched", "collected", "substeps", "scatterings", "interfaces", "runaway", "aborted", "lanes")
        return dict(zip(keys, buf[:]))

    def autotune(self, iL=0, nPackets=0):
        """mcm_mc_plan_autotune: pick the fastest number of XCD-private
        fluence copies for this model; clears the tallies.  Returns the trial
        times in ms for 1, 2, 4, 8 copies (0 = skipped)."""
        buf = (C.c_double * 4)()
        check(lib().mcm_mc_plan_autotune(self.h, int(iL), int(nPackets), buf))
        return dict(zip((1, 2, 4, 8), buf[:]))

    def set_fluence_copies(self, copies):
        """mcm_mc_plan_set_fluence_copies: 1, 2, 4 or 8 XCD-private copies."""
        check(lib().mcm_mc_plan_set_fluence_copies(self.h, int(copies)))

    def batch(self, items):
        """mcm_mc_batch_run: items = [(iL, nPhotons, seed, stream), ...] run
        in one sequence of persistent launches; returns one output dict per
        item (one wavelength each, normalised like MCmatlab)."""
        n = len(items)
        arr = (abi.mcm_mc_item * n)()
        for k, (iL, nph, seed, stream) in enumerate(items):
            arr[k].iL, arr[k].nPhotons, arr[k].seed, arr[k].stream, arr[k].reserved = int(iL), int(nph), int(seed), int(stream), 0
        outs = [abi.MCOutputs(self._inp, nL=1) for _ in range(n)]
        structs = (abi.mcm_mc_output * n)(*[o.struct for o in outs])
        check(lib().mcm_mc_batch_run(self.h, arr, n, structs))
        res = []
        for o, st in zip(outs, structs):
            o.struct = st
            res.append(o.result())
        return res

    def absorbed_power(self, box, nPhotons, seed=0, muaMap=None, iL=0, grad=True):
        """mcm_mc_grad_run: absorbed power in the voxel box ((x0, x1), (y0,
        y1), (z0, z1), 0-based, half-open) per watt, its gradient with
        respect to the per-voxel absorption map (muaMap [nx,ny,nz] 1/cm, or
        None for the media table), and d/d log(mus scale).  Returns a dict
        J, J_se, dJ_dmua (Fortran-ordered [nx,ny,nz] or None), dJ_dlogmus,
        dJ_dlogmus_se."""
        shape = self._inp.shape
        gi = abi.mcm_mc_grad_input()
        keep = None
        if muaMap is not None:
            keep = np.ascontiguousarray(np.asarray(muaMap, dtype=np.float32).reshape(shape).ravel(order="F"))
            gi.muaMap = keep.ctypes.data_as(C.POINTER(C.c_float))
        (x0, x1), (y0, y1), (z0, z1) = box
        for k, v in enumerate((x0, x1, y0, y1, z0, z1)):
            gi.box[k] = int(v)
        gi.iL, gi.nPhotons, gi.seed = int(iL), int(nPhotons), int(seed)
        go = abi.mcm_mc_grad_output()
        g = np.zeros(int(np.prod(shape)), dtype=np.float64) if grad else None
        if g is not None:
            go.dJ_dmua = g.ctypes.data_as(C.POINTER(C.c_double))
        check(lib().mcm_mc_grad_run(self.h, C.byref(gi), C.byref(go)))
        del keep
        return {"J": go.J, "J_se": go.J_se, "dJ_dmua": None if g is None else g.reshape(shape, order="F"),
                "dJ_dlogmus": go.dJ_dlogmus, "dJ_dlogmus_se": go.dJ_dlogmus_se}

    def close(self):
        if self.h:
            lib().mcm_mc_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def absorbed_power_function():
    """A torch.autograd.Function over MCPlan.absorbed_power: forward(muaMap)
    returns J (a 0-d float64 tensor), backward multiplies dJ/dmua by the
    incoming gradient.  Usage:
        AP = absorbed_power_function()
        J = AP.apply(mua_tensor, plan, box, nPhotons, seed)
        J.backward()   # mua_tensor.grad = dJ/dmua
    """
    import torch

    class AbsorbedPower(torch.autograd.Function):
        @staticmethod
        def forward(ctx, mua, plan, box, nPhotons, seed):
            r = plan.absorbed_power(box, nPhotons, seed=seed, muaMap=mua.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(r["dJ_dmua"])).to(mua.device, mua.dtype))
            return torch.tensor(r["J"], dtype=torch.float64, device=mua.device)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return gJ.to(g.dtype) * g, None, None, None, None

    return AbsorbedPower
