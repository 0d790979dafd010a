#!/usr/bin/env python3
"""Benchmark: Monte Carlo photon packets per second on the 200^3 layered-skin
workload with Fresnel interfaces and a divergent top-hat beam, on one
MI355X, through the library's device-resident plan API (the C ABI of
include/mcm_abi.h).

A step is one full run of the workload: 1e8 packets (FP32 transport, FP64
fluence) launched as kernel launches of 2^24 packets into the plan's FP64
tallies; inputs are resident in HBM before timing starts.  Prints one JSON
line (see the keys below), including:
  roofline      the transport kernel's algorithmic bytes per launch (packets
                per launch x 17 E[S] + 72 E[X], E from the oracle's event
                counts in tests/golden/roofline_counts.json) over its mean
                launch time from HIP events on the launch stream, against the
                8 TB/s HBM peak;
  cpu_baseline  the FP64 CPU restatement (OpenMP, dSFMT) timed on this host's
                cores on a bounded sample of the same workload.
  heat          the heat propagator on the same grid (2000 FP64 DG-ADI steps);
  heat_explicit the same schedule with the FP64 explicit scheme.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--packets N] [--copies R]
       python bench.py --copy-probe BYTES   (plain copy kernel for counter scaling)
       python bench.py --workload spectral64|heat200|heat200explicit|grad128|adjoint128|sweep32|binning200   (secondary lines)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "layered_tissue_200_fresnel_tophat"
HBM_PEAK_GBS = 8000.0
LAUNCH_PACKETS = 1 << 24


class Hip:
    """The few HIP runtime calls the timing needs (events and a stream)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    def stream(self):
        s = C.c_void_p()
        self._ok(self.h.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s

    def event(self):
        e = C.c_void_p()
        self._ok(self.h.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e

    def record(self, e, s):
        self._ok(self.h.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self._ok(self.h.hipEventSynchronize(b), "hipEventSynchronize")
        self._ok(self.h.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value

    def copy_probe(self, nbytes, repeats=3):
        """Device-to-device copies of nbytes (the runtime's plain copy
        kernel): the reference point that scales the HBM byte counters."""
        a, b = C.c_void_p(), C.c_void_p()
        self._ok(self.h.hipMalloc(C.byref(a), C.c_size_t(nbytes)), "hipMalloc")
        self._ok(self.h.hipMalloc(C.byref(b), C.c_size_t(nbytes)), "hipMalloc")
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self._ok(self.h.hipMemset(a, 1, C.c_size_t(nbytes)), "hipMemset")
        for _ in range(repeats):
            self._ok(self.h.hipMemcpy(b, a, C.c_size_t(nbytes), 3), "hipMemcpy")  # 3 = device to device
        self.sync()
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipFree(a)
        self.h.hipFree(b)

    def sync(self, s=None):
        if s is None:
            self._ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        else:
            self._ok(self.h.hipStreamSynchronize(s), "hipStreamSynchronize")


def measured_traffic(packets_per_launch, copies):
    """HBM bytes per transport launch from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a copy probe), or None
    when it was measured for another launch size."""
    path = os.path.join(ROOT, "profiles", "mc_transport_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None
    if (d.get("workload") != WORKLOAD or int(d.get("packets_per_launch", 0)) != packets_per_launch
            or int(d.get("fluence_copies", 1)) != copies):
        return None
    return float(d["traffic_bytes_per_launch"])


def roofline_counts():
    with open(os.path.join(ROOT, "tests", "golden", "roofline_counts.json")) as f:
        return json.load(f)["workloads"][WORKLOAD]


def host_cores():
    """Cores this process may run on (the OpenMP thread count of the CPU
    baseline, like the reference's useAllCPUs, MCmatlab.c:727-728)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle(fast=False):
    from mcmatlab_amd import abi

    lib = C.CDLL(os.path.join(ROOT, "oracle", "_build", "liboracle_fast.so" if fast else "liboracle_omp.so"))
    lib.oracle_mc_run.argtypes = [C.POINTER(abi.mcm_mc_input), C.POINTER(abi.mcm_mc_output), C.c_int, C.c_int, C.c_uint64, C.c_void_p]
    lib.oracle_heat_run.argtypes = [C.POINTER(abi.mcm_heat_input), C.POINTER(abi.mcm_heat_output), C.c_int, C.c_int]
    return lib


def _median(xs):
    xs = sorted(xs)
    k = len(xs)
    return xs[k // 2] if k % 2 else 0.5 * (xs[k // 2 - 1] + xs[k // 2])


def cpu_mc_rate(model, seconds=2.0, runs=5, threads=None, fast=False, want_absorbed=False):
    """photons/s of the FP64 CPU restatement (dSFMT streams seeded seed +
    thread) on `threads` cores: one warm-up run sizes the sample to about
    `seconds`, then the median of `runs` runs, each timed like the
    reference's simulationTime (the photon loop only)."""
    import numpy as np

    from mcmatlab_amd import abi

    lib = _oracle(fast)
    threads = threads or host_cores()

    def run(n, seed):
        model.MC.nPhotonsRequested = float(n)
        inp = abi.MCInputs(model, seed=0)
        out = abi.MCOutputs(inp)
        rc = lib.oracle_mc_run(C.byref(inp.struct), C.byref(out.struct), threads, 0, seed, None)
        if rc != 0:
            raise RuntimeError("oracle failed")
        # the reference's simulationTime (minutes): the photon loop alone,
        # not set-up or normalisation (MCmatlab.c:725, 742)
        return out, out.struct.simulationTime * 60.0

    out, dt0 = run(500 * threads, 1)
    out, dt0 = run(max(1000, int(out.struct.nPhotons / max(dt0, 1e-3) * 0.5)), 2)  # the second warm-up sizes the sample
    n = max(1000, int(out.struct.nPhotons / max(dt0, 1e-3) * seconds))
    rates, absorbed = [], []
    mua = np.asarray(model.MC.mediaProperties["mua"])[:, 0]
    for k in range(runs):
        out, dt = run(n, 1000 * (k + 2))
        rates.append(out.struct.nPhotons / dt)
        if want_absorbed:
            nfr = np.asarray(out.result()["NFR"], np.float64).reshape(model.MC.M.shape + (-1,), order="F")[..., 0]
            absorbed.append((float(np.sum(nfr * mua[model.MC.M - 1]) * model.G.dx * model.G.dy * model.G.dz), int(out.struct.nPhotons)))
    r = {"value": _median(rates), "runs": [round(x, 1) for x in rates], "packets_per_run": n, "threads": threads}
    if want_absorbed:
        r["absorbed"] = absorbed
    return r


def cpu_heat_rate(n=200, steps=(20, 8), runs=5, threads=None):
    """voxel-steps/s of the CPU restatement of the heat propagator on the
    200^3 workload: FP64 explicit (the benchmark's scheme) and the
    reference's own FP32 Douglas-Gunn ADI, median of `runs` runs."""
    from mcmatlab_amd import abi
    from mcmatlab_amd.workloads import layered_tissue_heat

    lib = _oracle()
    threads = threads or host_cores()
    T0, Om0, Pon, _ = layered_tissue_heat(n=n, steps_on=1, steps_off=1)
    out = {}
    for name, variant, k in (("fp64_explicit", 2, steps[0]), ("fp32_dgadi", 0, steps[1])):
        P = dict(Pon, steps=k)
        inp = abi.HeatInputs(T0, Om0, P)
        rates = []
        for _ in range(runs + 1):
            o = abi.HeatOutputs(inp, want_f64=False)
            t = time.perf_counter()
            if lib.oracle_heat_run(C.byref(inp.struct), C.byref(o.struct), variant, threads) != 0:
                raise RuntimeError("heat oracle failed")
            rates.append(n ** 3 * k / (time.perf_counter() - t))
        out[name] = {"voxel_steps_per_s": _median(rates[1:]), "steps_per_run": k}
    return out


def cpu_baseline(model):
    """The reference algorithm's CPU path as a baseline beside the GPU: the
    FP64 restatement (dSFMT, OpenMP over every core this process may use) on
    a bounded sample of the same workload, -O3 and the shipped -Ofast flags
    (MCmatlab.c:60); the larger rate is `value`.  Also one thread, and the
    heat restatement on the same grid."""
    o3 = cpu_mc_rate(model, want_absorbed=True)
    fast = cpu_mc_rate(model, fast=True)
    one = cpu_mc_rate(model, seconds=1.5, runs=3, threads=1)
    best = max(o3["value"], fast["value"])
    return {"value": best, "unit": "photons/s", "cores": o3["threads"], "kind": "port", "cpu": cpu_model(),
            "sample": f"{o3['packets_per_run']} packets of the {WORKLOAD} workload per run, median of 5 runs after a warm-up; "
                      "FP64 transport, dSFMT, OpenMP",
            "o3": o3["value"], "ofast": fast["value"], "runs_o3": o3["runs"], "runs_ofast": fast["runs"],
            "value_1thread": one["value"], "heat": cpu_heat_rate(), "_absorbed": o3["absorbed"]}


def bench_heat(hip, steps_on=667, steps_off=1333, scheme="dgadi"):
    """Config-2 heat schedule on the same 200^3 grid through the heat plan
    (FP64; scheme "dgadi", the reference's Douglas-Gunn ADI, or "explicit",
    forward Euler, stable here at dt = 0.18 dtmax)."""
    from mcmatlab_amd import abi
    from mcmatlab_amd._lib import check, lib
    from mcmatlab_amd.workloads import layered_tissue_heat

    T0, Om0, Pon, Poff = layered_tissue_heat(n=200, steps_on=steps_on, steps_off=steps_off)
    inp = abi.HeatInputs(T0, Om0, Pon, scheme=abi.MCM_HEAT_EXPLICIT if scheme == "explicit" else abi.MCM_HEAT_DGADI)
    h = C.c_void_p()
    check(lib().mcm_heat_plan_create(C.byref(inp.struct), C.byref(h)))
    try:
        # warm-up: 10 steps
        check(lib().mcm_heat_plan_advance(h, 10, Pon["dt"], 1))
        hip.sync()
        t = time.perf_counter()
        check(lib().mcm_heat_plan_advance(h, steps_on, Pon["dt"], 1))
        check(lib().mcm_heat_plan_advance(h, steps_off, Poff["dt"], 0))
        hip.sync()
        dt = time.perf_counter() - t
    finally:
        lib().mcm_heat_plan_destroy(h)
    L = 200 ** 3
    gbs = L * (steps_on * 41 + steps_off * 33) / dt / 1e9
    return {"voxel_steps_per_s": L * (steps_on + steps_off) / dt, "steps": steps_on + steps_off, "seconds": dt,
            "scheme": "explicit" if scheme == "explicit" else "DG-ADI", "dtype": "f64", "algorithmic_GBs": gbs,
            "frac": gbs / HBM_PEAK_GBS}


def bench_spectral(seed=7):
    """The 64-wavelength spectral workload (150^3, matched, pencil beam,
    per-wavelength budgets round(10^U(5,7))): one batched persistent launch
    (mcm_mc_batch_run) against 64 plain launches on the same plan and
    against 64 whole single-wavelength calls (mcm_mc_run each).  Wall clock
    includes normalisation and the download of every wavelength's outputs."""
    import numpy as np

    from mcmatlab_amd import api
    from mcmatlab_amd.workloads import layered_tissue_spectral_mc

    model, budgets = layered_tissue_spectral_mc(n=150, nL=64)
    total = sum(budgets)
    plan = api.MCPlan(model)
    items = [(iL, n, seed, iL) for iL, n in enumerate(budgets)]
    plan.batch([(0, 100000, seed + 1, 0)])  # warm-up
    t = time.perf_counter()
    got = plan.batch(items)
    t_batch = time.perf_counter() - t
    t = time.perf_counter()
    for iL, n in enumerate(budgets):
        plan.simulate(n, seed=seed, iL=iL, firstPhoton=0)
        plan.normalize(iL)
    t_plain = time.perf_counter() - t
    plan.close()
    t = time.perf_counter()
    for iL, n in enumerate(budgets):
        m1 = model.copy()
        mp = model.MC.mediaProperties
        m1.MC.mediaProperties = {k: np.asarray(v)[:, iL:iL + 1] for k, v in mp.items()}
        m1.MC.spectrum = np.ones(1)
        m1.MC.nPhotonsRequested = float(n)
        api.MCmatlab(m1, 1, seed=seed)
    t_calls = time.perf_counter() - t
    absorbed = [float(np.sum(np.asarray(g["NFR"], np.float64)[..., 0] * np.asarray(model.MC.mediaProperties["mua"])[:, iL][model.MC.M - 1])
                      * model.G.dx ** 3) for iL, g in enumerate(got)]
    return {"metric": "photons/s, 64-wavelength batched launch (150^3 layered skin, pencil beam)", "value": total / t_batch,
            "unit": "photons/s", "packets": total, "batched_s": t_batch, "plain_launches_s": t_plain, "plain_calls_s": t_calls,
            "speedup_vs_plain_launches": t_plain / t_batch, "speedup_vs_plain_calls": t_calls / t_batch,
            "absorbed_fraction_first_last": [absorbed[0], absorbed[-1]]}


def bench_config1(packets=100_000_000):
    """The CPU-runnable workload: 100^3 matched layered skin, pencil beam.
    GPU photons/s (autotuned fluence copies, inputs resident) and the CPU
    restatement on every core (median of 5 runs)."""
    from mcmatlab_amd.api import MCPlan
    from mcmatlab_amd.workloads import layered_tissue_mc

    model = layered_tissue_mc(n=100, matched=True, beam="pencil", nPhotons=packets)
    plan = MCPlan(model)
    tune = plan.autotune()
    best = min(t for t in tune.values() if t > 0)
    copies = min(R for R, t in tune.items() if 0 < t <= 1.10 * best)
    plan.set_fluence_copies(copies)
    hip = Hip()
    plan.simulate(1 << 20, seed=1)
    plan.normalize()
    hip.sync()
    t = time.perf_counter()
    done = 0
    while done < packets:
        n = min(LAUNCH_PACKETS, packets - done)
        plan.simulate(n, seed=2, firstPhoton=done)
        done += n
    hip.sync()
    dt = time.perf_counter() - t
    plan.close()
    cpu = cpu_mc_rate(layered_tissue_mc(n=100, matched=True, beam="pencil", nPhotons=1000), fast=True)
    return {"workload": "layered_tissue_100_matched_pencil", "photons_per_s": packets / dt, "fluence_copies": copies,
            "cpu_photons_per_s": cpu["value"], "cpu_cores": cpu["threads"], "cpu_flags": "-Ofast", "vs_cpu": packets / dt / cpu["value"]}


def bench_heat_line(scheme="dgadi"):
    """The heat propagator alone: 667 lights-on + 1333 lights-off FP64 steps
    on the 200^3 grid, against the compulsory traffic of 41 / 33 bytes per
    voxel-step."""
    r = bench_heat(Hip(), scheme=scheme)
    name = "explicit" if scheme == "explicit" else "DG-ADI"
    return {"metric": f"voxel-steps/s, 200^3 heat (FP64 {name}, 667 on + 1333 off)", "value": r["voxel_steps_per_s"],
            "unit": "voxel-steps/s", "ms_per_step": r["seconds"] / r["steps"] * 1e3, "dtype": "f64",
            "roofline": {"bound": "hbm", "achieved": r["algorithmic_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": r["frac"]}}


def bench_grad(n=128, packets=10_000_000, seed=3):
    """Absorbed power in the 16^3 box of the 128^3 blob absorber and its
    per-voxel mua gradient (two passes per packet)."""
    from mcmatlab_amd import api
    from mcmatlab_amd.workloads import blob_absorber_mc

    m, mua, box = blob_absorber_mc(n=n)
    plan = api.MCPlan(m)
    plan.absorbed_power(box, 100000, seed=seed + 1, muaMap=mua)  # warm-up
    t = time.perf_counter()
    r = plan.absorbed_power(box, packets, seed=seed, muaMap=mua)
    dt = time.perf_counter() - t
    # the forward alone: J (and the one-pass likelihood-ratio dJ/dlog mus)
    t = time.perf_counter()
    f = plan.absorbed_power(box, packets, seed=seed, muaMap=mua, grad=False)
    dt_fwd = time.perf_counter() - t
    # the production transport (fluence and escapes) on the same grid and
    # medium; it reads the media table (mua 1/cm), not the per-voxel map
    plan.simulate(100000, seed=seed + 2)
    plan.normalize()
    t = time.perf_counter()
    plan.simulate(packets, seed=seed)
    plan.normalize()
    dt_tr = time.perf_counter() - t
    plan.close()
    return {"metric": "packets/s, absorbed power + dJ/dmua (128^3 blob absorber, 16^3 box)", "value": packets / dt,
            "unit": "packets/s", "seconds": dt, "forward_s": dt_fwd, "grad_over_forward": dt / dt_fwd,
            "transport_forward_s": dt_tr, "grad_over_transport_forward": dt / dt_tr, "J": r["J"],
            "J_forward": f["J"], "J_rel_se": r["J_se"] / r["J"], "dJ_dlogmus": r["dJ_dlogmus"],
            "dJ_dlogmus_rel_se": abs(r["dJ_dlogmus_se"] / r["dJ_dlogmus"])}


def bench_adjoint(n=128, steps=3000, every=100):
    """Final blood damage and its gradient w.r.t. the heating rate: 3000 FP64
    DG-ADI steps on 128^3 (heat parameters of the 200^3 workload),
    checkpoints every 100 steps."""
    from mcmatlab_amd import api
    from mcmatlab_amd.workloads import layered_tissue_heat

    T0, Om0, Pon, _ = layered_tissue_heat(n=n, steps_on=steps, steps_off=1)
    t = time.perf_counter()
    r = api.heat_adjoint(T0, Om0, Pon, segments=[(steps, 5e-3 / 667, True)], checkpointEvery=every, grads=False)
    t_fwd = time.perf_counter() - t
    t = time.perf_counter()
    r = api.heat_adjoint(T0, Om0, Pon, segments=[(steps, 5e-3 / 667, True)], checkpointEvery=every)
    dt = time.perf_counter() - t
    return {"metric": "seconds, 3000-step heat adjoint (128^3, FP64 DG-ADI, checkpoints every 100)", "value": dt, "unit": "s",
            "higher_is_better": False, "forward_only_s": t_fwd, "adjoint_over_forward": dt / t_fwd, "J": r["J"],
            "dJ_dQ_max": float(abs(r["dJ_dQ"]).max())}


def bench_binning(n=200, bins=150, sample_div=16):
    """Media binning (getDownsampledParamVals) of a temperature-dependent
    medium filling the 200^3 grid into 150 sub-media (Example14's nBins):
    one call, host input to host output (host sums, upload, 151 passes over
    the 8e6 sets, download).  CPU: the oracle (the reference's sequential
    algorithm, one thread) on every 16th set with the same number of bins."""
    import numpy as np

    from mcmatlab_amd import api
    from mcmatlab_amd.workloads import thermochromic_params

    pv = thermochromic_params(n=n)
    api.getDownsampledParamVals(pv[:, :4096], 4)  # device initialisation
    t = time.perf_counter()
    out, bins_ = api.getDownsampledParamVals(pv, bins)
    dt = time.perf_counter() - t
    n1 = pv.shape[1]
    res = {"metric": "set-passes/s, media binning of 200^3 voxel parameter sets into 150 bins", "value": n1 * bins / dt,
           "unit": "set-passes/s", "seconds": dt, "n1": n1, "n2": bins, "dtype": "f32",
           "labels_used": int(len(np.unique(bins_)))}
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    try:
        import oracle_lib as OL

        sample = np.ascontiguousarray(pv[:, ::sample_div])
        t = time.perf_counter()
        OL.oracle_binning(sample, bins)
        dc = time.perf_counter() - t
        res["cpu_baseline"] = {"value": sample.shape[1] * bins / dc, "unit": "set-passes/s", "cores": 1, "kind": "port",
                               "sample": f"every {sample_div}th set ({sample.shape[1]}), {bins} bins ({dc:.1f} s)"}
    except OSError:
        res["cpu_baseline"] = None
    return res


def bench_sweep(n_models=32, seed=3, heat_steps=1000, every=100, parts=("mc", "grad", "heat", "adjoint")):
    """Batched forward + backward sweep over independent models on uneven
    grids (multi_grid_sweep: sides 64..247 from seed 3, cyclic x/y, 1e6-1e7
    packets each): forward MC of every model as one multi-grid launch
    sequence (and, for comparison, as plain launches one model after
    another), the absorbed-power gradient of each model (four at a time from
    host threads, and one after another), 1000 heat steps of every model as
    one plan group (and one plan after another), and each model's heat
    adjoint (four at a time).  `parts` selects stages (the heat and adjoint
    stages need "mc" for their heating rates).  Stages print progress to
    stderr."""
    from mcmatlab_amd import api
    from mcmatlab_amd.workloads import multi_grid_sweep

    def log(msg):
        print(f"[sweep] {msg}", file=sys.stderr, flush=True)

    t = time.perf_counter()
    W = multi_grid_sweep(n_models=n_models, seed=seed)
    log(f"models built in {time.perf_counter() - t:.1f}s")
    plans = [api.MCPlan(w["model"]) for w in W]
    hip = Hip()
    packets = sum(w["nPhotons"] for w in W)
    voxels = sum(w["side"] ** 3 for w in W)
    api.multi_run(plans, [(k, 0, 20000, 7, 0) for k in range(len(W))])  # warm-up
    items = [(k, 0, w["nPhotons"], 11 + k, 0) for k, w in enumerate(W)]
    hip.sync()
    t = time.perf_counter()
    outs = api.multi_run(plans, items)
    t_multi = time.perf_counter() - t
    log(f"multi-grid MC {t_multi:.2f}s")
    t = time.perf_counter()
    for p, (k, iL, n, sd, st) in zip(plans, items):
        p.simulate(n, seed=sd, iL=iL)
        p.normalize(iL)
    t_plain = time.perf_counter() - t
    log(f"plain MC {t_plain:.2f}s")
    res = {"metric": f"seconds, {n_models}-grid sweep (sides 64..256): " + " + ".join(parts), "unit": "s",
           "higher_is_better": False, "models": n_models, "voxels": voxels, "packets": packets, "mc_multi_s": t_multi,
           "mc_plain_s": t_plain, "mc_multi_photons_per_s": packets / t_multi}
    total = t_multi
    if "grad" in parts:
        t = time.perf_counter()
        for p, w in zip(plans, W):
            p.absorbed_power(w["box"], w["nPhotons"], seed=5)
        t_grad1 = time.perf_counter() - t
        log(f"gradients one after another {t_grad1:.2f}s")
        t = time.perf_counter()
        api.absorbed_powers([(p, {"box": w["box"], "nPhotons": w["nPhotons"], "seed": 5}) for p, w in zip(plans, W)], workers=4)
        t_grad = time.perf_counter() - t
        log(f"gradients (4 at a time) {t_grad:.2f}s")
        res["grad_serial_s"] = t_grad1
        res.update(grad_s=t_grad, grad_packets_per_s=packets / t_grad)
        total += t_grad
    for p in plans:
        p.close()
    if "heat" not in parts and "adjoint" not in parts:
        res["value"] = total
        return res
    cases = [w["heat"](o["NFR"][..., 0]) for w, o in zip(W, outs)]
    if "heat" in parts:
        hplans = [api.HeatPlan(T0, Om0, P) for T0, Om0, P in cases]
        dts = [P["dt"] for _, _, P in cases]
        hip.sync()
        t = time.perf_counter()
        with api.HeatGroup(hplans, steps_per_graph=32) as g:
            g.advance(heat_steps, dts, [1] * len(hplans))
            hip.sync()
        t_group = time.perf_counter() - t
        log(f"heat group {t_group:.2f}s")
        t = time.perf_counter()
        for h, d in zip(hplans, dts):
            h.advance(heat_steps, d, 1)
            hip.sync()
        t_serial = time.perf_counter() - t
        log(f"heat one plan after another {t_serial:.2f}s")
        for h in hplans:
            h.close()
        res.update(heat_group_s=t_group, heat_serial_s=t_serial, heat_group_voxel_steps_per_s=voxels * heat_steps / t_group)
        total += t_group
    if "adjoint" in parts:
        t = time.perf_counter()
        rs = api.heat_adjoints([(T0, Om0, P, [(heat_steps, P["dt"], True)]) for T0, Om0, P in cases], workers=4,
                               checkpointEvery=every)
        t_adj = time.perf_counter() - t
        log(f"adjoints (4 at a time) {t_adj:.2f}s")
        res.update(adjoint_s=t_adj, damage_sum=sum(r["J"] for r in rs))
        total += t_adj
    res["value"] = total
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=float, default=1e8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-heat", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched-vs-plain, gradient-cost and 100^3 lines")
    ap.add_argument("--workload", default="layered_tissue_200_fresnel_tophat",
                    choices=["layered_tissue_200_fresnel_tophat", "spectral64", "heat200", "heat200explicit", "grad128", "adjoint128",
                             "sweep32", "binning200"],
                    help="secondary workloads (each prints its own line): spectral64 batched vs plain launches of the "
                         "64-wavelength workload; heat200 (heat200explicit) the DG-ADI (explicit) heat propagator alone; "
                         "grad128 absorbed-power gradients; "
                         "adjoint128 the heat adjoint; sweep32 the 32-grid forward + backward sweep; "
                         "binning200 the media binning of a 200^3 temperature-dependent medium")
    ap.add_argument("--copies", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="force the number of XCD-private fluence copies (0: the autotuner's choice)")
    ap.add_argument("--sweep-parts", default="mc,grad,heat,adjoint",
                    help="stages of the sweep32 workload (comma-separated subset of mc,grad,heat,adjoint)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    secondary = {"spectral64": bench_spectral, "heat200": bench_heat_line, "heat200explicit": lambda: bench_heat_line("explicit"),
                 "grad128": bench_grad, "adjoint128": bench_adjoint, "sweep32": bench_sweep, "binning200": bench_binning}
    if args.workload in secondary:
        kw = {"parts": tuple(args.sweep_parts.split(","))} if args.workload == "sweep32" else {}
        print(json.dumps(secondary[args.workload](**kw)), flush=True)
        return
    if args.copy_probe:
        Hip().copy_probe(args.copy_probe)
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "repeats": 3}), flush=True)
        return

    from mcmatlab_amd.api import MCPlan
    from mcmatlab_amd.workloads import layered_tissue_mc

    t_setup = time.perf_counter()
    model = layered_tissue_mc(n=200, matched=False, beam="tophat", nPhotons=args.packets)
    hip = Hip()
    stream = hip.stream()
    plan = MCPlan(model)
    tune_ms = plan.autotune()  # number of XCD-private fluence copies for this model
    best = min(t for t in tune_ms.values() if t > 0)
    copies = min(R for R, t in tune_ms.items() if 0 < t <= 1.10 * best)
    if args.copies:
        copies = args.copies
        plan.set_fluence_copies(copies)
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s, fluence copies {copies} (trial ms {tune_ms})", file=sys.stderr,
          flush=True)
    npk = int(args.packets)
    launches = []
    first = 0

    def step(record):
        nonlocal first
        done = 0
        while done < npk:
            n = min(LAUNCH_PACKETS, npk - done)
            if record:
                e0, e1 = hip.event(), hip.event()
                hip.record(e0, stream)
            plan.simulate(n, seed=12345, firstPhoton=first, stream=stream)
            if record:
                hip.record(e1, stream)
                launches.append((e0, e1, n))
            first += n
            done += n

    for _ in range(args.warmup):
        step(False)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    hip.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = args.steps * npk / elapsed

    durs = [(hip.elapsed_ms(a, b), n) for a, b, n in launches]
    full = [d for d, n in durs if n == LAUNCH_PACKETS] or [d for d, _ in durs]
    n_per = LAUNCH_PACKETS if any(n == LAUNCH_PACKETS for _, n in durs) else durs[0][1]
    avg_ms = sum(full) / len(full)
    rc = roofline_counts()
    alg_bytes = rc["bytes_per_packet"] * n_per
    achieved = alg_bytes / (avg_ms / 1e3) / 1e9
    # sanity of the physics on the timed tallies
    res = plan.normalize()
    import numpy as np

    mua = np.asarray(model.MC.mediaProperties["mua"])[:, 0]
    absorbed = float(np.sum(np.asarray(res["NFR"], dtype=np.float64)[..., 0] * mua[model.MC.M - 1]) * model.G.dx ** 3)
    plan.close()

    line = {
        "metric": "photons/s at 200^3 tissue (Fresnel, top-hat beam), FP32 transport with FP64 fluence",
        "value": value, "unit": "photons/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOAD, "grid": [200, 200, 200], "packets_per_step": npk, "fresnel": True,
                   "beam": "top-hat r=0.03 cm, Gaussian 5 deg divergence", "boundaryType": 1, "fluence_copies": copies,
                   "absorbed_fraction": absorbed},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n_per, copies),
                     "kernel": "mc_transport", "launch_ms": avg_ms, "packets_per_launch": n_per,
                     "bytes_per_packet": rc["bytes_per_packet"]},
    }
    if not args.no_heat:
        try:
            line["heat"] = bench_heat(hip)
            line["heat_explicit"] = bench_heat(hip, scheme="explicit")
        except Exception as e:  # heat kernels not available in this build
            line["heat"] = {"error": str(e)[:200]}
    if not args.no_extra:
        # the metric's other two parts: batched vs plain launches, and the
        # cost of a forward + gradient against a forward
        sp = bench_spectral()
        line["batched_vs_plain"] = {"workload": "layered_tissue_150_spectral64", "packets": sp["packets"], "batched_s": sp["batched_s"],
                                    "plain_launches_s": sp["plain_launches_s"], "plain_calls_s": sp["plain_calls_s"],
                                    "speedup_vs_plain_launches": sp["speedup_vs_plain_launches"],
                                    "speedup_vs_plain_calls": sp["speedup_vs_plain_calls"]}
        gr = bench_grad()
        line["grad_over_forward"] = {"workload": "blob_absorber_128", "packets": 10_000_000, "forward_plus_grad_s": gr["seconds"],
                                     "forward_s": gr["forward_s"], "ratio": gr["grad_over_forward"],
                                     "transport_forward_s": gr["transport_forward_s"],
                                     "ratio_vs_transport": gr["grad_over_transport_forward"]}
    if not args.no_cpu:
        cb = cpu_baseline(layered_tissue_mc(n=200, matched=False, beam="tophat", nPhotons=1000))
        # the GPU's absorbed fraction against the oracle's on the same
        # workload: per packet the absorbed fraction lies in [0, 1], so its
        # variance is at most p (1 - p)
        ab = cb.pop("_absorbed")
        n_cpu = sum(n for _, n in ab)
        p_cpu = sum(a * n for a, n in ab) / n_cpu
        se = math.sqrt(max(p_cpu * (1 - p_cpu), 1e-12) / n_cpu)
        line["config"]["absorbed_fraction_oracle"] = p_cpu
        line["config"]["absorbed_fraction_z"] = (absorbed - p_cpu) / se
        line["cpu_baseline"] = cb
        line["vs_cpu"] = value / cb["value"]
        if not args.no_extra:
            line["config1"] = bench_config1()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
