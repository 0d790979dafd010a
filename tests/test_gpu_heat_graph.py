"""Heat steps replayed from captured HIP graphs (mcm_heat_plan_set_graph_steps)
against the same steps launched one by one.

Both run the same kernels with the same arguments in the same order, so
every state, sensor record and per-medium maximum is compared bit for bit,
after every run of the schedule (a download synchronises between replays).
The schedule mixes runs long enough for graph replays with short runs of
step launches, odd run lengths (the state buffers swap parity, so a cached
graph of the other parity must not be reused) and a repeated setting (the
cached graph is replayed again).  The step launches themselves are pinned
to the oracle in test_gpu_heat.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from mcmatlab_amd.workloads import layered_tissue_heat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def run(api, T0, Om0, P, schedule, precision, scheme, graph_steps):
    out = []
    with api.HeatPlan(T0, Om0, P, precision=precision, scheme=scheme, graph_steps=graph_steps) as h:
        for steps, dt, on in schedule:
            h.advance(steps, dt, on)
            out.append(h.download(f64=True))
    return out


@pytest.mark.parametrize("precision,scheme,bt", [
    ("fp64", "dgadi", 1), ("fp64", "dgadi", 0), ("fp32", "dgadi", 1), ("fp64", "explicit", 1), ("fp32", "explicit", 0),
])
def test_graph_replays_equal_step_launches(api, precision, scheme, bt):
    print(f"[gpu-heat-graph] {precision} {scheme} bt={bt}", flush=True)
    T0, Om0, Pon, Poff = layered_tissue_heat(n=40, steps_on=100, steps_off=100)
    P = dict(Pon, heatBoundaryType=bt)
    don, doff = Pon["dt"], Poff["dt"]
    schedule = [(70, don, 1), (9, don, 1), (66, doff, 0), (70, don, 1), (64, doff, 0), (33, don, 1)]
    eager = run(api, T0, Om0, P, schedule, precision, scheme, 0)
    graph = run(api, T0, Om0, P, schedule, precision, scheme, 16)
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert a["stepsDone"] == b["stepsDone"] == sum(s for s, _, _ in schedule[: k + 1])
        for key in ("T_f64", "Omega_f64", "sensorTemps", "maxMediaTemps"):
            assert np.array_equal(a[key], b[key]), (k, key)
    assert eager[-1]["sensorTemps"].shape == (3, 313)
    assert np.all(np.isfinite(eager[-1]["sensorTemps"]))


def run_no_sync(api, T0, Om0, P, schedule, graph_steps):
    with api.HeatPlan(T0, Om0, P, graph_steps=graph_steps) as h:
        for steps, dt, on in schedule:
            h.advance(steps, dt, on)
        return h.download(f64=True)


def test_graph_cache_eviction_with_replays_in_flight(api):
    """Six settings in a row with no download in between: the plan keeps four
    captured graphs, so the fifth and sixth captures evict graphs whose
    replays may still be queued on the stream; the eviction waits for them.
    The result equals the step launches bit for bit."""
    T0, Om0, Pon, _ = layered_tissue_heat(n=40, steps_on=100, steps_off=100)
    dt0 = Pon["dt"]
    schedule = [(40, dt0 * (1 + 0.1 * k), k % 2) for k in range(6)] + [(40, dt0, 1)]
    a = run_no_sync(api, T0, Om0, Pon, schedule, 0)
    b = run_no_sync(api, T0, Om0, Pon, schedule, 16)
    for key in ("T_f64", "Omega_f64", "sensorTemps", "maxMediaTemps"):
        assert np.array_equal(a[key], b[key]), key


def test_graph_steps_validation(api):
    T0, Om0, Pon, _ = layered_tissue_heat(n=16, steps_on=10, steps_off=10)
    with api.HeatPlan(T0, Om0, Pon) as h:
        with pytest.raises(RuntimeError):
            h.set_graph_steps(-2)
        h.set_graph_steps(7)  # odd: rounded down to 6 steps per graph
        h.advance(20, Pon["dt"], 1)
        r = h.download()
    assert r["stepsDone"] == 20


def grid_case(shape, bt, seed):
    """The layered tissue heat model cut to `shape` with a random heating rate."""
    n = max(shape)
    T0, Om0, Pon, Poff = layered_tissue_heat(n=n, steps_on=100, steps_off=100)
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    P = dict(Pon, heatBoundaryType=bt)
    P["M"] = np.asarray(Pon["M"])[:nx, :ny, :nz].copy(order="F")
    P["dTdt_abs"] = (2e4 * rng.random(shape)).astype(np.float32)
    P["tempSensorCornerIdxs"] = np.zeros(0)
    P["tempSensorInterpWeights"] = np.zeros((0, 3))
    T = (37 + 5 * rng.random(shape)).astype(np.float32)
    return T, np.zeros(shape, np.float32), P, Pon["dt"], Poff["dt"]


def test_group_equals_plans_advanced_alone(api):
    """Three grids (one with sensors, one FP32, one with fixed-temperature
    boundaries) advanced as a group: every plan's state after every run
    equals the same plan advanced alone with step launches."""
    cases = [grid_case((24, 24, 24), 1, 1), grid_case((40, 32, 28), 0, 2), grid_case((48, 20, 36), 1, 3)]
    T0, Om0, Pon, Poff = layered_tissue_heat(n=32, steps_on=100, steps_off=100)
    cases.append((T0, Om0, dict(Pon), Pon["dt"], Poff["dt"]))  # with sensors
    precisions = ["fp64", "fp64", "fp32", "fp64"]
    runs = [(70, 1), (9, 1), (40, 0), (64, 1)]

    def alone(i):
        T, Om, P, don, doff = cases[i]
        out = []
        with api.HeatPlan(T, Om, P, precision=precisions[i], graph_steps=0) as h:
            for steps, on in runs:
                h.advance(steps, don if on else doff, on)
                out.append(h.download())
        return out

    ref = [alone(i) for i in range(len(cases))]
    plans = [api.HeatPlan(T, Om, P, precision=pr) for (T, Om, P, _, _), pr in zip(cases, precisions)]
    try:
        with api.HeatGroup(plans, steps_per_graph=16) as g:
            for k, (steps, on) in enumerate(runs):
                g.advance(steps, [c[3] if on else c[4] for c in cases], [on] * len(cases))
                for i, p in enumerate(plans):
                    got = p.download()
                    assert got["stepsDone"] == ref[i][k]["stepsDone"]
                    for key in ("T_f64", "Omega_f64", "sensorTemps", "maxMediaTemps"):
                        assert np.array_equal(got[key], ref[i][k][key]), (i, k, key)
    finally:
        for p in plans:
            p.close()
    assert ref[3][-1]["sensorTemps"].shape == (3, 184)


def test_group_validation(api):
    T0, Om0, Pon, _ = layered_tissue_heat(n=16, steps_on=10, steps_off=10)
    with api.HeatPlan(T0, Om0, Pon) as h:
        with pytest.raises(RuntimeError):
            api.HeatGroup([h, h])
        with api.HeatGroup([h]) as g:
            with pytest.raises(RuntimeError):
                g.advance(-1, [Pon["dt"]], [1])
            g.advance(20, [Pon["dt"]], [1])
        assert h.download()["stepsDone"] == 20
