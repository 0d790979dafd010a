"""Statistical comparison of two independent Monte Carlo implementations
run on the same model with different random streams.

A run's scalar summaries are the quantities the reference prints or derives:
  absorbed fraction  = sum(NFR * mua[M]) * dV            (plotMCmatlabMC.m:145)
  escape per face    = sum(NI_face) * dA                 (README.md:616-617)
  collected          = sum(LC.image) * pixel area (lens) or sum(image) (fiber)
  far-field total    = sum(farField)
Each side is run with S seeds; the difference of the means is compared with
its standard error from the seed-to-seed scatter (Welch).  The per-voxel
NFR is compared the same way voxel by voxel.
"""
from __future__ import annotations

import math

import numpy as np

# two-sided tail of 3 sigma for a normal variable
P_3SIGMA = 0.0026997960632601866


def t_critical(dof, p=P_3SIGMA):
    """Critical |t| of Student's t with `dof` degrees of freedom at two-sided
    tail probability p (bisection on the regularized incomplete beta)."""
    from scipy.stats import t as tdist

    return float(tdist.ppf(1 - p / 2, dof))


def summaries(res, model, simType=1, iL=0):
    mc = model.FMC if simType == 2 else model.MC
    mua = np.asarray(mc.mediaProperties["mua"], dtype=np.float64)
    if mua.ndim == 1:
        mua = mua[:, None]
    M = np.asarray(mc.M)
    dx, dy, dz = model.G.dx, model.G.dy, model.G.dz
    out = {}
    if "NFR" in res:
        nfr = np.asarray(res["NFR"], dtype=np.float64).reshape(M.shape + (-1,), order="F")[..., iL]
        out["absorbed"] = float(np.sum(nfr * mua[M - 1, iL]) * dx * dy * dz)
    areas = {"NI_xpos": dy * dz, "NI_xneg": dy * dz, "NI_ypos": dx * dz, "NI_yneg": dx * dz, "NI_zpos": dx * dy, "NI_zneg": dx * dy}
    for k, a in areas.items():
        if k in res and np.asarray(res[k]).size > 1:
            arr = np.asarray(res[k], dtype=np.float64)
            arr = arr.reshape(arr.shape[0], arr.shape[1], -1, order="F")[..., iL]
            out[k] = float(np.sum(arr) * a)
    if "image" in res and np.asarray(res["image"]).size >= 1 and mc.useLightCollector:
        img = np.asarray(res["image"], dtype=np.float64)
        out["collected"] = float(np.sum(img))
    if "farField" in res and np.asarray(res["farField"]).size > 1:
        out["farField"] = float(np.sum(res["farField"]))
    return out


def welch(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    va, vb = a.var(ddof=1) / len(a), b.var(ddof=1) / len(b)
    se = math.sqrt(va + vb)
    diff = a.mean() - b.mean()
    if se == 0:
        return 0.0 if diff == 0 else math.inf, len(a) + len(b) - 2
    dof = (va + vb) ** 2 / ((va ** 2 / (len(a) - 1) if va else 0) + (vb ** 2 / (len(b) - 1) if vb else 0) or 1e-300)
    return diff / se, dof


def compare_scalars(runs_a, runs_b, model, simType=1, iL=0, rel_floor=1e-6):
    """[(name, t, t_crit, mean_a, mean_b)] for every scalar summary."""
    sa = [summaries(r, model, simType, iL) for r in runs_a]
    sb = [summaries(r, model, simType, iL) for r in runs_b]
    rows = []
    for k in sa[0]:
        a = [s[k] for s in sa]
        b = [s[k] for s in sb]
        if max(abs(np.mean(a)), abs(np.mean(b))) < rel_floor:
            continue
        t, dof = welch(a, b)
        rows.append((k, t, t_critical(max(dof, 2)), float(np.mean(a)), float(np.mean(b))))
    return rows


def compare_voxels(runs_a, runs_b, key="NFR", min_nonzero=None):
    """(fraction of voxels with |t| > 3, mean t^2, voxels tested) for the
    per-voxel field `key`, over voxels that every run of both sides hit."""
    A = np.stack([np.asarray(r[key], dtype=np.float64).ravel() for r in runs_a])
    B = np.stack([np.asarray(r[key], dtype=np.float64).ravel() for r in runs_b])
    need = min_nonzero if min_nonzero is not None else A.shape[0]
    ok = (np.count_nonzero(A, axis=0) >= need) & (np.count_nonzero(B, axis=0) >= B.shape[0] if min_nonzero is None else
                                                 np.count_nonzero(B, axis=0) >= need)
    se = np.sqrt(A.var(0, ddof=1) / A.shape[0] + B.var(0, ddof=1) / B.shape[0])
    ok &= se > 0
    t = (A.mean(0) - B.mean(0))[ok] / se[ok]
    if t.size == 0:
        return 0.0, 0.0, 0
    return float(np.mean(np.abs(t) > 3)), float(np.mean(t ** 2)), int(t.size)


def check_scalars_two_stage(run_a, run_b, model, seeds=8, extra=24, simType=1, iL=0, runs=None):
    """3-sigma check of every scalar summary with a second stage: a summary
    outside the bound after `seeds` runs per side is re-tested on
    `seeds + extra` runs per side, and only fails if it is still outside.
    Keeps the false-alarm rate of many simultaneous 3-sigma comparisons low
    while a real bias (which does not shrink with more runs) still fails.
    Returns the list of failures [(name, t, t_crit, mean_a, mean_b)]."""
    A = runs[0] if runs else [run_a(s) for s in range(seeds)]
    B = runs[1] if runs else [run_b(s) for s in range(seeds)]
    bad = [r for r in compare_scalars(A, B, model, simType, iL) if not abs(r[1]) <= r[2]]
    if not bad:
        return []
    A = A + [run_a(s) for s in range(seeds, seeds + extra)]
    B = B + [run_b(s) for s in range(seeds, seeds + extra)]
    names = {r[0] for r in bad}
    return [r for r in compare_scalars(A, B, model, simType, iL) if r[0] in names and not abs(r[1]) <= r[2]]
