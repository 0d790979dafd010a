"""Host-side builders for the hot path's input arrays that the reference
produces in MATLAB before calling the MEX files.

- interface_normals: the (theta, phi) per-voxel surface-normal field that
  MCmatlab.c reads as MC.interfaceNormals (getOpticalMediaProperties.m:222-244
  builds it from Sobel gradients of each refractive-index class, smoothed).
  Here each class indicator is Gaussian-smoothed and differentiated with
  central differences; the normal points into the voxel's own class, as in the
  reference, and is stored as the same spherical angles.
- heat_coefficients / heat_dtmax: dTdtperdeltaT and the explicit stability
  limit (simulateHeatDistribution.m:469-514); conductivity_gradient: the
  chain rule from dJ/d dTdtperdeltaT to dJ/dTC.
- sensor_weights: tempSensorCornerIdxs / tempSensorInterpWeights
  (simulateHeatDistribution.m:543-609).
- bin_medium: the sub-media of an FR/T/FD-dependent medium
  (getOpticalMediaProperties.m:148-163, getThermalMediaProperties.m:75-84):
  the GPU binning followed by the callers' sortrows relabelling.
"""
from __future__ import annotations

import numpy as np


def interface_normals(M, n_of_medium, d, smoothing_voxels=1.5):
    """MC.interfaceNormals single [2, L] for media map M (1-based, [nx,ny,nz])
    and per-medium refractive index n_of_medium[nM]."""
    from scipy.ndimage import gaussian_filter

    M = np.asarray(M)
    ri = np.asarray(n_of_medium, dtype=np.float64)[M - 1]
    classes = np.unique(ri)
    G = np.zeros((3,) + M.shape)
    for c in classes:
        ind = (ri == c).astype(np.float64)
        sm = gaussian_filter(ind, smoothing_voxels, mode="nearest") if smoothing_voxels > 0 else ind
        gx, gy, gz = np.gradient(sm, d[0], d[1], d[2])
        sel = ri == c
        G[0][sel], G[1][sel], G[2][sel] = gx[sel], gy[sel], gz[sel]
    phi = np.arctan2(G[1], G[0])
    elev = np.arctan2(G[2], np.hypot(G[0], G[1]))
    theta = np.pi / 2 - elev
    return np.stack([theta.ravel(order="F"), phi.ravel(order="F")]).astype(np.float32)


def heat_coefficients(VHC, TC, dx, dy, dz):
    """dTdtperdeltaT single [nM, nM, 3]: c[a, b, d] = TC_eff(a, b) / VHC(a) / d^2
    with TC_eff the harmonic mean (simulateHeatDistribution.m:469-483)."""
    VHC = np.asarray(VHC, dtype=np.float32)
    TC = np.asarray(TC, dtype=np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        TCe = (2 * np.outer(TC, TC) / (TC[None, :] + TC[:, None])).astype(np.float32)
    TCe[np.isnan(TCe)] = 0
    base = TCe / VHC[:, None]
    return np.stack([base / np.float32(dx ** 2), base / np.float32(dy ** 2), base / np.float32(dz ** 2)], axis=2).astype(np.float32)


def conductivity_gradient(dJ_dc, VHC, TC, dx, dy, dz):
    """dJ/dTC [nM] from dJ/dc [nM, nM, 3] (the heat adjoint's gradient with
    respect to dTdtperdeltaT), by the chain rule through
    c[a, b, d] = TC_eff(a, b) / VHC(a) / d^2, TC_eff = 2 TC_a TC_b / (TC_a + TC_b)
    (simulateHeatDistribution.m:469-483): dTC_eff/dTC_a = 2 TC_b^2 / (TC_a + TC_b)^2."""
    g = np.asarray(dJ_dc, dtype=np.float64)
    VHC = np.asarray(VHC, dtype=np.float64)
    TC = np.asarray(TC, dtype=np.float64)
    w = (g[:, :, 0] / dx ** 2 + g[:, :, 1] / dy ** 2 + g[:, :, 2] / dz ** 2) / VHC[:, None]  # dJ / dTC_eff[a, b]
    s = TC[:, None] + TC[None, :]
    with np.errstate(invalid="ignore", divide="ignore"):
        da = np.where(s > 0, 2 * TC[None, :] ** 2 / s ** 2, 0.0)  # dTC_eff[a, b] / dTC_a
        db = np.where(s > 0, 2 * TC[:, None] ** 2 / s ** 2, 0.0)  # dTC_eff[a, b] / dTC_b
    return (w * da).sum(axis=1) + (w * db).sum(axis=0)


def heat_dtmax(M, VHC, TC, dx, dy, dz):
    """Largest stable forward-Euler step over all voxels (simulateHeatDistribution.m:496-514)."""
    M = np.asarray(M) - 1
    TCv = np.asarray(TC, dtype=np.float32)[M]
    VHCv = np.asarray(VHC, dtype=np.float32)[M]
    tot = np.zeros(M.shape, np.float32)
    for ax, d in ((0, dx), (1, dy), (2, dz)):
        a = np.moveaxis(TCv, ax, 0)
        with np.errstate(invalid="ignore", divide="ignore"):
            e = 2 * a[:-1] * a[1:] / (a[:-1] + a[1:])
        e[np.isnan(e)] = 0
        t = np.moveaxis(tot, ax, 0)
        t[:-1] += e / np.float32(d ** 2)
        t[1:] += e / np.float32(d ** 2)
    with np.errstate(divide="ignore"):
        return float(np.min(VHCv / tot))


def sensor_weights(positions, x, y, z):
    """(tempSensorCornerIdxs [nS], tempSensorInterpWeights [nS, 3]) for sensor
    positions [nS, 3] on voxel-centre axes x, y, z (simulateHeatDistribution.m:543-609)."""
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    axes = (np.asarray(x), np.asarray(y), np.asarray(z))
    nS = positions.shape[0]
    li = np.zeros(nS)
    w = np.zeros((nS, 3))
    for i in range(nS):
        idx = []
        for a in range(3):
            ax = axes[a]
            n, dd = ax.size, ax[1] - ax[0]
            rel = positions[i, a] - ax[0]
            if rel < 0:
                k, ww = 0, 0.0
            elif rel < ax[-1] - ax[0]:
                k, ww = int(np.floor(rel / dd)), np.mod(rel / dd, 1)
            else:
                k, ww = n - 2, 1.0
            idx.append(k)
            w[i, a] = ww
        li[i] = idx[0] + idx[1] * axes[0].size + idx[2] * axes[0].size * axes[1].size
    return li, w


def bin_medium(paramVals, nBins, device=0):
    """Sub-media of one FR/T/FD-dependent medium as the MATLAB callers build
    them (getOpticalMediaProperties.m:148-163, getThermalMediaProperties.m:
    75-84): getDownsampledParamVals on the GPU, then the nBins downsampled
    sets sorted as `sortrows` does (ascending by the first row, then the
    second, then the third; equal sets keep their order) and every voxel's
    label renumbered to the sorted order.  paramVals is 3 x N1 (mua, mus, g
    or VHC, TC, 1) single; returns (sorted sets 3 x nBins single, labels 1..
    nBins per set, int); the medium's voxels then get M = firstSubmedium +
    label - 1."""
    from . import api

    if not 1 <= int(nBins) <= 255:
        raise ValueError("nBins must be 1..255 (the labels are uint8)")
    ds, b = api.getDownsampledParamVals(paramVals, int(nBins), device)
    order = np.lexsort((ds[2], ds[1], ds[0]))  # stable, first row most significant
    rank = np.empty(int(nBins), np.int64)
    rank[order] = np.arange(1, int(nBins) + 1)
    return ds[:, order], rank[b.astype(np.int64) - 1]

