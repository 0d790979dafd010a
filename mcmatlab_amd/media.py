"""Host-side builders for the hot path's input arrays that the reference
produces in MATLAB before calling the MEX files.

- interface_normals: the (theta, phi) per-voxel surface-normal field that
  MCmatlab.c reads as MC.interfaceNormals, built like
  getOpticalMediaProperties.m:222-244 (Sobel gradient of each
  refractive-index class, smoothn-smoothed); interface_normals_fast, a
  Gaussian-smoothed variant for large synthetic sweeps.
- phase_function_cdf / unique_cdfs: the CDF table of custom phase functions
  (getOpticalMediaProperties.m:94-116).
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


def _sobel_gradient(ind):
    """The reference's 3-D Sobel gradient of an indicator field
    (@model/private/imgradientxyz.m): along each axis the difference of the
    two neighbours, weighted over the orthogonal 3 x 3 neighbourhood by
    [[1,3,1],[3,6,3],[1,3,1]], edges replicated; returns (G_x, G_y, G_z) along
    the array's axes 0, 1, 2."""
    from scipy.ndimage import correlate

    S = np.array([[1.0, 3.0, 1.0], [3.0, 6.0, 3.0], [1.0, 3.0, 1.0]])
    out = []
    for ax in range(3):
        k = np.zeros((3, 3, 3))
        others = [a for a in range(3) if a != ax]
        for i, sgn in ((0, -1.0), (2, 1.0)):
            sl = [slice(None)] * 3
            sl[ax] = i
            k[tuple(sl)] = sgn * S  # S spans the two other axes in their order
        out.append(correlate(ind, k, mode="nearest"))
    return out


def _smoothn(fields, s, spacing):
    """Penalised least-squares smoothing with a given smoothness s, as
    @model/private/smoothn.m computes it for complete, equally weighted data
    (no robust iterations): z = IDCT(Gamma * DCT(y)) with the orthonormal
    3-D DCT-II, Gamma = 1 / (1 + s Lambda^2) and
    Lambda = sum_i (2 - 2 cos(pi k_i / n_i)) / dI_i^2, dI = spacing / max(spacing)."""
    from scipy.fft import dctn, idctn

    shape = fields[0].shape
    dI = np.asarray(spacing, np.float64) / max(spacing)
    lam = np.zeros(shape)
    for i, n in enumerate(shape):
        sh = [1] * len(shape)
        sh[i] = n
        lam = lam + ((2 - 2 * np.cos(np.pi * np.arange(n) / n)) / dI[i] ** 2).reshape(sh)
    gamma = 1.0 / (1.0 + s * lam ** 2)
    return [idctn(gamma * dctn(f, type=2, norm="ortho", workers=-1), type=2, norm="ortho", workers=-1) for f in fields]


def interface_normals(M, n_of_medium, d, smoothingLengthScale=0.1):
    """MC.interfaceNormals single [2, L] for the media map M (1-based,
    [nx,ny,nz]) and per-medium refractive indices n_of_medium [nM], as
    getOpticalMediaProperties.m:222-244 builds them: for each distinct
    refractive index, the Sobel gradient of its indicator divided by the voxel
    sizes, smoothed by smoothn with smoothness max(1, round(smoothingLengthScale
    / max(d))) (skipped for a non-positive length scale), kept on that class's
    voxels; then (theta, phi) = (pi/2 - elevation, azimuth) of cart2sph."""
    M = np.asarray(M)
    ri = np.asarray(n_of_medium, dtype=np.float64)[M - 1]
    dx, dy, dz = d
    G = np.full((3,) + M.shape, np.nan)
    for c in np.unique(ri):
        sel = ri == c
        gx, gy, gz = _sobel_gradient(sel.astype(np.float64))
        comps = [gx / dx, gy / dy, gz / dz]
        if smoothingLengthScale > 0:
            comps = _smoothn(comps, max(1, round(smoothingLengthScale / max(dx, dy, dz))), (dx, dy, dz))
        for k in range(3):
            G[k][sel] = comps[k][sel]
    phi = np.arctan2(G[1], G[0])
    elev = np.arctan2(G[2], np.hypot(G[0], G[1]))
    theta = np.pi / 2 - elev
    return np.stack([theta.ravel(order="F"), phi.ravel(order="F")]).astype(np.float32)


def interface_normals_fast(M, n_of_medium, d, smoothing_voxels=1.5):
    """A cheaper normal field for large synthetic sweeps: each refractive-index
    class's indicator Gaussian-smoothed and differentiated by central
    differences (normals point into the voxel's own class, like the
    reference's).  Not the reference's construction (interface_normals)."""
    from scipy.ndimage import gaussian_filter

    M = np.asarray(M)
    ri = np.asarray(n_of_medium, dtype=np.float64)[M - 1]
    G = np.zeros((3,) + M.shape)
    for c in np.unique(ri):
        ind = (ri == c).astype(np.float64)
        sm = gaussian_filter(ind, smoothing_voxels, mode="nearest") if smoothing_voxels > 0 else ind
        gx, gy, gz = np.gradient(sm, d[0], d[1], d[2])
        sel = ri == c
        G[0][sel], G[1][sel], G[2][sel] = gx[sel], gy[sel], gz[sel]
    phi = np.arctan2(G[1], G[0])
    elev = np.arctan2(G[2], np.hypot(G[0], G[1]))
    theta = np.pi / 2 - elev
    return np.stack([theta.ravel(order="F"), phi.ravel(order="F")]).astype(np.float32)


def phase_function_cdf(customPhaseFunc, wavelength=None, nThetas=200):
    """The 201-point CDF of the polar angle for a medium with a custom phase
    function, as getOpticalMediaProperties.m:94-104 builds it:
    theta_k = pi/nThetas (k - 1/2), PDF_k = sin(theta_k) p(theta_k),
    CDF = [0, cumsum(PDF) / sum(PDF)].  customPhaseFunc(wavelength, theta)."""
    th = np.pi / nThetas * (np.arange(1, nThetas + 1) - 0.5)
    pdf = np.array([np.sin(t) * float(customPhaseFunc(wavelength, t)) for t in th])
    if not np.all(np.isfinite(pdf)) or np.any(pdf < 0):
        raise ValueError("the custom phase function must return real, finite, non-negative numbers")
    return np.concatenate([[0.0], np.cumsum(pdf) / pdf.sum()])


def unique_cdfs(cdfs):
    """(CDFs [201, nCDF], CDFidx per medium) with identical CDFs shared, in
    order of first use (getOpticalMediaProperties.m:106-116); None entries
    (media with a Henyey-Greenstein g) get CDFidx NaN."""
    table, idx = [], []
    for c in cdfs:
        if c is None:
            idx.append(np.nan)
            continue
        c = np.asarray(c, np.float64)
        for k, t in enumerate(table):
            if np.array_equal(t, c):
                idx.append(float(k))
                break
        else:
            table.append(c)
            idx.append(float(len(table) - 1))
    return (np.stack(table, axis=1) if table else np.zeros((201, 0))), np.array(idx)


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

