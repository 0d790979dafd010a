// Host-side preparation of a Monte Carlo launch from an mcm_mc_input:
// validation, the per-wavelength media tables, beam profile tables, source
// basis vectors, light-collector constants and deposition criteria.  Follows
// the MEX body's preprocessing (+MCmatlab/src/MCmatlab.c:322-615) and is
// shared by the GPU plan and the host debug harness.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mc_photon.h"
#include "mcm_abi.h"

namespace mcm {

struct McTables {
  int nx = 0, ny = 0, nz = 0, nM = 0, nL = 0;
  int64_t L = 0;
  std::vector<uint8_t> M0;                  // 0-based media map
  bool fresnelPossible = false;             // some wavelength has media of different n
  std::vector<float4> normals;              // unit normal per voxel (w filled per wavelength)
  std::vector<Medium> media;                // [nL * 256]
  std::vector<uint8_t> cdfIdx;              // [nL * 256]
  std::vector<float> cdfs;                  // [201 * nCDF] (at least 201)
  std::vector<float> fpid1, aid1, fpid2, aid2;  // beam profile tables (at least 1 element each)
  std::vector<double> spectrum;             // power per wavelength (beam sources)
  std::vector<double> muaNorm;              // [nL * nM] FP64 mua for normalisation
  McParams base;                            // all scalar fields; pointers left null
  bool timed = false;
  double nPhotonsRequested = 0, simulationTimeRequested = 0;
  bool requestCollected = false;
  bool calcNFR = false;
  int lcRes0 = 0, lcRes1 = 1;
  int64_t nImage = 0, nFF = 0, nNI[6] = {0, 0, 0, 0, 0, 0};
  int64_t nExamplePaths = 0;
  double V = 0, dA[3] = {0, 0, 0};          // voxel volume; face pixel areas yz, xz, xy
};

inline bool finite_or_inf_nonneg(double x) { return !std::isnan(x) && x >= 0; }

// infCast (MCmatlablib.c:138) into the kernel's 32-bit counters
inline uint32_t cast_count(double x) {
  if (std::isinf(x) || x >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

// Beam profile table (MCmatlab.c:539-575): radius-weighted cumulative
// trapezoids for sourceType 4, plain ones for 5; one element d means a named
// profile, code -1 - d.
inline void beam_table(const double* src, int64_t L, bool radial, std::vector<float>& out, float& code) {
  out.assign(L > 1 ? (size_t)L : 1, 0.f);
  if (L <= 1) {
    code = (float)(-1.0 - (src ? src[0] : 0.0));
    return;
  }
  code = 0.f;
  std::vector<double> d((size_t)L);
  d[0] = 0;
  for (int64_t k = 1; k < L; ++k) d[k] = d[k - 1] + (radial ? (double)(k - 1) : 1.0) * src[k - 1] + (radial ? (double)k : 1.0) * src[k];
  for (int64_t k = 0; k < L; ++k) out[k] = (float)(d[k] / d[L - 1]);
  out[L - 1] = 1.0f;
}

inline void unit_cross_d(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  const double n = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int k = 0; k < 3; ++k) c[k] /= n;
}

inline void rotate_d(const double* r, const double* a, double th, double* o) {
  const double st = std::sin(th), ct = std::cos(th);
  o[0] = (a[0] * a[0] * (1 - ct) + ct) * r[0] + (a[0] * a[1] * (1 - ct) - a[2] * st) * r[1] + (a[0] * a[2] * (1 - ct) + a[1] * st) * r[2];
  o[1] = (a[1] * a[0] * (1 - ct) + a[2] * st) * r[0] + (a[1] * a[1] * (1 - ct) + ct) * r[1] + (a[1] * a[2] * (1 - ct) - a[0] * st) * r[2];
  o[2] = (a[2] * a[0] * (1 - ct) - a[1] * st) * r[0] + (a[2] * a[1] * (1 - ct) + a[0] * st) * r[1] + (a[2] * a[2] * (1 - ct) + ct) * r[2];
}

// Returns an empty string on success, else the reason the input is invalid.
inline std::string build_mc_tables(const mcm_mc_input* in, McTables& T) {
  if (!in) return "null input";
  if (in->nx <= 0 || in->ny <= 0 || in->nz <= 0) return "grid dimensions must be positive";
  if (in->nx > 4096 || in->ny > 4096 || in->nz > 4096 || in->nx * in->ny * in->nz > (int64_t)1 << 31)
    return "grid too large for 32-bit voxel indices";
  if (in->nM <= 0 || in->nM > 256) return "number of media must be 1..256";
  if (in->nL <= 0) return "number of wavelengths must be positive";
  if (!in->M || !in->mua || !in->mus || !in->g || !in->n) return "missing M or mediaProperties";
  if (!(in->dx > 0 && in->dy > 0 && in->dz > 0)) return "voxel pitch must be positive";
  if (in->boundaryType < 0 || in->boundaryType > 3) return "boundaryType must be 0..3";
  T.nx = (int)in->nx;
  T.ny = (int)in->ny;
  T.nz = (int)in->nz;
  T.nM = (int)in->nM;
  T.nL = (int)in->nL;
  T.L = in->nx * in->ny * in->nz;
  T.M0.resize((size_t)T.L);
  for (int64_t k = 0; k < T.L; ++k) {
    const int m = in->M[k];
    if (m < 1 || m > T.nM) return "M holds a medium index outside 1..nM";
    T.M0[k] = (uint8_t)(m - 1);
  }
  T.timed = std::isnan(in->nPhotonsRequested);
  T.nPhotonsRequested = in->nPhotonsRequested;
  T.simulationTimeRequested = in->simulationTimeRequested;
  if (!T.timed && !(in->nPhotonsRequested >= 1)) return "nPhotonsRequested must be positive";
  if (T.timed && !(in->simulationTimeRequested > 0)) return "simulationTimeRequested must be positive";
  T.requestCollected = in->requestCollectedPhotons != 0;
  T.calcNFR = in->calcNFR != 0;
  const bool src3d = in->sourceDistribution != nullptr;
  const int sourceType = src3d ? -1 : (int)in->sourceType;
  if (!src3d && (sourceType < 0 || sourceType > 5)) return "sourceType must be 0..5";

  // media per wavelength, padded to 256 entries so any uint8 index is in bounds
  T.media.assign((size_t)T.nL * 256, Medium{0.f, 0.f, 0.f, 1.f});
  T.cdfIdx.assign((size_t)T.nL * 256, 0);
  T.muaNorm.assign((size_t)T.nL * T.nM, 0.0);
  const int64_t nCDF = in->nCDFs / kCdfSize;
  for (int iL = 0; iL < T.nL; ++iL) {
    for (int m = 0; m < T.nM; ++m) {
      const int64_t k = m + (int64_t)iL * T.nM;
      Medium md{(float)in->mua[k], (float)in->mus[k], (float)in->g[k], (float)in->n[k]};
      if (!(md.mus >= 0) && !std::isnan(md.g)) return "mus must be non-negative";
      T.media[(size_t)iL * 256 + m] = md;
      T.muaNorm[k] = in->mua[k];
      int ci = 0;
      if (in->CDFidx && !std::isnan(in->CDFidx[k])) ci = (int)in->CDFidx[k];
      if (std::isnan(in->g[k])) {
        if (ci < 0 || ci >= nCDF) return "CDFidx refers to a missing phase-function table";
      } else {
        ci = 0;
      }
      T.cdfIdx[(size_t)iL * 256 + m] = (uint8_t)ci;
      if (md.ri != T.media[(size_t)iL * 256].ri) T.fresnelPossible = true;
    }
  }
  T.cdfs.assign((size_t)std::max<int64_t>(nCDF, 1) * kCdfSize, 0.f);
  for (int64_t k = 0; k < nCDF * kCdfSize; ++k) T.cdfs[k] = (float)in->CDFs[k];
  for (int64_t c = 0; c < nCDF; ++c) T.cdfs[c * kCdfSize + kCdfSize - 1] = 1.0f;
  if (T.fresnelPossible) {
    if (!in->interfaceNormals) return "media with different refractive indices need interfaceNormals";
    T.normals.resize((size_t)T.L);
    for (int64_t k = 0; k < T.L; ++k) {
      const double th = in->interfaceNormals[2 * k], ph = in->interfaceNormals[2 * k + 1];
      T.normals[k] = make_float4((float)(std::sin(th) * std::cos(ph)), (float)(std::sin(th) * std::sin(ph)), (float)std::cos(th), 1.f);
    }
  }

  McParams& p = T.base;
  std::memset(&p, 0, sizeof p);
  p.n[0] = T.nx;
  p.n[1] = T.ny;
  p.n[2] = T.nz;
  p.d[0] = (float)in->dx;
  p.d[1] = (float)in->dy;
  p.d[2] = (float)in->dz;
  for (int a = 0; a < 3; ++a) p.inv_d[a] = (float)(1.0 / (a == 0 ? in->dx : a == 1 ? in->dy : in->dz));
  p.boundaryType = in->boundaryType;
  p.nM = T.nM;
  p.L = T.L;
  p.calcNFR = T.calcNFR;
  p.nfrReplicas = 1;
  T.V = in->dx * in->dy * in->dz;
  T.dA[0] = in->dy * in->dz;
  T.dA[1] = in->dx * in->dz;
  T.dA[2] = in->dx * in->dy;

  // source (MCmatlab.c:363-370, 534-614)
  p.beamType = sourceType;
  p.emitterLength = src3d ? 0.f : (float)in->emitterLength;
  const int64_t L_FPID1 = sourceType >= 4 ? (sourceType == 4 ? in->FPID.n_radialDistr : in->FPID.n_XDistr) : 0;
  const int64_t L_AID1 = sourceType >= 4 ? (sourceType == 4 ? in->AID.n_radialDistr : in->AID.n_XDistr) : 0;
  const int64_t L_FPID2 = sourceType == 5 ? in->FPID.n_YDistr : 0;
  const int64_t L_AID2 = sourceType == 5 ? in->AID.n_YDistr : 0;
  T.fpid1.assign(1, 0.f);
  T.aid1.assign(1, 0.f);
  T.fpid2.assign(1, 0.f);
  T.aid2.assign(1, 0.f);
  if (sourceType >= 4) {
    const double* f1 = sourceType == 4 ? in->FPID.radialDistr : in->FPID.XDistr;
    const double* a1 = sourceType == 4 ? in->AID.radialDistr : in->AID.XDistr;
    if (L_FPID1 < 1 || L_AID1 < 1 || !f1 || !a1) return "beam profile arrays missing";
    beam_table(f1, L_FPID1, sourceType == 4, T.fpid1, p.FPIDcode1);
    beam_table(a1, L_AID1, sourceType == 4, T.aid1, p.AIDcode1);
    if (sourceType == 5) {
      if (L_FPID2 < 1 || L_AID2 < 1 || !in->FPID.YDistr || !in->AID.YDistr) return "beam profile arrays missing";
      beam_table(in->FPID.YDistr, L_FPID2, false, T.fpid2, p.FPIDcode2);
      beam_table(in->AID.YDistr, L_AID2, false, T.aid2, p.AIDcode2);
    }
  }
  p.L_FPID1 = (int)L_FPID1;
  p.L_AID1 = (int)L_AID1;
  p.L_FPID2 = (int)L_FPID2;
  p.L_AID2 = (int)L_AID2;
  p.FPIDwidth1 = (float)(src3d ? 0 : (sourceType == 5 ? in->FPID.XWidth : in->FPID.radialWidth));
  p.FPIDwidth2 = (float)(src3d ? 0 : (sourceType == 5 ? in->FPID.YWidth : 0));
  p.AIDwidth1 = (float)(src3d ? 0 : (sourceType == 5 ? in->AID.XWidth : in->AID.radialWidth));
  p.AIDwidth2 = (float)(src3d ? 0 : (sourceType == 5 ? in->AID.YWidth : 0));
  p.tanAID1 = (float)std::tan((double)p.AIDwidth1);
  p.tanAID2 = (float)std::tan((double)p.AIDwidth2);
  const double tb = src3d ? 0 : in->LS_theta, pb = src3d ? 0 : in->LS_phi, psib = src3d ? 0 : in->LS_psi;
  double u[3] = {std::sin(tb) * std::cos(pb), std::sin(tb) * std::sin(pb), std::cos(tb)};
  double w[3] = {1, 0, 0}, v[3] = {0, 0, 1};
  if (u[2] != 1) unit_cross_d(u, v, w);
  rotate_d(w, u, psib, v);
  unit_cross_d(u, v, w);
  for (int k = 0; k < 3; ++k) {
    p.u[k] = (float)u[k];
    p.v[k] = (float)v[k];
    p.w[k] = (float)w[k];
  }
  p.focus[0] = (float)(src3d ? 0 : in->xFocus);
  p.focus[1] = (float)(src3d ? 0 : in->yFocus);
  p.focus[2] = (float)(src3d ? 0 : in->zFocus);
  if (!src3d && (sourceType == 0 || sourceType >= 3) && p.u[2] == 0.f) return "beam axis parallel to the surface";
  T.spectrum.assign((size_t)T.nL, 1.0);
  if (!src3d) {
    if (!in->spectrum) return "spectrum missing";
    for (int iL = 0; iL < T.nL; ++iL) T.spectrum[iL] = in->spectrum[iL];
  }

  // light collector (MCmatlab.c:427-447)
  p.useLC = in->useLightCollector != 0;
  {
    const double f = in->LC_f, th = in->LC_theta, ph = in->LC_phi;
    const bool fin = std::isfinite(f);
    p.lcR[0] = (float)(in->LC_x - (fin ? f * std::sin(th) * std::cos(ph) : 0));
    p.lcR[1] = (float)(in->LC_y - (fin ? f * std::sin(th) * std::sin(ph) : 0));
    p.lcR[2] = (float)(in->LC_z - (fin ? f * std::cos(th) : 0));
    p.lcTheta = (float)th;
    p.lcPhi = (float)ph;
    p.lcSinT = (float)std::sin(th);
    p.lcCosT = (float)std::cos(th);
    p.lcSinP = (float)std::sin(ph);
    p.lcCosP = (float)std::cos(ph);
    p.lcF = (float)f;
    p.lcFinite = fin;
    p.lcDiam = (float)in->LC_diam;
    p.lcFSorNA = (float)(fin ? in->LC_fieldSize : in->LC_NA);
    p.lcNAangle = (float)std::asin(std::fmin(1.0, fin ? 1.0 : in->LC_NA));
    const int64_t nTB = src3d ? 0 : (int64_t)in->LC_nTimeBins;
    T.lcRes0 = (int)in->LC_res;
    T.lcRes1 = nTB ? (int)nTB + 2 : 1;
    p.lcRes0 = T.lcRes0;
    p.lcRes1 = T.lcRes1;
    p.lcTStart = (float)(src3d ? 0 : in->LC_tStart);
    p.lcTEnd = (float)(src3d ? 0 : in->LC_tEnd);
    if (p.useLC) {
      if (fin && T.lcRes0 < 1) return "lightCollector.res must be >= 1 for a lens";
      if (!fin && T.lcRes0 != 1) return "lightCollector.res must be 1 when lightCollector.f is Inf";  // runMonteCarlo.m:364-365
      T.nImage = (int64_t)T.lcRes0 * T.lcRes0 * T.lcRes1;
    }
  }
  p.farFieldRes = (int)in->farFieldRes;
  if (p.farFieldRes < 0) return "farFieldRes must be non-negative";
  T.nFF = (int64_t)p.farFieldRes * p.farFieldRes;
  p.tooOld = (float)(1000.0 / 29979245800.0 * std::sqrt(std::pow(in->dx * in->nx, 2) + std::pow(in->dy * in->ny, 2)));
  const int64_t nxy = in->nx * in->ny, nxz = in->nx * in->nz, nyz = in->ny * in->nz;
  if (in->boundaryType == 1) {
    T.nNI[0] = T.nNI[1] = nyz;
    T.nNI[2] = T.nNI[3] = nxz;
    T.nNI[4] = T.nNI[5] = nxy;
  } else if (in->boundaryType == 2) {
    T.nNI[5] = nxy * kKillRange * kKillRange;
  } else if (in->boundaryType == 3) {
    T.nNI[4] = T.nNI[5] = nxy;
  }

  // deposition criteria (MCmatlab.c:402-424)
  p.minS = cast_count(in->minScatterings);
  p.maxS = cast_count(in->maxScatterings);
  p.minRefr = cast_count(in->minRefractions);
  p.maxRefr = cast_count(in->maxRefractions);
  p.minRefl = cast_count(in->minReflections);
  p.maxRefl = cast_count(in->maxReflections);
  p.minI = cast_count(in->minInterfaceTransitions);
  p.maxI = cast_count(in->maxInterfaceTransitions);
  const uint32_t mi = cast_count(in->minSubmediaIdx), ma = cast_count(in->maxSubmediaIdx);
  p.minIdx = mi == 0xFFFFFFFFu ? 0xFFFFFFFEu : mi - 1;
  p.maxIdx = ma == 0xFFFFFFFFu ? 0xFFFFFFFEu : ma - 1;
  p.onlyCollected = in->onlyCollected != 0;
  p.evalAtEnd = in->evaluateOnlyAtEndOfLife != 0;
  const bool unrestrictive = p.minS == 0 && p.maxS == 0xFFFFFFFFu && p.minRefr == 0 && p.maxRefr == 0xFFFFFFFFu && p.minRefl == 0 &&
                             p.maxRefl == 0xFFFFFFFFu && p.minI == 0 && p.maxI == 0xFFFFFFFFu && !p.onlyCollected;
  if (unrestrictive) p.evalAtEnd = 0;
  p.criteriaActive = !unrestrictive;
  // the time of flight is read only by time-binned light collectors and the
  // boundary type 3 age limit
  p.needTime = (p.useLC && p.lcRes1 > 1) || p.boundaryType == 3;
  T.nExamplePaths = in->nExamplePaths;
  return "";
}

}  // namespace mcm
