// FP32 photon-packet physics of the voxel Monte Carlo transport.
//
// Restates +MCmatlab/src/MCmatlablib.c (launchPhoton :519-732, propagatePhoton
// :1019-1119, checkEscape :821-889, getInterpolatedNormal :920-1014,
// checkRoulette :1124-1133, scatterPhoton :1138-1184, formImage/formFarField/
// formEdgeFluxes :737-816) for one packet, in FP32 like the reference's own
// GPU build (MCmatlab.c:91-108), with FP64 tallies.
//
// Everything here is __host__ __device__ and free of GPU intrinsics: the GPU
// kernel schedules these steps across a wavefront, and a host build of the
// same functions runs packets one at a time for debugging and tests.  Tally
// writes go through a Sink (device atomics or host adds).
//
// Departures from the reference, all FP32-safety guards:
//   - table searches are bounded bisections returning the unique j with
//     D[j] < x <= D[j+1], clamped to the table (the reference's probe loop
//     (:263-276) can run away when the last FP32 entry rounds below 1);
//   - tally indices are clamped into their arrays;
//   - the scatter rotation falls back to the polar formula when ux = uy = 0
//     while |uz| < 1 (possible after FP32 refraction of an axial ray), where
//     the reference formula divides by zero;
//   - a packet is retired after MAX_SUBSTEPS sub-steps (never reached by a
//     physical packet; it bounds every lane's work).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "philox.h"

namespace mcm {

constexpr float kC = 29979245800.0f;          // cm/s, MCmatlab.c:159
constexpr float kThreshold = 0.01f;           // roulette, MCmatlab.c:160
constexpr float kChance = 0.1f;               // MCmatlab.c:161
constexpr int kKillRange = 5;                 // MCmatlab.c:165
constexpr int kCdfSize = 201;                 // MCmatlab.c:171
constexpr float kPi = 3.14159265358979323846f;
constexpr float kEps = 1.1920928955078125e-07f;  // FLT_EPSILON
constexpr int kLaunchAttempts = 1000000;      // MCmatlablib.c:708
constexpr uint32_t kMaxSubsteps = 1u << 26;

// Per-medium optical constants of one wavelength.
struct Medium {
  float mua, mus, g, ri;
};

// Everything a launch needs; passed by value as the kernel argument.
struct McParams {
  // grid
  int n[3];
  float d[3];
  float inv_d[3];
  int boundaryType;
  int nM;
  const uint8_t* M;            // 0-based medium per voxel, x fastest
  const float4* normalRI;      // unit interface normal (x,y,z) and refractive index per voxel; null if matched
  const Medium* media;         // [nM] (the kernel stages it into LDS)
  const uint8_t* cdfIdx;       // [nM] custom phase-function table per medium
  const float* cdfs;           // [201 * nCDF]
  // source
  int beamType;                // 0..5; -1 = 3D source distribution
  float emitterLength;
  float focus[3], u[3], v[3], w[3];
  const float* FPIDdist1; int L_FPID1; float FPIDwidth1;
  const float* FPIDdist2; int L_FPID2; float FPIDwidth2;
  const float* AIDdist1; int L_AID1; float AIDwidth1;
  const float* AIDdist2; int L_AID2; float AIDwidth2;
  float FPIDcode1, FPIDcode2, AIDcode1, AIDcode2;  // -1 top-hat, -2 Gaussian, -3 Lambertian, 0 custom
  float tanAID1, tanAID2;
  const double* S;             // 3D source cumulative table [L + 1]
  int64_t L;
  // light collector
  int useLC;
  float lcR[3], lcTheta, lcPhi, lcF, lcDiam, lcFSorNA;
  float lcSinT, lcCosT, lcSinP, lcCosP;
  int lcRes0, lcRes1;
  float lcTStart, lcTEnd;
  int lcFinite;
  float lcNAangle;             // asin(min(1, NA)) for fibers
  // far field
  int farFieldRes;
  // bt 3 runaway limit on packet age
  float tooOld;
  // deposition criteria (counters count only media in [minIdx, maxIdx])
  uint32_t minS, maxS, minRefr, maxRefr, minRefl, maxRefl, minI, maxI, minIdx, maxIdx;
  int onlyCollected, evalAtEnd, criteriaActive;
  int needTime;                // track the time of flight
  // outputs (FP64 tallies)
  double* NFR;
  double* image;
  double* FF;
  double* NI[6];  // xpos, xneg, ypos, yneg, zpos, zneg
  int calcNFR;
  int nfrReplicas;  // copies of NFR (1, 2, 4 or 8), one per group of XCDs; folded before use
};

struct Photon {
  float i[3], u[3], D[3];
  float Dfull[3];  // d[a] / |u[a]|: voxel crossing length per axis for the current direction
  int j;
  float mua, mus, g, ri;
  int cdf;
  float stepLeft, weight, time;
  bool insideVolume, alive, sameVoxel;
  uint32_t scatterings, refractions, reflections, transitions;
  int kec;  // 0 killed, 1 escaped, 2 collected
  uint32_t substeps;
};

// Result of the movement half of a sub-step.
struct MoveResult {
  float s;       // geometric length of the sub-step [cm]
  float absorb;
  int jNew;
  int mNew;
  bool fresnel;  // refractive index changes: interface event pending
};

MCM_FN int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

MCM_FN int clamp_axis(float i, int n) { return i < 0.f ? 0 : (i >= (float)n ? n - 1 : (int)floorf(i)); }

// getNewj (MCmatlablib.c:281-285)
MCM_FN int voxel_of(const McParams& p, const float* i) {
  return clamp_axis(i[2], p.n[2]) * (p.n[0] * p.n[1]) + clamp_axis(i[1], p.n[1]) * p.n[0] + clamp_axis(i[0], p.n[0]);
}

// Unique j in [0, N-1] with D[j] < x <= D[j+1] for a nondecreasing table
// D[0..N] (the reference's binaryTreeSearch contract, MCmatlablib.c:263-276).
template <typename T>
MCM_FN int table_search(T x, int N, const T* D) {
  int lo = 0, hi = N;  // invariant: D[lo] < x (for lo > 0) and x <= D[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (D[mid] < x) lo = mid;
    else hi = mid;
  }
  return clampi(lo, 0, N - 1);
}

// axisrotate (MCmatlablib.c:251-258): rotate r about unit axis a by angle th
MCM_FN void rotate_about(const float* r, const float* a, float th, float* o) {
  float st, ct;
  sincosf(th, &st, &ct);
  const float omc = 1.f - ct;
  o[0] = (a[0] * a[0] * omc + ct) * r[0] + (a[0] * a[1] * omc - a[2] * st) * r[1] + (a[0] * a[2] * omc + a[1] * st) * r[2];
  o[1] = (a[1] * a[0] * omc + a[2] * st) * r[0] + (a[1] * a[1] * omc + ct) * r[1] + (a[1] * a[2] * omc - a[0] * st) * r[2];
  o[2] = (a[2] * a[0] * omc - a[1] * st) * r[0] + (a[2] * a[1] * omc + a[0] * st) * r[1] + (a[2] * a[2] * omc + ct) * r[2];
}

MCM_FN float boundary_distance(float i, float u, float d) {
  return u != 0.f ? (floorf(i) + (u > 0.f ? 1.f : 0.f) - i) * d / u : INFINITY;
}

// Per-axis crossing lengths of the current direction (after every change
// of direction).
MCM_FN void set_crossing_lengths(Photon& P, const McParams& p) {
  for (int a = 0; a < 3; ++a) P.Dfull[a] = p.d[a] / fabsf(P.u[a]);
}

MCM_FN bool inside_box(const McParams& p, const float* i) {
  return i[0] < (float)p.n[0] && i[0] >= 0.f && i[1] < (float)p.n[1] && i[1] >= 0.f && i[2] < (float)p.n[2] && i[2] >= 0.f;
}

// Inside the extended region a packet may wander in for boundary types 0 and 2
// (KILLRANGE times the cuboid, MCmatlablib.c:687-700, :826-840).
MCM_FN bool inside_killrange(const McParams& p, const float* i, bool topOpen) {
  const float h = kKillRange / 2.0f;
  const bool xy = fabsf(i[0] / (float)p.n[0] - 0.5f) < h && fabsf(i[1] / (float)p.n[1] - 0.5f) < h;
  if (topOpen) return xy && i[2] / (float)p.n[2] - 0.5f < h && i[2] >= 0.f;
  return xy && fabsf(i[2] / (float)p.n[2] - 0.5f) < h;
}

// depositionCriteriaMet (MCmatlablib.c:126-136)
MCM_FN bool criteria_met(const Photon& P, const McParams& p) {
  if (!p.criteriaActive) return true;
  const bool deadKilled = !P.alive && P.kec == 0;
  const uint32_t big = 0xFFFFFFFFu;
  return P.scatterings >= p.minS && (deadKilled ? big : P.scatterings) <= p.maxS && P.refractions >= p.minRefr &&
         (deadKilled ? big : P.refractions) <= p.maxRefr && P.reflections >= p.minRefl &&
         (deadKilled ? big : P.reflections) <= p.maxRefl && P.transitions >= p.minI && (deadKilled ? big : P.transitions) <= p.maxI &&
         (p.onlyCollected ? P.kec == 2 : true);
}

MCM_FN bool in_count_range(int m, const McParams& p) {
  return (uint32_t)m >= p.minIdx && (uint32_t)m <= p.maxIdx;
}

// Lambert W lower branch W_{-1}(x), -1/e <= x < 0 (LG01 radius sampling,
// MCmatlablib.c:586-588): Halley iteration (Corless et al. 1996, eq. 5.12)
// from the asymptotic start near 0 or the branch-point series near -1/e.
MCM_FN float lambert_wm1(float x) {
  if (x >= 0.f) return 0.f;
  const float q = x + 0.36787944117144233f;
  if (q <= 0.f) return -1.f;
  float w;
  if (q < 1e-3f) {
    const float pp = -sqrtf(2.0f * 2.718281828459045f * q);
    w = -1.0f + pp * (1.0f + pp * (-1.0f / 3.0f + pp * 11.0f / 72.0f));
  } else {
    const float l1 = logf(-x), l2 = logf(-l1);
    w = l1 - l2 + l2 / l1;
  }
  for (int it = 0; it < 12; ++it) {
    const float e = expf(w), pw = w + 1.0f;
    float t = w * e - x;
    const float den = e * pw - 0.5f * (pw + 1.0f) * t / pw;
    if (den == 0.f) break;
    t /= den;
    w -= t;
    if (fabsf(t) < 4.f * kEps * fmaxf(fabsf(w), 1.0f)) break;
  }
  return w;
}

MCM_FN float radial_or_tophat_draw(float code, float width, const float* dist, int L, PacketRng& R) {
  if (code == -1.f) return width * sqrtf(R.next());
  if (code == -2.f) return width * sqrtf(-0.5f * logf(R.next()));
  const float a = R.next();
  const int jb = table_search(a, L - 1, dist);
  return width * ((float)jb + R.next()) / (float)(L - 1);
}

// launchPhoton (MCmatlablib.c:519-732).  Returns false when 10^6 attempts
// all start outside the allowed region (the reference then aborts the run).
MCM_FN bool launch(Photon& P, const McParams& p, const Medium* media, PacketRng& R) {
  P.sameVoxel = false;
  P.weight = 1.f;
  P.kec = 0;
  P.substeps = 0;
  float target[3] = {0.f, 0.f, 0.f}, w0[3];
  int attempts = 0;
  const float hx = (float)p.n[0] * 0.5f, hy = (float)p.n[1] * 0.5f;
  do {
    float t2 = 0.f;  // distance-to-focus flag: time = -RI/C * |start - target|
    bool timeFromTarget = false;
    if (p.beamType < 0) {
      const int64_t L = p.L;
      const double x = R.next53();
      // bounded bisection over the FP64 cumulative table S[0..L]
      int64_t lo = 0, hi = L;
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (p.S[mid] < x) lo = mid;
        else hi = mid;
      }
      const int64_t j = lo;
      const int64_t nxy = (int64_t)p.n[0] * p.n[1];
      P.i[0] = (float)(j % p.n[0] + 1) - R.next();
      P.i[1] = (float)(j / p.n[0] % p.n[1] + 1) - R.next();
      P.i[2] = (float)(j / nxy + 1) - R.next();
      const float ct = 1.f - 2.f * R.next();
      const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
      float sp, cp;
      sincosf(2.f * kPi * R.next(), &sp, &cp);
      P.u[0] = st * cp;
      P.u[1] = st * sp;
      P.u[2] = ct;
      P.time = 0.f;
    } else {
      switch (p.beamType) {
        case 0:  // pencil beam through the focus
          P.i[0] = (p.focus[0] - p.focus[2] * p.u[0] / p.u[2]) * p.inv_d[0] + hx;
          P.i[1] = (p.focus[1] - p.focus[2] * p.u[1] / p.u[2]) * p.inv_d[1] + hy;
          P.i[2] = 0.f;
          for (int k = 0; k < 3; ++k) P.u[k] = p.u[k];
          for (int k = 0; k < 3; ++k) target[k] = p.focus[k];
          timeFromTarget = true;
          break;
        case 1: {  // isotropic point or line emitter
          const float r = R.next() - 0.5f;
          P.i[0] = (p.focus[0] + p.u[0] * r * p.emitterLength) * p.inv_d[0] + hx;
          P.i[1] = (p.focus[1] + p.u[1] * r * p.emitterLength) * p.inv_d[1] + hy;
          P.i[2] = (p.focus[2] + p.u[2] * r * p.emitterLength) * p.inv_d[2];
          const float ct = 1.f - 2.f * R.next();
          const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
          float sp, cp;
          sincosf(2.f * kPi * R.next(), &sp, &cp);
          P.u[0] = st * cp;
          P.u[1] = st * sp;
          P.u[2] = ct;
          P.time = 0.f;
          break;
        }
        case 2: {  // infinite plane wave over the cuboid (or the kill range)
          const float span = (p.boundaryType == 1 || p.boundaryType == 3) ? 1.f : (float)kKillRange;
          P.i[0] = span * (float)p.n[0] * (R.next() - 0.5f) + hx;
          P.i[1] = span * (float)p.n[1] * (R.next() - 0.5f) + hy;
          P.i[2] = 0.f;
          for (int k = 0; k < 3; ++k) P.u[k] = p.u[k];
          t2 = ((P.i[0] - hx) * p.d[0] * p.u[0] + (P.i[1] - hy) * p.d[1] * p.u[1]);
          break;
        }
        case 3: {  // Laguerre-Gaussian LG01
          rotate_about(p.v, p.u, R.next() * 2.f * kPi, w0);
          const float r = p.FPIDwidth1 * sqrtf((lambert_wm1(-R.next() * 0.36787944117144233f) + 1.f) / (-2.f)) / 1.50087f;
          for (int k = 0; k < 3; ++k) target[k] = p.focus[k] + r * w0[k];
          const float ph = p.AIDwidth1 * sqrtf((lambert_wm1(-R.next() * 0.36787944117144233f) + 1.f) / (-2.f)) / 1.50087f;
          rotate_about(p.u, w0, ph, P.u);
          timeFromTarget = true;
          break;
        }
        case 4: {  // radially factorizable beam
          rotate_about(p.v, p.u, R.next() * 2.f * kPi, w0);
          const float r = radial_or_tophat_draw(p.FPIDcode1, p.FPIDwidth1, p.FPIDdist1, p.L_FPID1, R);
          for (int k = 0; k < 3; ++k) target[k] = p.focus[k] + r * w0[k];
          rotate_about(p.v, p.u, R.next() * 2.f * kPi, w0);
          float ph;
          if (p.AIDcode1 == -1.f) ph = atanf(p.tanAID1 * sqrtf(R.next()));
          else if (p.AIDcode1 == -2.f) ph = atanf(p.tanAID1 * sqrtf(-0.5f * logf(R.next())));
          else if (p.AIDcode1 == -3.f) ph = asinf(sqrtf(R.next()));
          else {
            const float a = R.next();
            const int jb = table_search(a, p.L_AID1 - 1, p.AIDdist1);
            ph = atanf(p.tanAID1 * ((float)jb + R.next()) / (float)(p.L_AID1 - 1));
          }
          rotate_about(p.u, w0, ph, P.u);
          timeFromTarget = true;
          break;
        }
        case 5: {  // X/Y factorizable beam
          float X, Y;
          if (p.FPIDcode1 == -1.f) X = p.FPIDwidth1 * (R.next() * 2.f - 1.f);
          else if (p.FPIDcode1 == -2.f) {
            const float a = R.next(), b = R.next();
            X = p.FPIDwidth1 * sqrtf(-0.5f * logf(a)) * cosf(2.f * kPi * b);
          } else {
            const float a = R.next();
            const int jb = table_search(a, p.L_FPID1 - 1, p.FPIDdist1);
            X = p.FPIDwidth1 * (((float)jb + R.next()) / (float)(p.L_FPID1 - 1) * 2.f - 1.f);
          }
          if (p.FPIDcode2 == -1.f) Y = p.FPIDwidth2 * (R.next() * 2.f - 1.f);
          else if (p.FPIDcode2 == -2.f) {
            const float a = R.next(), b = R.next();
            Y = p.FPIDwidth2 * sqrtf(-0.5f * logf(a)) * cosf(2.f * kPi * b);
          } else {
            const float a = R.next();
            const int jb = table_search(a, p.L_FPID2 - 1, p.FPIDdist2);
            Y = p.FPIDwidth2 * (((float)jb + R.next()) / (float)(p.L_FPID2 - 1) * 2.f - 1.f);
          }
          for (int k = 0; k < 3; ++k) target[k] = p.focus[k] + X * p.v[k] + Y * p.w[k];
          if (p.AIDcode1 == -3.f) {
            rotate_about(p.v, p.u, R.next() * 2.f * kPi, w0);
            rotate_about(p.u, w0, asinf(sqrtf(R.next())), P.u);
          } else {
            float tx, ty;
            if (p.AIDcode1 == -1.f) tx = p.tanAID1 * (R.next() * 2.f - 1.f);
            else if (p.AIDcode1 == -2.f) {
              const float a = R.next(), b = R.next();
              tx = p.tanAID1 * sqrtf(-0.5f * logf(a)) * cosf(2.f * kPi * b);
            } else {
              const float a = R.next();
              const int jb = table_search(a, p.L_AID1, p.AIDdist1);
              tx = p.tanAID1 * (((float)jb + R.next()) / (float)p.L_AID1 * 2.f - 1.f);
            }
            if (p.AIDcode2 == -1.f) ty = p.tanAID2 * (R.next() * 2.f - 1.f);
            else if (p.AIDcode2 == -2.f) {
              const float a = R.next(), b = R.next();
              ty = p.tanAID2 * sqrtf(-0.5f * logf(a)) * cosf(2.f * kPi * b);
            } else {
              const float a = R.next();
              const int jb = table_search(a, p.L_AID2, p.AIDdist2);
              ty = p.tanAID2 * (((float)jb + R.next()) / (float)p.L_AID2 * 2.f - 1.f);
            }
            rotate_about(p.v, p.u, atan2f(tx, ty), w0);
            rotate_about(p.u, w0, atanf(sqrtf(tx * tx + ty * ty)), P.u);
          }
          timeFromTarget = true;
          break;
        }
        default:
          break;
      }
      if (p.beamType >= 3 || p.beamType == 0) {
        if (p.beamType != 0) {
          // start where the ray through the target crosses z = 0
          P.i[0] = (target[0] - target[2] * P.u[0] / P.u[2]) * p.inv_d[0] + hx;
          P.i[1] = (target[1] - target[2] * P.u[1] / P.u[2]) * p.inv_d[1] + hy;
          P.i[2] = 0.f;
        }
      }
    }
    P.j = voxel_of(p, P.i);
    P.ri = media[p.M[P.j]].ri;
    if (timeFromTarget) {
      // so that the packet crosses the focal plane at t = 0
      const float ex = (P.i[0] - hx) * p.d[0] - target[0], ey = (P.i[1] - hy) * p.d[1] - target[1], ez = P.i[2] * p.d[2] - target[2];
      P.time = -P.ri / kC * sqrtf(ex * ex + ey * ey + ez * ez);
    } else if (p.beamType == 2) {
      P.time = P.ri / kC * t2;
    }
    switch (p.boundaryType) {
      case 0: P.alive = inside_killrange(p, P.i, false); break;
      case 2: P.alive = inside_killrange(p, P.i, true); break;
      default: P.alive = inside_box(p, P.i); break;
    }
  } while (!P.alive && ++attempts < kLaunchAttempts);
  for (int k = 0; k < 3; ++k) P.D[k] = boundary_distance(P.i[k], P.u[k], p.d[k]);
  set_crossing_lengths(P, p);
  P.insideVolume = inside_box(p, P.i);
  P.stepLeft = -logf(R.next());
  P.scatterings = P.refractions = P.reflections = P.transitions = 0;
  return attempts < kLaunchAttempts;
}

// A launched packet as handed from the launch kernel to the transport
// kernel: position and refractive index, direction and time, distances to
// the next voxel boundaries and the remaining optical depth, and where the
// packet's random stream stands (block counter, words left in the current
// block, the unused words) plus flags.  The transport continues the stream
// exactly where launchPhoton left it, so the trajectory is the same as with
// an inline launch.  A launch draws at most a few words per attempt and makes
// at most 10^6 attempts, so the block counter stays below 2^24.
struct LaunchRecord {
  float4 a, b, c;
  uint4 d;
};

MCM_FN void pack_launch(const Photon& P, const PacketRng& R, bool ok, LaunchRecord& L) {
  L.a = make_float4(P.i[0], P.i[1], P.i[2], P.ri);
  L.b = make_float4(P.u[0], P.u[1], P.u[2], P.time);
  L.c = make_float4(P.D[0], P.D[1], P.D[2], P.stepLeft);
  const uint32_t flags = (R.block & 0xFFFFFFu) | ((uint32_t)R.have << 24) | (P.insideVolume ? 1u << 28 : 0u) | (ok ? 1u << 29 : 0u) |
                         (P.alive ? 1u << 30 : 0u);
  L.d = make_uint4(flags, R.b0, R.b1, R.b2);
}

// Photon state after launchPhoton from a record (R must have been started on
// the same packet).  Returns whether the launch succeeded.
MCM_FN bool unpack_launch(const LaunchRecord& L, Photon& P, PacketRng& R) {
  P.i[0] = L.a.x;
  P.i[1] = L.a.y;
  P.i[2] = L.a.z;
  P.ri = L.a.w;
  P.u[0] = L.b.x;
  P.u[1] = L.b.y;
  P.u[2] = L.b.z;
  P.time = L.b.w;
  P.D[0] = L.c.x;
  P.D[1] = L.c.y;
  P.D[2] = L.c.z;
  P.stepLeft = L.c.w;
  const uint32_t f = L.d.x;
  R.block = f & 0xFFFFFFu;
  R.have = (int)((f >> 24) & 7u);
  R.b0 = L.d.y;
  R.b1 = L.d.z;
  R.b2 = L.d.w;
  P.insideVolume = (f >> 28) & 1u;
  P.alive = (f >> 30) & 1u;
  P.sameVoxel = false;
  P.weight = 1.f;
  P.kec = 0;
  P.substeps = 0;
  P.scatterings = P.refractions = P.reflections = P.transitions = 0;
  return (f >> 29) & 1u;
}

// A packet that left the cuboid or died on a boundary, for the tally kernel
// (light collector image, far field, edge fluxes): position after
// checkEscape, direction, weight, time, refractive index, whether it
// escaped / adds to the edge fluxes, and its work item in a batched launch.
struct EscapeRecord {
  float4 a, b, c;
};

MCM_FN void pack_escape(const Photon& P, bool escaped, bool edgeFlux, int item, EscapeRecord& E) {
  E.a = make_float4(P.i[0], P.i[1], P.i[2], P.weight);
  E.b = make_float4(P.u[0], P.u[1], P.u[2], P.time);
  E.c = make_float4(P.ri, (escaped ? 1.f : 0.f) + (edgeFlux ? 2.f : 0.f), (float)item, 0.f);
}

// getNewVoxelProperties (MCmatlablib.c:894-904); the refractive index is only
// changed by refraction.
MCM_FN void new_voxel_properties(Photon& P, const McParams& p, const Medium* media, const uint8_t* cdfIdx) {
  P.j = voxel_of(p, P.i);
  const int m = p.M[P.j];
  const Medium md = media[m];
  P.mua = md.mua;
  P.mus = md.mus;
  P.g = md.g;
  P.cdf = cdfIdx[m];
}

// Movement half of propagatePhoton (MCmatlablib.c:1022-1047, :1086-1093):
// step to the next boundary or interaction, absorb continuously.
template <bool FRESNEL>
MCM_FN MoveResult move(Photon& P, const McParams& p, const Medium* media) {
  MoveResult r;
  P.sameVoxel = true;
  ++P.substeps;
  const float q = P.stepLeft / P.mus;
  const float s = fminf(q, fminf(P.D[0], fminf(P.D[1], P.D[2])));
  P.stepLeft = s == q ? 0.f : P.stepLeft - s * P.mus;
  if (p.needTime) P.time += s * P.ri / kC;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float fi = floorf(P.i[a]);
    if (s == P.D[a]) {
      P.i[a] = P.u[a] > 0.f ? fi + 1.f : fi - kEps * (fabsf(fi) + 1.f);
      P.D[a] = P.Dfull[a];
      P.sameVoxel = false;
    } else {
      float ni = P.i[a] + s * P.u[a] * p.inv_d[a];
      if (floorf(ni) != fi) ni = P.u[a] > 0.f ? fi + 1.f - kEps * (fabsf(fi) + 1.f) : fi;
      P.i[a] = ni;
      P.D[a] -= s;
    }
  }
  r.fresnel = false;
  r.jNew = P.j;
  r.mNew = -1;
  if (!P.sameVoxel && (FRESNEL || p.criteriaActive)) {
    r.jNew = voxel_of(p, P.i);
    r.mNew = p.M[r.jNew];
    const float riNew = media[r.mNew].ri;
    if (FRESNEL && P.ri != riNew) {
      r.fresnel = true;
    } else if (p.criteriaActive) {
      const int mOld = p.M[P.j];
      if (r.mNew != mOld && media[mOld].ri == P.ri && in_count_range(r.mNew, p)) ++P.transitions;
    }
  }
  r.s = s;
  r.absorb = -P.weight * expm1f(-P.mua * s);
  P.weight -= r.absorb;
  return r;
}

// Interface half of propagatePhoton (MCmatlablib.c:1047-1085) with the
// interpolated normal of getInterpolatedNormal (:920-1014).  The reference's
// interpolation cell is always voxels (ic-1, ic) per axis with weight 1/2
// (it floors an integer minus 0.5), so the normal is the mean of the unit
// normals of the up-to-8 voxels around the lower corner of the clamped voxel
// that share the target voxel's refractive index.
MCM_FN void fresnel(Photon& P, const McParams& p, const Medium* media, int jNew, int mNew, PacketRng& R) {
  const float riNew = media[mNew].ri;
  const float mu = P.ri / riNew;
  int lo[3];
  for (int a = 0; a < 3; ++a) lo[a] = clamp_axis(P.i[a], p.n[a]) - 1;
  float nx = 0.f, ny = 0.f, nz = 0.f;
  const int nxy = p.n[0] * p.n[1];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int cx = (c >> 2) & 1, cy = (c >> 1) & 1, cz = c & 1;
    const int x = lo[0] + cx, y = lo[1] + cy, z = lo[2] + cz;
    if (x < 0 || y < 0 || z < 0 || x > p.n[0] - 1 || y > p.n[1] - 1 || z > p.n[2] - 1) continue;
    const float4 nr = p.normalRI[x + y * p.n[0] + z * nxy];
    if (nr.w != riNew) continue;
    nx += 0.125f * nr.x;
    ny += 0.125f * nr.y;
    nz += 0.125f * nr.z;
  }
  const float inv = 1.f / sqrtf(nx * nx + ny * ny + nz * nz);
  nx *= inv;
  ny *= inv;
  nz *= inv;
  const float cin = nx * P.u[0] + ny * P.u[1] + nz * P.u[2];
  if (!(cin > 0.f)) return;  // normal says the packet is heading away from the medium
  const float cos2 = 1.f - mu * mu * (1.f - cin * cin);
  bool reflect = true;
  if (cos2 > 0.f) {
    const float cout = sqrtf(cos2);
    const float rs = (mu * cin - cout) / (mu * cin + cout);
    const float rp = (mu * cout - cin) / (mu * cout + cin);
    const float Rf = 0.5f * rs * rs + 0.5f * rp * rp;
    reflect = R.next() <= Rf;
  }
  if (reflect) {
    P.u[0] -= 2.f * nx * cin;
    P.u[1] -= 2.f * ny * cin;
    P.u[2] -= 2.f * nz * cin;
  } else {
    const float nc = sqrtf(cos2) - mu * cin;
    P.u[0] = nc * nx + mu * P.u[0];
    P.u[1] = nc * ny + mu * P.u[1];
    P.u[2] = nc * nz + mu * P.u[2];
    P.ri = riNew;
  }
  // counters updated without branches (branch-local increments of different
  // fields get merged into a store through a selected pointer)
  const uint32_t counted = (p.criteriaActive && in_count_range(mNew, p)) ? 1u : 0u;
  P.reflections += reflect ? counted : 0u;
  P.refractions += reflect ? 0u : counted;
  P.transitions += reflect ? 0u : counted;
  for (int a = 0; a < 3; ++a) P.D[a] = boundary_distance(P.i[a], P.u[a], p.d[a]);
  set_crossing_lengths(P, p);
}

// xyztoXYZ (MCmatlablib.c:239-246) with precomputed sin/cos
MCM_FN void to_lc_frame(const McParams& p, const float* r, float* o) {
  o[0] = p.lcSinP * r[0] - p.lcCosP * r[1];
  o[1] = p.lcCosT * p.lcCosP * r[0] + p.lcCosT * p.lcSinP * r[1] - p.lcSinT * r[2];
  o[2] = p.lcSinT * p.lcCosP * r[0] + p.lcSinT * p.lcSinP * r[1] + p.lcCosT * r[2];
}

// Light collector hit test (formImage, MCmatlablib.c:737-788).  Marks the
// packet collected and returns the image bin, or -1.
MCM_FN int collector_bin(Photon& P, const McParams& p) {
  float U[3];
  to_lc_frame(p, P.u, U);
  if (!(U[2] < 0.f)) return -1;
  const float hx = (float)p.n[0] * 0.5f, hy = (float)p.n[1] * 0.5f;
  const float resc[3] = {(P.i[0] - hx) * p.d[0] - p.lcR[0], (P.i[1] - hy) * p.d[1] - p.lcR[1], P.i[2] * p.d[2] - p.lcR[2]};
  float Re[3];
  to_lc_frame(p, resc, Re);
  if (!(Re[2] > 0.f)) return -1;
  const float X = Re[0] - Re[2] * U[0] / U[2], Y = Re[1] - Re[2] * U[1] / U[2];
  if (!(sqrtf(X * X + Y * Y) < p.lcDiam / 2.f)) return -1;
  int ti = 0;
  if (p.lcFinite) {
    const float IX = X + p.lcF * U[0] / U[2], IY = Y + p.lcF * U[1] / U[2];
    if (!(sqrtf(IX * IX + IY * IY) < p.lcFSorNA / 2.f)) return -1;
    const int xi = clampi((int)((float)p.lcRes0 * (IX / p.lcFSorNA + 0.5f)), 0, p.lcRes0 - 1);
    const int yi = clampi((int)((float)p.lcRes0 * (IY / p.lcFSorNA + 0.5f)), 0, p.lcRes0 - 1);
    if (p.lcRes1 > 1) {
      const float t = 1.f + (float)(p.lcRes1 - 2) * (P.time - (Re[2] - p.lcF) / U[2] * P.ri / kC - p.lcTStart) / (p.lcTEnd - p.lcTStart);
      ti = t < 0.f ? 0 : (t > (float)(p.lcRes1 - 1) ? p.lcRes1 - 1 : (int)t);
    }
    P.kec = 2;
    return xi + yi * p.lcRes0 + ti * p.lcRes0 * p.lcRes0;
  }
  const float th = atanf(-sqrtf(U[0] * U[0] + U[1] * U[1]) / U[2]);
  if (!(th < p.lcNAangle)) return -1;
  if (p.lcRes1 > 1) {
    const float t = 1.f + (float)(p.lcRes1 - 2) * (P.time - Re[2] / U[2] * P.ri / kC - p.lcTStart) / (p.lcTEnd - p.lcTStart);
    ti = t < 0.f ? 0 : (t > (float)(p.lcRes1 - 1) ? p.lcRes1 - 1 : (int)t);
  }
  P.kec = 2;
  return ti;
}

// formFarField bin (MCmatlablib.c:793-797)
MCM_FN int far_field_bin(const Photon& P, const McParams& p) {
  const float th = (1.f - kEps) * acosf(fminf(1.f, fmaxf(-1.f, P.u[2])));
  const float ph = (1.f - kEps) * (kPi + atan2f(P.u[1], P.u[0]));
  const int r = p.farFieldRes;
  const int a = clampi((int)floorf(th / kPi * (float)r), 0, r - 1);
  const int b = clampi((int)floorf(ph / (2.f * kPi) * (float)r), 0, r - 1);
  return a + r * b;
}

// formEdgeFluxes target (MCmatlablib.c:802-816): face 0..5 = xpos, xneg,
// ypos, yneg, zpos, zneg, and the index in that face's array; face -1 = none.
MCM_FN int edge_flux_bin(const Photon& P, const McParams& p, int* face) {
  const int nx = p.n[0], ny = p.n[1], nz = p.n[2];
  *face = -1;
  if (p.boundaryType == 1) {
    if (P.i[2] < 0.f) { *face = 5; return clampi((int)P.i[0], 0, nx - 1) + nx * clampi((int)P.i[1], 0, ny - 1); }
    if (P.i[2] >= (float)nz) { *face = 4; return clampi((int)P.i[0], 0, nx - 1) + nx * clampi((int)P.i[1], 0, ny - 1); }
    if (P.i[1] < 0.f) { *face = 3; return clampi((int)P.i[0], 0, nx - 1) + nx * clampi((int)P.i[2], 0, nz - 1); }
    if (P.i[1] >= (float)ny) { *face = 2; return clampi((int)P.i[0], 0, nx - 1) + nx * clampi((int)P.i[2], 0, nz - 1); }
    if (P.i[0] < 0.f) { *face = 1; return clampi((int)P.i[1], 0, ny - 1) + ny * clampi((int)P.i[2], 0, nz - 1); }
    if (P.i[0] >= (float)nx) { *face = 0; return clampi((int)P.i[1], 0, ny - 1) + ny * clampi((int)P.i[2], 0, nz - 1); }
    return 0;
  }
  if (p.boundaryType == 2) {
    if (P.i[2] < 0.f) {
      *face = 5;
      const int X = clampi((int)(P.i[0] + (float)(nx * (kKillRange - 1)) / 2.0f), 0, kKillRange * nx - 1);
      const int Y = clampi((int)(P.i[1] + (float)(ny * (kKillRange - 1)) / 2.0f), 0, kKillRange * ny - 1);
      return X + kKillRange * nx * Y;
    }
    return 0;
  }
  if (P.i[2] < 0.f) { *face = 5; return clampi((int)P.i[0], 0, nx - 1) + nx * clampi((int)P.i[1], 0, ny - 1); }
  if (P.i[2] >= (float)nz) { *face = 4; return clampi((int)P.i[0], 0, nx - 1) + nx * clampi((int)P.i[1], 0, ny - 1); }
  return 0;
}

// checkEscape (MCmatlablib.c:821-889) minus the tally writes: updates
// alive/insideVolume, wraps cyclic x/y, and reports what to tally.
struct EscapeResult {
  bool escaped;    // left through an escaping face into n = 1
  bool edgeFlux;   // died on a boundary: add weight to the NI_* face array
};

MCM_FN EscapeResult check_escape(Photon& P, const McParams& p) {
  EscapeResult e;
  e.escaped = false;
  switch (p.boundaryType) {
    case 0:
      P.alive = inside_killrange(p, P.i, false);
      break;
    case 1:
      P.alive = inside_box(p, P.i);
      e.escaped = !P.alive && P.ri == 1.f;
      break;
    case 2:
      P.alive = inside_killrange(p, P.i, true);
      e.escaped = P.i[2] < 0.f && P.ri == 1.f;
      break;
    default: {
      const bool tooOld = P.time > p.tooOld;
      P.alive = P.i[2] < (float)p.n[2] && P.i[2] >= 0.f && !tooOld;
      e.escaped = !P.alive && P.ri == 1.f && !tooOld;
      if (P.i[0] >= (float)p.n[0]) P.i[0] = 0.f;
      if (P.i[0] < 0.f) P.i[0] = (float)p.n[0] * (1.f - kEps);
      if (P.i[1] >= (float)p.n[1]) P.i[1] = 0.f;
      if (P.i[1] < 0.f) P.i[1] = (float)p.n[1] * (1.f - kEps);
      break;
    }
  }
  P.insideVolume = inside_box(p, P.i);
  if (e.escaped) P.kec = 1;
  e.edgeFlux = !P.alive && p.boundaryType != 0;
  return e;
}

// checkRoulette (MCmatlablib.c:1124-1133)
MCM_FN void roulette(Photon& P, PacketRng& R) {
  if (P.weight < kThreshold) {
    if (R.next() <= kChance) P.weight /= kChance;
    else P.alive = false;
  }
}

// scatterPhoton (MCmatlablib.c:1138-1184)
MCM_FN void scatter(Photon& P, const McParams& p, const float* cdfs, int mCur, PacketRng& R) {
  float ct;
  if (isnan(P.g)) {
    const float* D = cdfs + P.cdf * kCdfSize;
    const int jt = table_search(R.next(), kCdfSize - 1, D);
    ct = cosf(((float)jt + R.next()) * kPi / (float)(kCdfSize - 1));
  } else {
    const float g = P.g;
    if (fabsf(g) == 1.f) ct = g;
    else if (fabsf(g) <= 3.4526698e-04f) ct = 2.f * R.next() - 1.f;  // sqrt(FLT_EPSILON)
    else {
      const float t = (1.f - g * g) / (1.f - g + 2.f * g * R.next());
      ct = (1.f + g * g - t * t) / (2.f * g);
    }
  }
  const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
  float sp, cp;
  sincosf(2.f * kPi * R.next(), &sp, &cp);
  const float rho2 = P.u[0] * P.u[0] + P.u[1] * P.u[1];
  if (fabsf(P.u[2]) < 1.f && rho2 > 0.f) {
    const float rho = sqrtf(rho2);
    const float ir = 1.f / rho;
    const float ux = st * (P.u[0] * P.u[2] * cp - P.u[1] * sp) * ir + P.u[0] * ct;
    const float uy = st * (P.u[1] * P.u[2] * cp + P.u[0] * sp) * ir + P.u[1] * ct;
    P.u[2] = -st * cp * rho + P.u[2] * ct;
    P.u[1] = uy;
    P.u[0] = ux;
  } else {
    P.u[0] = st * cp;
    P.u[1] = st * sp;
    P.u[2] = ct * (P.u[2] >= 0.f ? 1.f : -1.f);
  }
  for (int a = 0; a < 3; ++a) P.D[a] = boundary_distance(P.i[a], P.u[a], p.d[a]);
  set_crossing_lengths(P, p);
  P.stepLeft = -logf(R.next());
  if (p.criteriaActive && in_count_range(mCur, p)) ++P.scatterings;
}

}  // namespace mcm
