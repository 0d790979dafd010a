// Heat diffusion + Arrhenius damage on gfx950: kernels and the C ABI
// (mcm_heat_plan_*, mcm_heat_run).  Replaces finiteElementHeatPropagator.c
// (+MCmatlab/src/finiteElementHeatPropagator.c:688-821).
//
// One Douglas-Gunn ADI step (:767-782) is three line-sweep kernels instead of
// the reference's six loops:
//   x sweep   substeps 1a + 1b: the explicit right-hand side (full 7-point
//             stencil, x half-weighted, plus the source) built on the fly and
//             solved along x;
//   y sweep   substeps 2a + 2b: the -dt/2 L_y correction and the y solve;
//   z sweep   substeps 3a + 3b: the -dt/2 L_z correction, the z solve, the
//             Arrhenius damage update and the per-medium maximum temperature.
// FP64 (default), lines of up to 512 voxels: each line is solved by one
// wavefront (partition method + parallel cyclic reduction across the 64
// lanes); the x sweep reads its lines directly (contiguous), the y and z
// sweeps stage tiles of lines through LDS.  FP32 (the reference's arithmetic)
// and longer lines: one lane per line running the Thomas recurrence over
// blocks of loaded elements.
// The explicit (forward Euler) scheme is one kernel per step.
//
// Arithmetic follows the reference's expressions (built without FMA
// contraction); in FP32 it is the reference's operation for operation.
// Per-medium maxima use an order-preserving integer key of the FP64 value:
// each lane keeps a running maximum over its current run of one medium and
// folds it into LDS when the medium changes; each workgroup folds LDS into
// the global keys at the end.
// The discrete adjoint of the FP64 DG-ADI step (transposed line solves and
// stencils) follows the forward kernels.
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "mcm_common.h"

namespace mcm {

constexpr int kHeatTPB = 256;

__host__ __device__ inline unsigned long long order_key(double v) {
  unsigned long long b;
  memcpy(&b, &v, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double key_value(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  double v;
  memcpy(&v, &b, 8);
  return v;
}

template <typename RT>
struct HeatArgs {
  int nx, ny, nz, nM;
  int sinked, lightsOn, calcDamage;
  RT dt;
  RT Rgas, czero;
  const uint8_t* M;
  const float* Q;        // dTdt_abs
  const RT* c;           // [nM * nM * 3]: c[a + b nM + d nM^2]
  const RT* logA;
  const RT* E;
  const RT* T1;          // temperature at the start of the step
  RT* T2;                // result (and intermediate) of the step
  RT* B;                 // Thomas scratch: forward-eliminated diagonal
  RT* Om;                // damage integral (calcDamage)
  unsigned long long* maxKey;  // [nM] order_key of the running maximum
  // sensor record of T1 (the state before the step), written by the first
  // workgroup of the step's first kernel: column sCol (+ *sColBase)
  int nS;
  const long long* sIdx;
  const RT* sW;
  double* sOut;
  const long long* sColBase;
  long long sCol;
};

// recordTempSensors (:652-663): trilinear interpolation of T into column
// `col` of the sensor record, sensors s = first, first + step, ...
template <typename RT>
__device__ inline void sensors_at(const RT* T, int nx, int nxy, const long long* idx, const RT* w, int nS, double* out, long long col,
                                  int first, int step) {
  for (int s = first; s < nS; s += step) {
    const long long k = idx[s];
    const RT wx = w[s], wy = w[s + nS], wz = w[s + 2 * nS];
    out[s + col * nS] = (double)((1 - wx) * (1 - wy) * (1 - wz) * T[k] + (1 - wx) * (1 - wy) * wz * T[k + nxy] +
                                 (1 - wx) * wy * (1 - wz) * T[k + nx] + (1 - wx) * wy * wz * T[k + nx + nxy] +
                                 wx * (1 - wy) * (1 - wz) * T[k + 1] + wx * (1 - wy) * wz * T[k + 1 + nxy] +
                                 wx * wy * (1 - wz) * T[k + 1 + nx] + wx * wy * wz * T[k + 1 + nx + nxy]);
  }
}

// The step kernels' share of the sensor record (first workgroup only).
template <typename RT>
__device__ inline void step_sensors(const HeatArgs<RT>& a) {
  if (a.nS == 0 || blockIdx.x != 0) return;
  const long long col = a.sCol + (a.sColBase ? *a.sColBase : 0);
  sensors_at(a.T1, a.nx, a.nx * a.ny, a.sIdx, a.sW, a.nS, a.sOut, col, (int)threadIdx.x, (int)blockDim.x);
}

template <typename RT>
__device__ inline bool frozen(const HeatArgs<RT>& a, int x, int y, int z) {
  return a.sinked && (x == 0 || x == a.nx - 1 || y == 0 || y == a.ny - 1 || z == 0 || z == a.nz - 1);
}

template <typename RT>
__device__ inline RT cc(const HeatArgs<RT>& a, int m, int n, int d) {
  return a.c[m + n * a.nM + d * a.nM * a.nM];
}

struct MaxRun {
  int m = -1;
  double v = -INFINITY;
  __device__ void add(int mm, double t, unsigned long long* s) {
    if (mm != m) {
      if (m >= 0) atomicMax(&s[m], order_key(v));
      m = mm;
      v = t;
    } else if (t > v) {
      v = t;
    }
  }
  __device__ void flush(unsigned long long* s) {
    if (m >= 0) atomicMax(&s[m], order_key(v));
    m = -1;
  }
};

__device__ inline void init_lds_max(unsigned long long* s, int nM) {
  for (int k = threadIdx.x; k < nM; k += blockDim.x) s[k] = order_key(-INFINITY);
  __syncthreads();
}
__device__ inline void fold_lds_max(unsigned long long* s, unsigned long long* g, int nM) {
  __syncthreads();
  const unsigned long long none = order_key(-INFINITY);
  for (int k = threadIdx.x; k < nM; k += blockDim.x)
    if (s[k] != none) atomicMax(&g[k], s[k]);
}

enum : int { SWEEP_X = 0, SWEEP_Y = 1, SWEEP_Z = 2 };
constexpr int kLdsCoef = 1536;  // coefficient tables up to 22 media live in LDS

template <typename RT>
__device__ inline RT coef(const RT* c, int nM, int m, int nb, int d) {
  return c[m + nb * nM + d * nM * nM];
}

// Per-workgroup copies of the small per-medium tables.
template <typename RT>
struct HeatLds {
  RT c[kLdsCoef];
  RT logA[256];
  RT E[256];
  unsigned long long maxKey[256];
};

template <typename RT>
__device__ inline const RT* stage_tables(const HeatArgs<RT>& a, HeatLds<RT>& s) {
  const int nc = a.nM * a.nM * 3;
  if (nc <= kLdsCoef)
    for (int k = threadIdx.x; k < nc; k += blockDim.x) s.c[k] = a.c[k];
  for (int k = threadIdx.x; k < a.nM; k += blockDim.x) {
    s.logA[k] = a.logA[k];
    s.E[k] = a.E[k];
    s.maxKey[k] = order_key(-INFINITY);
  }
  __syncthreads();
  return nc <= kLdsCoef ? s.c : a.c;
}

// One grid line: the explicit part of the substep (KIND X: the full
// right-hand side of substep 1a; Y/Z: the -dt/2 L_d T correction of 2a/3a),
// Thomas forward elimination, back substitution, and for Z the Arrhenius
// update and the running maxima.  Expressions and their order are those of
// the reference (finiteElementHeatPropagator.c:288-646); only the memory
// access is reorganised: each lane loads a block of U consecutive elements
// of its line (and their stencil neighbours) before running the sequential
// recurrence over them, so a wavefront has U loads in flight instead of one.
//   i0, st, n : first voxel, stride and length of the line (axis d)
//   lf        : the whole line lies on a fixed-temperature boundary
//   s1/k1/n1, s2/k2/n2 : stride, coordinate and extent of the two other axes
template <typename RT, int KIND, int U>
__device__ void sweep_line(const HeatArgs<RT>& a, const RT* cT, const HeatLds<RT>& L, int i0, int st, int n, bool lf, int d, int s1,
                           int k1, int n1, int s2, int k2, int n2, MaxRun* run, unsigned long long* sMax) {
  const int nM = a.nM;
  const RT dt = a.dt;
  const bool sk = a.sinked;
  RT tprev = 0, bprev = 1;
  for (int k0 = 0; k0 < n; k0 += U) {
    RT Tw[U + 2];
    int Mw[U + 2];
    RT R0[U], Ta[U], Tb[U], Tc[U], Td[U];
    int Ma[U], Mb[U], Mc[U], Md[U];
    float Qv[U];
#pragma unroll
    for (int u = 0; u < U + 2; ++u) {
      const int k = k0 - 1 + u;
      const bool ok = k >= 0 && k < n;
      const int i = i0 + k * st;
      Tw[u] = ok ? a.T1[i] : (RT)0;
      Mw[u] = ok ? (int)a.M[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      const bool ok = k < n;
      const int i = i0 + k * st;
      if (KIND == SWEEP_X) {
        const bool a1 = ok && k1 > 0, b1 = ok && k1 < n1 - 1, a2 = ok && k2 > 0, b2 = ok && k2 < n2 - 1;
        Ta[u] = a1 ? a.T1[i - s1] : (RT)0;
        Ma[u] = a1 ? (int)a.M[i - s1] : 0;
        Tb[u] = b1 ? a.T1[i + s1] : (RT)0;
        Mb[u] = b1 ? (int)a.M[i + s1] : 0;
        Tc[u] = a2 ? a.T1[i - s2] : (RT)0;
        Mc[u] = a2 ? (int)a.M[i - s2] : 0;
        Td[u] = b2 ? a.T1[i + s2] : (RT)0;
        Md[u] = b2 ? (int)a.M[i + s2] : 0;
        Qv[u] = ok ? a.Q[i] : 0.f;
      } else {
        R0[u] = ok ? a.T2[i] : (RT)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (k >= n) break;
      const int i = i0 + k * st;
      const int m = Mw[u + 1];
      const bool fz = lf || (sk && (k == 0 || k == n - 1));
      RT t;
      if (KIND == SWEEP_X) {
        RT dTdt = 0;
        if (!fz) {
          dTdt = a.lightsOn ? (RT)Qv[u] : (RT)0;
          if (k != 0) dTdt += (Tw[u] - Tw[u + 1]) * coef(cT, nM, m, Mw[u], 0) / 2;
          if (k != n - 1) dTdt += (Tw[u + 2] - Tw[u + 1]) * coef(cT, nM, m, Mw[u + 2], 0) / 2;
          if (k1 != 0) dTdt += (Ta[u] - Tw[u + 1]) * coef(cT, nM, m, Ma[u], 1);
          if (k1 != n1 - 1) dTdt += (Tb[u] - Tw[u + 1]) * coef(cT, nM, m, Mb[u], 1);
          if (k2 != 0) dTdt += (Tc[u] - Tw[u + 1]) * coef(cT, nM, m, Mc[u], 2);
          if (k2 != n2 - 1) dTdt += (Td[u] - Tw[u + 1]) * coef(cT, nM, m, Md[u], 2);
        }
        t = Tw[u + 1] + dt * dTdt;
      } else {
        t = R0[u];
        if (!fz) {
          RT dTdt = 0;
          if (k != 0) dTdt -= (Tw[u] - Tw[u + 1]) * coef(cT, nM, m, Mw[u], d) / 2;
          if (k != n - 1) dTdt -= (Tw[u + 2] - Tw[u + 1]) * coef(cT, nM, m, Mw[u + 2], d) / 2;
          t = t + dt * dTdt;
        }
      }
      RT b = 1;
      if (!fz) {
        if (k < n - 1) b += dt / 2 * coef(cT, nM, m, Mw[u + 2], d);
        if (k > 0) {
          const int mp = Mw[u];
          b += dt / 2 * coef(cT, nM, m, mp, d);
          const RT w = -dt / 2 * coef(cT, nM, m, mp, d) / bprev;
          if (k > 1 || !sk) b += w * dt / 2 * coef(cT, nM, mp, m, d);
          t -= w * tprev;
        }
      }
      a.T2[i] = t;
      a.B[i] = b;
      tprev = t;
      bprev = b;
    }
  }
  RT xnext = 0;
  for (int kt = n - 1; kt >= 0; kt -= U) {
    RT Dv[U], Bv[U], T1v[U], Omv[U];
    int Mi[U], Mn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kt - u;
      const bool ok = k >= 0;
      const int i = i0 + k * st;
      Dv[u] = ok ? a.T2[i] : (RT)0;
      Bv[u] = ok ? a.B[i] : (RT)1;
      Mi[u] = ok ? (int)a.M[i] : 0;
      Mn[u] = ok && k < n - 1 ? (int)a.M[i + st] : 0;
      if (KIND == SWEEP_Z) {
        T1v[u] = ok ? a.T1[i] : (RT)0;
        Omv[u] = ok && a.calcDamage ? a.Om[i] : (RT)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kt - u;
      if (k < 0) break;
      const int i = i0 + k * st;
      const int m = Mi[u];
      const bool fz = lf || (sk && (k == 0 || k == n - 1));
      RT t = Dv[u];
      if (!fz) {
        t = (t + (k == n - 1 ? (RT)0 : dt / 2 * coef(cT, nM, m, Mn[u], d) * xnext)) / Bv[u];
        a.T2[i] = t;
      }
      xnext = t;
      if (KIND == SWEEP_Z) {
        if (a.calcDamage && L.E[m] != 0) a.Om[i] = Omv[u] + dt * (RT)exp(L.logA[m] - L.E[m] / (a.Rgas * ((t + T1v[u]) / 2 + a.czero)));
        run->add(m, (double)t, sMax);
      }
    }
  }
}

constexpr int kSweepU = 8;

// substeps 1a + 1b (finiteElementHeatPropagator.c:288-312, :352-389): lane per (y, z) x-line
template <typename RT, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_x(HeatArgs<RT> a) {
  __shared__ HeatLds<RT> L;
  step_sensors(a);
  (void)stage_tables(a, L);
  const RT* cT = LDS_COEF ? L.c : a.c;  // LDS or global, never a generic pointer
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= a.ny * a.nz) return;
  const int y = line % a.ny, z = line / a.ny;
  const bool lf = a.sinked && (y == 0 || y == a.ny - 1 || z == 0 || z == a.nz - 1);
  sweep_line<RT, SWEEP_X, kSweepU>(a, cT, L, y * a.nx + z * a.nx * a.ny, 1, a.nx, lf, 0, a.nx, y, a.ny, a.nx * a.ny, z, a.nz, nullptr,
                                   nullptr);
}

// substeps 2a + 2b (:429-448, :486-521): lane per (x, z) y-line, lanes along x
template <typename RT, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_y(HeatArgs<RT> a) {
  __shared__ HeatLds<RT> L;
  (void)stage_tables(a, L);
  const RT* cT = LDS_COEF ? L.c : a.c;  // LDS or global, never a generic pointer
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= a.nx * a.nz) return;
  const int x = line % a.nx, z = line / a.nx;
  const bool lf = a.sinked && (x == 0 || x == a.nx - 1 || z == 0 || z == a.nz - 1);
  sweep_line<RT, SWEEP_Y, kSweepU>(a, cT, L, x + z * a.nx * a.ny, a.nx, a.ny, lf, 1, 0, 0, 1, 0, 0, 1, nullptr, nullptr);
}

// substeps 3a + 3b (:546-565, :606-646) with the Arrhenius update and the
// per-medium maximum: lane per (x, y) z-line, lanes along x
template <typename RT, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_z(HeatArgs<RT> a) {
  __shared__ HeatLds<RT> L;
  (void)stage_tables(a, L);
  const RT* cT = LDS_COEF ? L.c : a.c;  // LDS or global, never a generic pointer
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line < a.nx * a.ny) {
    const int x = line % a.nx, y = line / a.nx;
    const bool lf = a.sinked && (x == 0 || x == a.nx - 1 || y == 0 || y == a.ny - 1);
    MaxRun run;
    sweep_line<RT, SWEEP_Z, kSweepU>(a, cT, L, line, a.nx * a.ny, a.nz, lf, 2, 0, 0, 1, 0, 0, 1, &run, L.maxKey);
    run.flush(L.maxKey);
  }
  fold_lds_max(L.maxKey, a.maxKey, a.nM);
}

// ---------------------------------------------------------------------------
// Wavefront-parallel tridiagonal solve (FP64 sweeps)
//
// A system  a[r] x[r-1] + b[r] x[r] + c[r] x[r+1] = d[r]  of up to W*S rows
// is spread over a segment of W lanes (W = 64, 32 or 16; a wavefront solves
// 64 / W systems side by side), S consecutive rows per lane (row
// r = lane * S + s, lane counted within the segment); unused rows are
// a = c = 0, b = 1, d = 0.  The matrices of the heat step are strictly
// diagonally dominant, so no pivoting is needed.  Partition method:
//   1. each lane eliminates its interior rows 1..S-2, expressing them through
//      its own first and last unknowns;
//   2. the first unknown of each lane is eliminated as well, leaving one
//      tridiagonal equation per lane in the lanes' last unknowns U;
//   3. that W-row system is solved by parallel cyclic reduction, log2(W)
//      steps of lane shuffles that never cross a segment's edge;
//   4. each lane recovers its first and interior unknowns from U.
// Same solution as the sequential Thomas recurrence up to rounding.
__device__ inline double lane_shfl(double v, int src) { return __shfl(v, src < 0 ? 0 : (src > 63 ? 63 : src), 64); }
__device__ inline int lane_shfl(int v, int src) { return __shfl(v, src < 0 ? 0 : (src > 63 ? 63 : src), 64); }

// DPP move within rows of 16 lanes (CTRL 0x100 + h: row_shl:h, lane p gets
// lane p + h; 0x110 + h: row_shr:h, lane p gets lane p - h; 0 where the
// source lies outside the row)
template <int CTRL>
__device__ inline int dpp_move(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true); }
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<CTRL>((int)b), hi = dpp_move<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// v of lane wl + h (UP) or wl - h of the same W-lane segment; a lane whose
// source lies outside its segment gets an unspecified value, which callers
// mask.  16-lane segments are the DPP unit's rows, so their shifts are VALU
// row moves rather than LDS-routed permutes.
template <int W, bool UP, typename V>
__device__ inline V seg_shift(V v, int h) {
  if constexpr (W == 16) {
    constexpr int base = UP ? 0x100 : 0x110;
    switch (h) {
      case 1: return dpp_move<base + 1>(v);
      case 2: return dpp_move<base + 2>(v);
      case 4: return dpp_move<base + 4>(v);
      case 8: return dpp_move<base + 8>(v);
      default: break;
    }
  }
  const int wl = __lane_id();
  return lane_shfl(v, UP ? wl + h : wl - h);
}

// a / b for the positive, well-scaled pivots of the solver: hardware
// reciprocal estimate refined by two Newton steps (a few ulp), instead of the
// correctly rounded division sequence.
__device__ inline double pdiv(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  return a * r;
}

template <int S, int W = 64>
__device__ inline void wave_tridiag(const double (&a)[S], const double (&b)[S], const double (&c)[S], const double (&d)[S],
                                    double (&x)[S]) {
  // a line per W-lane segment of the wavefront; lane is the index within it
  static_assert(W == 16 || W == 32 || W == 64, "segments of 16, 32 or 64 lanes");
  const int wl = __lane_id(), lane = wl & (W - 1);
  double al[S], be[S], ga[S], de[S];
  double A1 = a[0], B1 = b[0], C1 = c[0], D1 = d[0];
  double A2 = a[S - 1], B2 = b[S - 1], C2 = c[S - 1], D2 = d[S - 1];
  if constexpr (S >= 3) {
    // forward elimination of the interior, first unknown kept symbolic
    al[1] = a[1];
    be[1] = b[1];
    ga[1] = c[1];
    de[1] = d[1];
#pragma unroll
    for (int s = 2; s <= S - 2; ++s) {
      const double r = pdiv(a[s], be[s - 1]);
      al[s] = -r * al[s - 1];
      be[s] = b[s] - r * ga[s - 1];
      ga[s] = c[s];
      de[s] = d[s] - r * de[s - 1];
    }
    // backward elimination, last unknown kept symbolic
#pragma unroll
    for (int s = S - 3; s >= 1; --s) {
      const double r = pdiv(ga[s], be[s + 1]);
      al[s] -= r * al[s + 1];
      ga[s] = -r * ga[s + 1];
      de[s] -= r * de[s + 1];
    }
    const double r1 = pdiv(c[0], be[1]);
    B1 = b[0] - r1 * al[1];
    C1 = -r1 * ga[1];
    D1 = d[0] - r1 * de[1];
    const double r2 = pdiv(a[S - 1], be[S - 2]);
    A2 = -r2 * al[S - 2];
    B2 = b[S - 1] - r2 * ga[S - 2];
    D2 = d[S - 1] - r2 * de[S - 2];
  }
  // one equation per lane in U = last unknowns:  pa U[l-1] + pb U[l] + pc U[l+1] = pd
  double pa, pb, pc, pd;
  if constexpr (S == 1) {
    pa = A1;
    pb = B1;
    pc = C1;
    pd = D1;
  } else {
    // shuffles run in every lane (a lane reading from a lane that is
    // inactive at the shuffle gets 0), then the edge lanes are masked
    const bool hasNext = lane < W - 1;
    const double sA = seg_shift<W, true>(A1, 1), sB = seg_shift<W, true>(B1, 1);
    const double sC = seg_shift<W, true>(C1, 1), sD = seg_shift<W, true>(D1, 1);
    const double nA = hasNext ? sA : 0.0, nB = hasNext ? sB : 1.0, nC = hasNext ? sC : 0.0, nD = hasNext ? sD : 0.0;
    const double iB1 = pdiv(1.0, B1), inB = pdiv(1.0, nB);
    pa = -A2 * A1 * iB1;
    pb = B2 - A2 * C1 * iB1 - C2 * nA * inB;
    pc = -C2 * nC * inB;
    pd = D2 - A2 * D1 * iB1 - C2 * nD * inB;
  }
#pragma unroll
  for (int h = 1; h < W; h <<= 1) {
    const bool lo = lane >= h, hi = lane + h < W;
    const double sam = seg_shift<W, false>(pa, h), sbm = seg_shift<W, false>(pb, h);
    const double scm = seg_shift<W, false>(pc, h), sdm = seg_shift<W, false>(pd, h);
    const double sap = seg_shift<W, true>(pa, h), sbp = seg_shift<W, true>(pb, h);
    const double scp = seg_shift<W, true>(pc, h), sdp = seg_shift<W, true>(pd, h);
    const double am = lo ? sam : 0.0, bm = lo ? sbm : 1.0, cm = lo ? scm : 0.0, dm = lo ? sdm : 0.0;
    const double ap = hi ? sap : 0.0, bp = hi ? sbp : 1.0, cp = hi ? scp : 0.0, dp = hi ? sdp : 0.0;
    const double k1 = pdiv(pa, bm), k2 = pdiv(pc, bp);
    pb = pb - k1 * cm - k2 * ap;
    pd = pd - k1 * dm - k2 * dp;
    pa = -k1 * am;
    pc = -k2 * cp;
  }
  const double U = pdiv(pd, pb);
  if constexpr (S == 1) {
    x[0] = U;
  } else {
    const double sU = seg_shift<W, false>(U, 1);
    const double Up = lane > 0 ? sU : 0.0;
    const double first = pdiv(D1 - A1 * Up - C1 * U, B1);
    x[0] = first;
    x[S - 1] = U;
    if constexpr (S >= 3) {
#pragma unroll
      for (int s = 1; s <= S - 2; ++s) x[s] = pdiv(de[s] - al[s] * first - ga[s] * U, be[s]);
    }
  }
}

// substeps 1a + 1b in FP64 with one wavefront per (y, z) x-line: lane l holds
// voxels x = l*S .. l*S+S-1, so every load of the wavefront is contiguous.
// The right-hand side is the full 7-point explicit stencil (x half-weighted)
// plus the source, in the expression order of the sequential kernel.
template <int S, int W, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_x_wave(HeatArgs<double> a) {
  // the coefficient table in LDS when it fits (LDS_COEF), else read from
  // global memory; one or the other, so every lookup is a plain LDS or
  // global load rather than a generic one
  __shared__ double sc[LDS_COEF ? kLdsCoef : 1];
  step_sensors(a);
  if (LDS_COEF) {
    const int nc = a.nM * a.nM * 3;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) sc[k] = a.c[k];
    __syncthreads();
  }
  const double* __restrict__ cT = LDS_COEF ? sc : a.c;
  // a line per W-lane segment; a segment past the last line solves the
  // last line again and stores nothing
  const int wl = __lane_id(), lane = wl & (W - 1);
  const int line = (blockIdx.x * (kHeatTPB / 64) + (threadIdx.x >> 6)) * (64 / W) + wl / W;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny, nM = a.nM;
  const bool valid = line < ny * nz;
  const int y = (valid ? line : ny * nz - 1) % ny, z = (valid ? line : ny * nz - 1) / ny;
  const int i0 = y * nx + z * nxy;
  const bool sk = a.sinked;
  const bool lf = sk && (y == 0 || y == ny - 1 || z == 0 || z == nz - 1);
  const double* __restrict__ T1 = a.T1;
  const uint8_t* __restrict__ Mg = a.M;
  const float* __restrict__ Q = a.Q;
  const double dt = a.dt;
  double T[S];
  int Mv[S];
  // neighbouring lines and the source, loaded up front (clamped addresses)
  double Tym[S], Typ[S], Tzm[S], Tzp[S];
  int Mym[S], Myp[S], Mzm[S], Mzp[S];
  float Qv[S];
  const int oym = y != 0 ? -nx : 0, oyp = y != ny - 1 ? nx : 0, ozm = z != 0 ? -nxy : 0, ozp = z != nz - 1 ? nxy : 0;
  const bool vec = S % 4 == 0 && (nx & 3) == 0 && lane * S + S <= nx;
  if (vec) {
    // a lane's S voxels of each line are consecutive and 32-byte aligned:
    // 16-byte loads for T and Q, 4-byte loads for M
    const int b = i0 + lane * S;
#pragma unroll
    for (int s = 0; s < S; s += 2) {
      const double2 t = *(const double2*)(T1 + b + s), tym = *(const double2*)(T1 + b + oym + s), typ = *(const double2*)(T1 + b + oyp + s);
      const double2 tzm = *(const double2*)(T1 + b + ozm + s), tzp = *(const double2*)(T1 + b + ozp + s);
      T[s] = t.x, T[s + 1] = t.y;
      Tym[s] = tym.x, Tym[s + 1] = tym.y;
      Typ[s] = typ.x, Typ[s + 1] = typ.y;
      Tzm[s] = tzm.x, Tzm[s + 1] = tzm.y;
      Tzp[s] = tzp.x, Tzp[s + 1] = tzp.y;
    }
#pragma unroll
    for (int s = 0; s < S; s += 4) {
      const float4 q = *(const float4*)(Q + b + s);
      Qv[s] = q.x, Qv[s + 1] = q.y, Qv[s + 2] = q.z, Qv[s + 3] = q.w;
      const uint32_t m = *(const uint32_t*)(Mg + b + s), mym = *(const uint32_t*)(Mg + b + oym + s);
      const uint32_t myp = *(const uint32_t*)(Mg + b + oyp + s), mzm = *(const uint32_t*)(Mg + b + ozm + s);
      const uint32_t mzp = *(const uint32_t*)(Mg + b + ozp + s);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        Mv[s + u] = (int)((m >> (8 * u)) & 0xffu);
        Mym[s + u] = (int)((mym >> (8 * u)) & 0xffu);
        Myp[s + u] = (int)((myp >> (8 * u)) & 0xffu);
        Mzm[s + u] = (int)((mzm >> (8 * u)) & 0xffu);
        Mzp[s + u] = (int)((mzp >> (8 * u)) & 0xffu);
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      const int i = i0 + (k < nx ? k : 0);
      T[s] = k < nx ? T1[i] : 0.0;
      Mv[s] = k < nx ? (int)Mg[i] : 0;
      Tym[s] = T1[i + oym];
      Typ[s] = T1[i + oyp];
      Tzm[s] = T1[i + ozm];
      Tzp[s] = T1[i + ozp];
      Mym[s] = Mg[i + oym];
      Myp[s] = Mg[i + oyp];
      Mzm[s] = Mg[i + ozm];
      Mzp[s] = Mg[i + ozp];
      Qv[s] = Q[i];
    }
  }
  const double Tl = seg_shift<W, false>(T[S - 1], 1), Tr = seg_shift<W, true>(T[0], 1);
  const int Ml = seg_shift<W, false>(Mv[S - 1], 1), Mr = seg_shift<W, true>(Mv[0], 1);
  double ca[S], cb[S], cc_[S], cd[S], xs[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int k = lane * S + s;
    ca[s] = 0.0;
    cb[s] = 1.0;
    cc_[s] = 0.0;
    cd[s] = 0.0;
    if (k >= nx) continue;
    const int m = Mv[s];
    const double TL = s > 0 ? T[s - 1] : Tl, TR = s < S - 1 ? T[s + 1] : Tr;
    const int ML = s > 0 ? Mv[s - 1] : Ml, MR = s < S - 1 ? Mv[s + 1] : Mr;
    const bool fz = lf || (sk && (k == 0 || k == nx - 1));
    if (fz) {
      cd[s] = T[s];
      continue;
    }
    double dTdt = a.lightsOn ? (double)Qv[s] : 0.0;
    if (k != 0) dTdt += (TL - T[s]) * coef(cT, nM, m, ML, 0) / 2;
    if (k != nx - 1) dTdt += (TR - T[s]) * coef(cT, nM, m, MR, 0) / 2;
    if (y != 0) dTdt += (Tym[s] - T[s]) * coef(cT, nM, m, Mym[s], 1);
    if (y != ny - 1) dTdt += (Typ[s] - T[s]) * coef(cT, nM, m, Myp[s], 1);
    if (z != 0) dTdt += (Tzm[s] - T[s]) * coef(cT, nM, m, Mzm[s], 2);
    if (z != nz - 1) dTdt += (Tzp[s] - T[s]) * coef(cT, nM, m, Mzp[s], 2);
    cd[s] = T[s] + dt * dTdt;
    if (k < nx - 1) {
      const double q = dt / 2 * coef(cT, nM, m, MR, 0);
      cb[s] += q;
      cc_[s] = -q;
    }
    if (k > 0) {
      const double q = dt / 2 * coef(cT, nM, m, ML, 0);
      cb[s] += q;
      ca[s] = -q;
    }
  }
  wave_tridiag<S, W>(ca, cb, cc_, cd, xs);
  double* __restrict__ T2 = a.T2;
  if (!valid) return;
  if (vec) {
#pragma unroll
    for (int s = 0; s < S; s += 2) *(double2*)(T2 + i0 + lane * S + s) = make_double2(xs[s], xs[s + 1]);
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      if (k < nx) T2[i0 + k] = xs[s];
    }
  }
}

// substeps 2a + 2b (KIND Y) and 3a + 3b with the Arrhenius update and the
// per-medium maxima (KIND Z) in FP64, lines staged through LDS: a workgroup
// takes TX consecutive x-lines at one z (Y) or y (Z), loads them row by row
// (TX consecutive x per row, coalesced; all of a thread's loads issued before
// its first LDS store) into LDS with each line contiguous, each wavefront
// solves TX/4 lines from LDS with the wavefront solver (lane l holds voxels
// l*S .. l*S+S-1), and the workgroup writes the rows back.  The right-hand
// side is T1 of the previous substep minus dt/2 L_d T(step start) along the
// line, in the expression order of the sequential kernel.
template <int S, int W>
struct YzTile {
  static constexpr int NP = W * S;            // padded line length
  static constexpr int LD = NP + 1;           // LDS pitch of a line (doubles)
  // lines per workgroup: 16 (128-byte rows) up to 128 voxels a line; longer
  // lines would need ~70 KB of LDS per workgroup at 16, leaving only two
  // workgroups per CU, so they take 8
  static constexpr int TX = NP <= 128 ? 16 : 8;
};

// Coefficient tables of up to 8 media (3 * 8 * 8 entries) go to LDS in the
// tile kernels, whose LDS is mostly taken by the lines.
constexpr int kTileCoef = 192;

template <int S, int W, int KIND, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_yz_tile(HeatArgs<double> a) {
  constexpr int TX = YzTile<S, W>::TX, NP = YzTile<S, W>::NP, LD = YzTile<S, W>::LD;
  __shared__ double sT[TX * LD], sR[TX * LD];
  __shared__ double sC[LDS_COEF ? kTileCoef : 1];
  if (LDS_COEF)
    for (int k = threadIdx.x; k < a.nM * a.nM * 3; k += blockDim.x) sC[k] = a.c[k];
  const double* __restrict__ cT = LDS_COEF ? sC : a.c;  // LDS or global, never a generic pointer
  __shared__ uint8_t sM[TX * NP];
  __shared__ double sLogA[KIND == SWEEP_Z ? 256 : 1], sE[KIND == SWEEP_Z ? 256 : 1];
  __shared__ unsigned long long sMax[KIND == SWEEP_Z ? 256 : 1];
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny, nM = a.nM;
  const int nLine = KIND == SWEEP_Y ? ny : nz;
  const int st = KIND == SWEEP_Y ? nx : nxy;
  const int d = KIND == SWEEP_Y ? 1 : 2;
  const int nOther = KIND == SWEEP_Y ? nz : ny;
  const int ntx = (nx + TX - 1) / TX;
  const int tx0 = (blockIdx.x % ntx) * TX, o = blockIdx.x / ntx;
  const int base = KIND == SWEEP_Y ? tx0 + o * nxy : tx0 + o * nx;
  const int cols = min(TX, nx - tx0);
  const double* __restrict__ T1 = a.T1;
  const uint8_t* __restrict__ Mg = a.M;
  if (KIND == SWEEP_Z)
    for (int k = threadIdx.x; k < nM; k += blockDim.x) {
      sLogA[k] = a.logA[k];
      sE[k] = a.E[k];
      sMax[k] = order_key(-INFINITY);
    }
  // all loads of a thread are issued before the first LDS store
  constexpr int kIt = (TX * NP + kHeatTPB - 1) / kHeatTPB;
  {
    double vT[kIt], vR[kIt];
    int vM[kIt];
#pragma unroll
    for (int q = 0; q < kIt; ++q) {
      const int e = threadIdx.x + q * kHeatTPB;
      const int c = e % TX, k = e / TX;
      const bool ok = c < cols && k < nLine;
      const int i = ok ? base + c + k * st : base;
      vT[q] = T1[i];
      vR[q] = a.T2[i];
      vM[q] = Mg[i];
    }
#pragma unroll
    for (int q = 0; q < kIt; ++q) {
      const int e = threadIdx.x + q * kHeatTPB;
      const int c = e % TX, k = e / TX;
      if (c < cols && k < nLine) {
        sT[c * LD + k] = vT[q];
        sR[c * LD + k] = vR[q];
        sM[c * NP + k] = (uint8_t)vM[q];
      }
    }
  }
  __syncthreads();
  // a line per W-lane segment: segment g of wavefront w solves lines
  // w * 64 / W + g, + 4 * 64 / W, ...; a segment past the last line solves
  // an empty system and stores nothing
  const int wl = __lane_id(), lane = wl & (W - 1), wave = threadIdx.x >> 6;
  const bool sk = a.sinked;
  const double dt = a.dt;
  for (int c0 = wave * (64 / W); c0 < cols; c0 += kHeatTPB / W) {
    const bool valid = c0 + wl / W < cols;
    const int c = valid ? c0 + wl / W : c0;
    const int n = valid ? nLine : 0;
    const int x = tx0 + c;
    const bool lf = sk && (x == 0 || x == nx - 1 || o == 0 || o == nOther - 1);
    double Tn[S], R0[S];
    int Mv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      Tn[s] = k < n ? sT[c * LD + k] : 0.0;
      R0[s] = k < n ? sR[c * LD + k] : 0.0;
      Mv[s] = k < n ? (int)sM[c * NP + k] : 0;
    }
    const double Tl = seg_shift<W, false>(Tn[S - 1], 1), Tr = seg_shift<W, true>(Tn[0], 1);
    const int Ml = seg_shift<W, false>(Mv[S - 1], 1), Mr = seg_shift<W, true>(Mv[0], 1);
    double ca[S], cb[S], cc_[S], cd[S], xs[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      ca[s] = 0.0;
      cb[s] = 1.0;
      cc_[s] = 0.0;
      cd[s] = 0.0;
      if (k >= n) continue;
      const int m = Mv[s];
      const double TL = s > 0 ? Tn[s - 1] : Tl, TR = s < S - 1 ? Tn[s + 1] : Tr;
      const int ML = s > 0 ? Mv[s - 1] : Ml, MR = s < S - 1 ? Mv[s + 1] : Mr;
      const bool fz = lf || (sk && (k == 0 || k == n - 1));
      double t = R0[s];
      if (!fz) {
        double dTdt = 0;
        if (k != 0) dTdt -= (TL - Tn[s]) * coef(cT, nM, m, ML, d) / 2;
        if (k != n - 1) dTdt -= (TR - Tn[s]) * coef(cT, nM, m, MR, d) / 2;
        t = t + dt * dTdt;
        if (k < n - 1) {
          const double q = dt / 2 * coef(cT, nM, m, MR, d);
          cb[s] += q;
          cc_[s] = -q;
        }
        if (k > 0) {
          const double q = dt / 2 * coef(cT, nM, m, ML, d);
          cb[s] += q;
          ca[s] = -q;
        }
      }
      cd[s] = t;
    }
    wave_tridiag<S, W>(ca, cb, cc_, cd, xs);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      if (k < n) sR[c * LD + k] = xs[s];
    }
  }
  __syncthreads();
  MaxRun run;
  double* __restrict__ T2 = a.T2;
  double* __restrict__ Om = a.Om;
  double vO[KIND == SWEEP_Z ? kIt : 1];
  if (KIND == SWEEP_Z && a.calcDamage) {
#pragma unroll
    for (int q = 0; q < kIt; ++q) {
      const int e = threadIdx.x + q * kHeatTPB;
      const int c = e % TX, k = e / TX;
      const bool ok = c < cols && k < nLine;
      vO[q] = Om[ok ? base + c + k * st : base];
    }
  }
#pragma unroll
  for (int q = 0; q < kIt; ++q) {
    const int e = threadIdx.x + q * kHeatTPB;
    const int c = e % TX, k = e / TX;
    if (c >= cols || k >= nLine) continue;
    const int i = base + c + k * st;
    const double t = sR[c * LD + k];
    T2[i] = t;
    if (KIND == SWEEP_Z) {
      const int m = sM[c * NP + k];
      if (a.calcDamage && sE[m] != 0) Om[i] = vO[q] + dt * exp(sLogA[m] - sE[m] / (a.Rgas * ((t + sT[c * LD + k]) / 2 + a.czero)));
      run.add(m, t, sMax);
    }
  }
  if (KIND == SWEEP_Z) {
    run.flush(sMax);
    fold_lds_max(sMax, a.maxKey, nM);
  }
}

// rows per lane of the wavefront x-solve for a line of n voxels (0: use the
// sequential lane-per-line kernel)
inline int wave_rows(int n) { return n <= 64 ? 1 : n <= 128 ? 2 : n <= 256 ? 4 : n <= 512 ? 8 : 0; }

// rows per lane S and lanes per line W of the forward wavefront solver for a
// line of n <= 512 voxels.  Fewer lanes per line mean more rows per lane and
// fewer cyclic-reduction steps, with 64 / W lines solved side by side in one
// wavefront.  Measured per step (200 on + 200 off steps, layered skin):
// y/z sweeps with W = 16 up to 128 voxels and W = 32 up to 256 took the DG-ADI
// step at 128^3 from 0.097 to 0.064 ms and at 200^3 from 0.259 to 0.232 ms;
// the x sweep, which also gathers four neighbouring lines, gains from W = 32
// only up to 128 voxels (more rows per lane cost it occupancy: 154 VGPRs at
// S = 8).
struct WaveShape {
  int S, W;
};
inline WaveShape wave_shape(int n, bool xAxis) {
  if (n <= 32 || n > 256) return {wave_rows(n), 64};
  if (xAxis) return n <= 128 ? WaveShape{n <= 64 ? 2 : 4, 32} : WaveShape{4, 64};
  return n <= 128 ? WaveShape{n <= 64 ? 4 : 8, 16} : WaveShape{8, 32};
}

// calls f(integral_constant S, integral_constant W) for the shapes
// wave_shape picks (X: the x sweep's, else the y/z sweeps')
template <bool X, class F>
inline void with_wave_shape(WaveShape ws, F&& f) {
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I4 = std::integral_constant<int, 4>;
  using I8 = std::integral_constant<int, 8>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  const int key = ws.W * 16 + ws.S;
  if (key == 64 * 16 + 1) return f(I1(), I64());
  if (key == 64 * 16 + 8) return f(I8(), I64());
  if constexpr (X) {
    if (key == 64 * 16 + 4) return f(I4(), I64());
    if (key == 32 * 16 + 2) return f(I2(), I32());
    if (key == 32 * 16 + 4) return f(I4(), I32());
  } else {
    if (key == 32 * 16 + 8) return f(I8(), I32());
    if (key == 16 * 16 + 4) return f(I4(), I16());
    if (key == 16 * 16 + 8) return f(I8(), I16());
  }
}

// Forward Euler step (no reference counterpart; the 7-point right-hand side
// of substep 1a with all three axes fully weighted, the damage update and
// the per-medium maxima of the DG-ADI step).  A workgroup owns a
// kTileX x kTileY column tile and a run of kZRun z planes (chosen per plan by
// timing, see tune_explicit), marching up in z:
// each plane of T and M (with its x/y halo) is staged through LDS once, so
// every voxel is loaded from memory once and its four in-plane neighbours
// come from LDS; the z neighbours are carried in registers.
constexpr int kTileX = 64, kTileY = 4;

template <typename RT, bool LDS_COEF>
__global__ void __launch_bounds__(kTileX * kTileY) heat_explicit_tile(HeatArgs<RT> a, int kZRun) {
  __shared__ HeatLds<RT> L;
  __shared__ RT sT[kTileY + 2][kTileX + 2];
  __shared__ uint8_t sM[kTileY + 2][kTileX + 2];
  step_sensors(a);
  (void)stage_tables(a, L);
  const RT* cT = LDS_COEF ? L.c : a.c;  // LDS or global, never a generic pointer
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny, nM = a.nM;
  const int ntx = (nx + kTileX - 1) / kTileX, nty = (ny + kTileY - 1) / kTileY;
  const int b = blockIdx.x;
  const int x0 = (b % ntx) * kTileX, y0 = (b / ntx % nty) * kTileY, z0 = b / (ntx * nty) * kZRun;
  const int tx = threadIdx.x % kTileX, ty = threadIdx.x / kTileX;
  const int x = x0 + tx, y = y0 + ty;
  const bool ok = x < nx && y < ny;
  const int col = ok ? x + y * nx : 0;
  const int zEnd = min(z0 + kZRun, nz);
  // halo duty: the tile's edge threads also stage the neighbouring column
  const int hx = tx == 0 ? x - 1 : (tx == kTileX - 1 ? x + 1 : -1);
  const int hy = ty == 0 ? y - 1 : (ty == kTileY - 1 ? y + 1 : -1);
  const bool hxOk = ok && hx >= 0 && hx < nx, hyOk = ok && hy >= 0 && hy < ny;
  const int hxCol = hxOk ? hx + y * nx : 0, hyCol = hyOk ? x + hy * nx : 0;
  const int sxh = tx == 0 ? 0 : kTileX + 1, syh = ty == 0 ? 0 : kTileY + 1;
  const RT dt = a.dt;
  MaxRun run;
  RT tBelow = 0, tHere = 0;
  int mBelow = 0, mHere = 0;
  if (ok) {
    tBelow = z0 > 0 ? a.T1[col + (z0 - 1) * nxy] : (RT)0;
    mBelow = z0 > 0 ? (int)a.M[col + (z0 - 1) * nxy] : 0;
    tHere = a.T1[col + z0 * nxy];
    mHere = a.M[col + z0 * nxy];
  }
  for (int z = z0; z < zEnd; ++z) {
    const int i = col + z * nxy;
    // this thread's loads for the step, all issued before the LDS stores
    RT tAbove = 0, hxT = 0, hyT = 0;
    int mAbove = 0, hxM = 0, hyM = 0;
    if (ok && z != nz - 1) {
      tAbove = a.T1[i + nxy];
      mAbove = a.M[i + nxy];
    }
    if (hxOk) {
      hxT = a.T1[hxCol + z * nxy];
      hxM = a.M[hxCol + z * nxy];
    }
    if (hyOk) {
      hyT = a.T1[hyCol + z * nxy];
      hyM = a.M[hyCol + z * nxy];
    }
    const RT q = ok && a.lightsOn ? (RT)a.Q[i] : (RT)0;
    __syncthreads();  // the previous plane's readers are done
    sT[ty + 1][tx + 1] = tHere;
    sM[ty + 1][tx + 1] = (uint8_t)mHere;
    if (hxOk) {
      sT[ty + 1][sxh] = hxT;
      sM[ty + 1][sxh] = (uint8_t)hxM;
    }
    if (hyOk) {
      sT[syh][tx + 1] = hyT;
      sM[syh][tx + 1] = (uint8_t)hyM;
    }
    __syncthreads();
    if (ok) {
      RT dTdt = 0;
      if (!frozen(a, x, y, z)) {
        dTdt = a.lightsOn ? q : (RT)0;
        if (x != 0) dTdt += (sT[ty + 1][tx] - tHere) * coef(cT, nM, mHere, (int)sM[ty + 1][tx], 0);
        if (x != nx - 1) dTdt += (sT[ty + 1][tx + 2] - tHere) * coef(cT, nM, mHere, (int)sM[ty + 1][tx + 2], 0);
        if (y != 0) dTdt += (sT[ty][tx + 1] - tHere) * coef(cT, nM, mHere, (int)sM[ty][tx + 1], 1);
        if (y != ny - 1) dTdt += (sT[ty + 2][tx + 1] - tHere) * coef(cT, nM, mHere, (int)sM[ty + 2][tx + 1], 1);
        if (z != 0) dTdt += (tBelow - tHere) * coef(cT, nM, mHere, mBelow, 2);
        if (z != nz - 1) dTdt += (tAbove - tHere) * coef(cT, nM, mHere, mAbove, 2);
      }
      const RT t = tHere + dt * dTdt;
      a.T2[i] = t;
      if (a.calcDamage && L.E[mHere] != 0)
        a.Om[i] += dt * (RT)exp(L.logA[mHere] - L.E[mHere] / (a.Rgas * ((t + tHere) / 2 + a.czero)));
      run.add(mHere, (double)t, L.maxKey);
    }
    tBelow = tHere;
    mBelow = mHere;
    tHere = tAbove;
    mHere = mAbove;
  }
  run.flush(L.maxKey);
  fold_lds_max(L.maxKey, a.maxKey, a.nM);
}

// Sensor record of T into column `col` on its own (the final state of a
// run, recorded when the record is read).
template <typename RT>
__global__ void heat_sensors(const RT* T, int nx, int nxy, const long long* idx, const RT* w, int nS, double* out, long long col) {
  sensors_at(T, nx, nxy, idx, w, nS, out, col, (int)threadIdx.x, (int)blockDim.x);
}

// step counter of graph-replayed steps: set before a series of replays,
// advanced by the last node of each replay
__global__ void heat_step_counter(long long* c, long long value, int add) {
  if (threadIdx.x == 0) *c = add ? *c + value : value;
}

template <typename RT>
__global__ void heat_load(RT* dst, const float* src32, const double* src64, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src64 ? (RT)src64[i] : (RT)src32[i];
}

template <typename RT>
__global__ void heat_store(const RT* src, float* dst32, double* dst64, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    if (dst32) dst32[i] = (float)src[i];
    if (dst64) dst64[i] = (double)src[i];
  }
}

__global__ void heat_init_keys(unsigned long long* k, int n) {
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i < n) k[i] = order_key(-INFINITY);
}


// ---------------------------------------------------------------------------
// Discrete adjoint of the FP64 DG-ADI step (no reference counterpart).
//
// With D the non-frozen voxels and L_d the conduction operators
// ((L_d u)_i = sum over the d-neighbours n of c[M_i, M_n, d] (u_n - u_i)),
// one step is
//   T*   = u + dt D (L_x/2 + L_y + L_z) u + dt lightsOn D Q
//   T1   = A_x^-1 T*,  T2 = A_y^-1 (T1 - dt/2 D L_y u),  T3 = A_z^-1 (T2 - dt/2 D L_z u)
// with A_d = I - dt/2 D L_d.  For an adjoint lam of T3 the adjoints are
//   mu3 = A_z^-T lam,  mu2 = A_y^-T mu3,  mu1 = A_x^-T mu2
//   lam_u = mu1 + dt (L_x/2 + L_y + L_z)^T D mu1 - dt/2 L_y^T D mu2 - dt/2 L_z^T D mu3
//   lam_Q = dt lightsOn D mu1
// and the damage of the step, dt A exp(-E / (R ((T3 + u)/2 + 273.15))),
// adds dt A exp(.) E / (R theta^2) / 2 to the adjoints of both u and T3.

// adjoint of the step's damage term for the target voxels
__global__ void heat_adj_damage(HeatArgs<double> a, const double* __restrict__ x0, const double* __restrict__ x1,
                                const uint8_t* __restrict__ target, double* lam, int64_t L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L || !target[i]) return;
  const int m = a.M[i];
  const double E = a.E[m];
  if (E == 0) return;
  const double th = (x0[i] + x1[i]) / 2 + a.czero;
  lam[i] += a.dt * exp(a.logA[m] - E / (a.Rgas * th)) * E / (a.Rgas * th * th) / 2;
}

// A_d^T y = r on every line along axis d: lane per line, Thomas recurrence;
// `beta` holds the eliminated diagonal.
template <int AXIS>
__global__ void __launch_bounds__(kHeatTPB) heat_adj_solve(HeatArgs<double> a, const double* __restrict__ r, double* y, double* beta) {
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  int i0, st, n;
  bool lf;
  if (AXIS == 0) {
    if (line >= ny * nz) return;
    const int yy = line % ny, zz = line / ny;
    i0 = yy * nx + zz * nxy, st = 1, n = nx;
    lf = a.sinked && (yy == 0 || yy == ny - 1 || zz == 0 || zz == nz - 1);
  } else if (AXIS == 1) {
    if (line >= nx * nz) return;
    const int xx = line % nx, zz = line / nx;
    i0 = xx + zz * nxy, st = nx, n = ny;
    lf = a.sinked && (xx == 0 || xx == nx - 1 || zz == 0 || zz == nz - 1);
  } else {
    if (line >= nxy) return;
    const int xx = line % nx, yy = line / nx;
    i0 = line, st = nxy, n = nz;
    lf = a.sinked && (xx == 0 || xx == nx - 1 || yy == 0 || yy == ny - 1);
  }
  const double h = a.dt / 2;
  // row k of A_d: (sub, diag, super) = (ak, bk, ck)
  auto row = [&](int k, double& ak, double& bk, double& ck) {
    ak = 0;
    bk = 1;
    ck = 0;
    if (lf || (a.sinked && (k == 0 || k == n - 1))) return;
    const int i = i0 + k * st, m = a.M[i];
    if (k > 0) {
      const double q = h * cc(a, m, a.M[i - st], AXIS);
      bk += q;
      ak = -q;
    }
    if (k < n - 1) {
      const double q = h * cc(a, m, a.M[i + st], AXIS);
      bk += q;
      ck = -q;
    }
  };
  double aPrev, bPrev, cPrev, ak, bk, ck;
  row(0, ak, bk, ck);
  double bet = bk, yy = r[i0];
  beta[i0] = bet;
  y[i0] = yy;
  for (int k = 1; k < n; ++k) {
    aPrev = ak;
    bPrev = bk;
    cPrev = ck;
    row(k, ak, bk, ck);
    // A^T row k: sub = c_(k-1), diag = b_k; A^T row k-1 has super = a_k
    const double w = cPrev / bet;
    bet = bk - w * ak;
    const int i = i0 + k * st;
    yy = r[i] - w * yy;
    beta[i] = bet;
    y[i] = yy;
    (void)aPrev;
    (void)bPrev;
  }
  double xn = y[i0 + (n - 1) * st] / beta[i0 + (n - 1) * st];
  y[i0 + (n - 1) * st] = xn;
  double aNext;
  {
    double t0, t1;
    row(n - 1, aNext, t0, t1);
  }
  for (int k = n - 2; k >= 0; --k) {
    const int i = i0 + k * st;
    xn = (y[i] - aNext * xn) / beta[i];
    y[i] = xn;
    double t0, t1;
    row(k, aNext, t0, t1);
  }
}

// A_d^T x = r along one line of n voxels held S per lane by a wavefront
// (lane l: voxels l*S .. l*S+S-1; rv the right-hand side, Mv the media; lf:
// the whole line is frozen).  The rows of A^T are (sub, diag, super) =
// (c_(k-1), b_k, a_(k+1)) of A's rows; the neighbouring lane's edge
// coefficients come by one shuffle each.
template <int S, int W>
__device__ inline void adj_line_solve(const HeatArgs<double>& a, const double* cT, int axis, int n, bool lf, const double (&rv)[S],
                                      const int (&Mv)[S], double (&xs)[S]) {
  const int wl = __lane_id(), lane = wl & (W - 1);  // lane within the line's W-lane segment
  const double h = a.dt / 2;
  double A[S], B[S], Cc[S];
  const int Ml = seg_shift<W, false>(Mv[S - 1], 1), Mr = seg_shift<W, true>(Mv[0], 1);
  // rows of A
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int k = lane * S + s;
    A[s] = 0.0;
    B[s] = 1.0;
    Cc[s] = 0.0;
    if (k >= n || lf || (a.sinked && (k == 0 || k == n - 1))) continue;
    const int m = Mv[s];
    const int ML = s > 0 ? Mv[s - 1] : Ml, MR = s < S - 1 ? Mv[s + 1] : Mr;
    if (k > 0) {
      const double q = h * coef(cT, a.nM, m, ML, axis);
      B[s] += q;
      A[s] = -q;
    }
    if (k < n - 1) {
      const double q = h * coef(cT, a.nM, m, MR, axis);
      B[s] += q;
      Cc[s] = -q;
    }
  }
  // rows of A^T
  const double cPrevLane = seg_shift<W, false>(Cc[S - 1], 1), aNextLane = seg_shift<W, true>(A[0], 1);
  double ta[S], tb[S], tc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int k = lane * S + s;
    const bool ok = k < n;
    ta[s] = ok && k > 0 ? (s > 0 ? Cc[s - 1] : cPrevLane) : 0.0;
    tb[s] = ok ? B[s] : 1.0;
    tc[s] = ok && k < n - 1 ? (s < S - 1 ? A[s + 1] : aNextLane) : 0.0;
  }
  wave_tridiag<S, W>(ta, tb, tc, rv, xs);
}

// A_d^T y = r with one wavefront per line (lines of up to 512 voxels),
// loading the line directly (x lines; strided along y and z).
template <int S, int W, int AXIS, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_adj_solve_wave(HeatArgs<double> a, const double* __restrict__ r, double* y) {
  __shared__ double sC[LDS_COEF ? kTileCoef : 1];
  if (LDS_COEF) {
    for (int k = threadIdx.x; k < a.nM * a.nM * 3; k += blockDim.x) sC[k] = a.c[k];
    __syncthreads();
  }
  const double* __restrict__ cT = LDS_COEF ? sC : a.c;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
  // a line per W-lane segment; a segment past the last line solves the
  // last line again and stores nothing
  const int wl = __lane_id(), lane = wl & (W - 1);
  const int seg = (blockIdx.x * (kHeatTPB / 64) + (threadIdx.x >> 6)) * (64 / W) + wl / W;
  const int nLines = AXIS == 0 ? ny * nz : AXIS == 1 ? nx * nz : nxy;
  const bool valid = seg < nLines;
  const int line = valid ? seg : nLines - 1;
  int i0, st, n;
  bool lf;
  if (AXIS == 0) {
    const int yy = line % ny, zz = line / ny;
    i0 = yy * nx + zz * nxy, st = 1, n = nx;
    lf = a.sinked && (yy == 0 || yy == ny - 1 || zz == 0 || zz == nz - 1);
  } else if (AXIS == 1) {
    const int xx = line % nx, zz = line / nx;
    i0 = xx + zz * nxy, st = nx, n = ny;
    lf = a.sinked && (xx == 0 || xx == nx - 1 || zz == 0 || zz == nz - 1);
  } else {
    i0 = line, st = nxy, n = nz;
    const int xx = line % nx, yy = line / nx;
    lf = a.sinked && (xx == 0 || xx == nx - 1 || yy == 0 || yy == ny - 1);
  }
  double rv[S], xs[S];
  int Mv[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int k = lane * S + s;
    const int i = i0 + (k < n ? k : 0) * st;
    rv[s] = k < n ? r[i] : 0.0;
    Mv[s] = a.M[i];
  }
  adj_line_solve<S, W>(a, cT, AXIS, n, lf, rv, Mv, xs);
  if (!valid) return;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int k = lane * S + s;
    if (k < n) y[i0 + k * st] = xs[s];
  }
}

// A_d^T y = r along y (AXIS 1) or z (AXIS 2) with the lines staged through
// LDS like the forward y/z sweeps: a workgroup takes TX consecutive x-lines
// at one z (AXIS 1) or y (AXIS 2), loads them row by row (coalesced), each
// wavefront solves TX/4 lines, and the rows are written back.
template <int S, int W, int AXIS, bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_adj_solve_tile(HeatArgs<double> a, const double* __restrict__ r, double* y) {
  constexpr int TX = YzTile<S, W>::TX, NP = YzTile<S, W>::NP, LD = YzTile<S, W>::LD;
  __shared__ double sR[TX * LD];
  __shared__ uint8_t sM[TX * NP];
  __shared__ double sC[LDS_COEF ? kTileCoef : 1];
  if (LDS_COEF)
    for (int k = threadIdx.x; k < a.nM * a.nM * 3; k += blockDim.x) sC[k] = a.c[k];
  const double* __restrict__ cT = LDS_COEF ? sC : a.c;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
  const int nLine = AXIS == 1 ? ny : nz;
  const int st = AXIS == 1 ? nx : nxy;
  const int nOther = AXIS == 1 ? nz : ny;
  const int ntx = (nx + TX - 1) / TX;
  const int tx0 = (blockIdx.x % ntx) * TX, o = blockIdx.x / ntx;
  const int base = AXIS == 1 ? tx0 + o * nxy : tx0 + o * nx;
  const int cols = min(TX, nx - tx0);
  constexpr int kIt = (TX * NP + kHeatTPB - 1) / kHeatTPB;
  {
    double vR[kIt];
    int vM[kIt];
#pragma unroll
    for (int q = 0; q < kIt; ++q) {
      const int e = threadIdx.x + q * kHeatTPB;
      const int c = e % TX, k = e / TX;
      const int i = c < cols && k < nLine ? base + c + k * st : base;
      vR[q] = r[i];
      vM[q] = a.M[i];
    }
#pragma unroll
    for (int q = 0; q < kIt; ++q) {
      const int e = threadIdx.x + q * kHeatTPB;
      const int c = e % TX, k = e / TX;
      if (c < cols && k < nLine) {
        sR[c * LD + k] = vR[q];
        sM[c * NP + k] = (uint8_t)vM[q];
      }
    }
  }
  __syncthreads();
  // a line per W-lane segment; a segment past the last line solves an
  // empty system and stores nothing
  const int wl = __lane_id(), lane = wl & (W - 1), wave = threadIdx.x >> 6;
  for (int c0 = wave * (64 / W); c0 < cols; c0 += kHeatTPB / W) {
    const bool valid = c0 + wl / W < cols;
    const int c = valid ? c0 + wl / W : c0;
    const int n = valid ? nLine : 0;
    const int x = tx0 + c;
    const bool lf = a.sinked && (x == 0 || x == nx - 1 || o == 0 || o == nOther - 1);
    double rv[S], xs[S];
    int Mv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      rv[s] = k < n ? sR[c * LD + k] : 0.0;
      Mv[s] = k < n ? (int)sM[c * NP + k] : 0;
    }
    adj_line_solve<S, W>(a, cT, AXIS, n, lf, rv, Mv, xs);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k = lane * S + s;
      if (k < n) sR[c * LD + k] = xs[s];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kIt; ++q) {
    const int e = threadIdx.x + q * kHeatTPB;
    const int c = e % TX, k = e / TX;
    if (c < cols && k < nLine) y[base + c + k * st] = sR[c * LD + k];
  }
}

__device__ inline bool heat_frozen(const HeatArgs<double>& a, int x, int y, int z) { return frozen(a, x, y, z); }

// lam_u of the step (plus the damage adjoint of u) and the source adjoint.
template <bool LDS_COEF>
__global__ void __launch_bounds__(kHeatTPB) heat_adj_combine(HeatArgs<double> a, const double* __restrict__ mu1, const double* __restrict__ mu2,
                                                            const double* __restrict__ mu3, const double* __restrict__ x0,
                                                            const double* __restrict__ x1, const uint8_t* __restrict__ target,
                                                            double* lamOut, double* dQ, int lightsOn) {
  __shared__ double sC[LDS_COEF ? kTileCoef : 1];
  if (LDS_COEF) {
    for (int k = threadIdx.x; k < a.nM * a.nM * 3; k += blockDim.x) sC[k] = a.c[k];
    __syncthreads();
  }
  const double* __restrict__ cT = LDS_COEF ? sC : a.c;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
  // one z plane per grid row: 32-bit in-plane index arithmetic
  const int p = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
  if (p >= nxy) return;
  const int x = p % nx, y = p / nx;
  const int64_t i = (int64_t)z * nxy + p;
  const int m = a.M[i];
  const double dt = a.dt;
  // L_d^T (D w) at i for w = mu1, mu2, mu3
  auto wv = [&](const double* mu, int xx, int yy, int zz) { return heat_frozen(a, xx, yy, zz) ? 0.0 : mu[xx + yy * nx + (int64_t)zz * nxy]; };
  double t1[3] = {0, 0, 0}, t2 = 0, t3 = 0;
  const int xyz[3] = {x, y, z}, ext[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    for (int s = -1; s <= 1; s += 2) {
      const int q = xyz[d] + s;
      if (q < 0 || q >= ext[d]) continue;
      const int xx = d == 0 ? q : x, yy = d == 1 ? q : y, zz = d == 2 ? q : z;
      const int mn = a.M[xx + yy * nx + (int64_t)zz * nxy];
      const double cin = coef(cT, a.nM, mn, m, d), cout = coef(cT, a.nM, m, mn, d);
      t1[d] += cin * wv(mu1, xx, yy, zz) - cout * wv(mu1, x, y, z);
      if (d == 1) t2 += cin * wv(mu2, xx, yy, zz) - cout * wv(mu2, x, y, z);
      if (d == 2) t3 += cin * wv(mu3, xx, yy, zz) - cout * wv(mu3, x, y, z);
    }
  }
  double lam = mu1[i] + dt * (t1[0] / 2 + t1[1] + t1[2]) - dt / 2 * t2 - dt / 2 * t3;
  if (target[i] && a.E[m] != 0) {
    const double E = a.E[m];
    const double th = (x0[i] + x1[i]) / 2 + a.czero;
    lam += dt * exp(a.logA[m] - E / (a.Rgas * th)) * E / (a.Rgas * th * th) / 2;
  }
  lamOut[i] = lam;
  if (lightsOn) dQ[i] += dt * wv(mu1, x, y, z);
}

// Gradient with respect to the conduction coefficients c[a, b, d] (the
// reference's dTdtperdeltaT).  Perturbing c in the step's equations and
// applying the adjoints mu1..mu3 gives, with dL_d v at i = sum over the
// d-neighbours n of dc[M_i, M_n, d] (v_n - v_i),
//   dJ = dt/2 mu1^T D dL_x (u + T1)
//      + mu1^T D dL_y (dt u) + mu2^T D dL_y (dt/2 (T2 - u))
//      + mu1^T D dL_z (dt u) + mu3^T D dL_z (dt/2 (T3 - u)).
// The intermediate states follow from the step's two ends by undoing the
// line solves: T2 = A_z T3 + dt/2 D L_z u = T3 - dt/2 D L_z (T3 - u) and
// T1 = T2 - dt/2 D L_y (T2 - u).

// lo = hi - dt/2 D L_AXIS (hi - u)
template <int AXIS>
__global__ void __launch_bounds__(kHeatTPB) heat_adj_undo(HeatArgs<double> a, const double* __restrict__ hi, const double* __restrict__ u,
                                                         double* __restrict__ lo) {
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
  const int p = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
  if (p >= nxy) return;
  const int x = p % nx, y = p / nx;
  const int64_t i = (int64_t)z * nxy + p;
  const double vi = hi[i] - u[i];
  double l = 0;
  if (!heat_frozen(a, x, y, z)) {
    const int q = AXIS == 1 ? y : z, n = AXIS == 1 ? ny : nz;
    const int64_t st = AXIS == 1 ? nx : nxy;
    const int m = a.M[i];
    if (q > 0) l += coef(a.c, a.nM, m, (int)a.M[i - st], AXIS) * ((hi[i - st] - u[i - st]) - vi);
    if (q < n - 1) l += coef(a.c, a.nM, m, (int)a.M[i + st], AXIS) * ((hi[i + st] - u[i + st]) - vi);
  }
  lo[i] = hi[i] - a.dt / 2 * l;
}

// One step's contributions to dJ/dc, summed per (a, b, d) in LDS (one FP64
// LDS atomic per wavefront where all its lanes hit the same coefficient,
// the usual case away from interfaces) and added to the global sums; a
// grid-stride loop keeps the number of global atomics per step small.
constexpr int kCoefBins = 192;  // LDS bins: tables of up to 8 media
template <bool LDS_BINS>
__global__ void __launch_bounds__(kHeatTPB) heat_adj_coef(HeatArgs<double> a, const double* __restrict__ mu1, const double* __restrict__ mu2,
                                                         const double* __restrict__ mu3, const double* __restrict__ u,
                                                         const double* __restrict__ T1, const double* __restrict__ T2,
                                                         const double* __restrict__ T3, double* dJdc) {
  __shared__ double sB[LDS_BINS ? kCoefBins : 1];
  const int nM = a.nM, nB = nM * nM * 3;
  if (LDS_BINS) {
    for (int k = threadIdx.x; k < nB; k += blockDim.x) sB[k] = 0;
    __syncthreads();
  }
  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
  const int64_t L = (int64_t)nxy * nz;
  const double dt = a.dt;
  const int lane = threadIdx.x & 63;
  const int64_t span = (int64_t)gridDim.x * blockDim.x;
  // whole wavefronts run the loop together (shuffles below need every lane)
  const int64_t iters = (L + span - 1) / span;
  for (int64_t it = 0; it < iters; ++it) {
    const int64_t i = it * span + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < L;
    int x = 0, y = 0, z = 0, m = 0;
    bool act = false;
    if (live) {
      z = (int)(i / nxy);
      const int p = (int)(i - (int64_t)z * nxy);
      x = p % nx, y = p / nx;
      m = a.M[i];
      act = !heat_frozen(a, x, y, z);
    }
    for (int d = 0; d < 3; ++d) {
      for (int s = -1; s <= 1; s += 2) {
        double g = 0;
        int k = -1;
        if (act) {
          const int q = (d == 0 ? x : d == 1 ? y : z) + s, n = d == 0 ? nx : d == 1 ? ny : nz;
          if (q >= 0 && q < n) {
            const int64_t st = d == 0 ? 1 : d == 1 ? nx : nxy;
            const int64_t j = i + s * st;
            const int mn = a.M[j];
            k = m + mn * nM + d * nM * nM;
            if (d == 0) {
              g = dt / 2 * mu1[i] * ((u[j] + T1[j]) - (u[i] + T1[i]));
            } else {
              const double* W = d == 1 ? mu2 : mu3;
              const double* V = d == 1 ? T2 : T3;
              g = dt * mu1[i] * (u[j] - u[i]) + dt / 2 * W[i] * ((V[j] - u[j]) - (V[i] - u[i]));
            }
          }
        }
        if (LDS_BINS) {
          const int k0 = __shfl(k, 0);
          if (__all(k == k0) && k0 >= 0) {
            for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off);
            if (lane == 0) atomicAdd(&sB[k0], g);
          } else if (k >= 0) {
            atomicAdd(&sB[k], g);
          }
        } else if (k >= 0 && g != 0) {
          atomicAdd(&dJdc[k], g);
        }
      }
    }
  }
  if (LDS_BINS) {
    __syncthreads();
    for (int k = threadIdx.x; k < nB; k += blockDim.x)
      if (sB[k] != 0) atomicAdd(&dJdc[k], sB[k]);
  }
}

}  // namespace mcm

using namespace mcm;

struct HeatBuf {
  void* p = nullptr;
  ~HeatBuf() {
    if (p) (void)hipFree(p);
  }
};

// An instantiated graph of `steps` time steps at one (dt, lights) setting,
// valid while the state and sensor buffers it captured stay in place.
struct HeatGraph {
  hipGraphExec_t exec = nullptr;
  int64_t steps = 0;
  double dt = 0;
  int lightsOn = 0;
  const void *T1 = nullptr, *sensors = nullptr;
};

struct mcm_heat_plan {
  int device = 0;
  hipStream_t stream = nullptr;
  int fp32 = 0, explicitScheme = 0;
  int explicitZRun = 16;  // z planes per workgroup of the explicit stencil
  int nx = 0, ny = 0, nz = 0, nM = 0;
  int64_t L = 0;
  int sinked = 0, calcDamage = 0;
  int64_t nS = 0;
  int64_t stepsDone = 0;
  int64_t sensorsRecorded = 0;  // sensor columns 0 .. sensorsRecorded-1 are in `sensors`
  int64_t sensorCap = 0;      // columns allocated in `sensors`
  bool interrupted = false;
  int (*interrupt)(void) = nullptr;
  HeatBuf M, Q, c, logA, E, T1, T2, B, Om, maxKey, sIdx, sW, sensors, stage32, stage64, stepDev;
  double constR = 0, constC = 0;
  int64_t graphSteps = 32;    // steps per captured graph (0: launch every step)
  std::vector<HeatGraph> graphs;
  void drop_graphs() {
    for (HeatGraph& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  ~mcm_heat_plan() {
    drop_graphs();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Plans advanced together: one graph holds K steps of every plan, each plan
// a branch on its own stream, so the small kernels of different grids run
// concurrently.
struct mcm_heat_group {
  int device = 0;
  std::vector<mcm_heat_plan*> plans;
  std::vector<hipEvent_t> events;  // [0] fork, [1 + i] join of plan i
  struct Graph {
    hipGraphExec_t exec = nullptr;
    int64_t steps = 0;
    std::vector<double> dt;
    std::vector<int> lightsOn;
    std::vector<const void*> T1, sensors;
  };
  std::vector<Graph> graphs;
  void drop_graphs() {
    for (Graph& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  ~mcm_heat_group() {
    drop_graphs();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
};

namespace {

size_t rt_size(const mcm_heat_plan* p) { return p->fp32 ? sizeof(float) : sizeof(double); }

int halloc(HeatBuf& b, size_t bytes) {
  MCM_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 8)));
  return MCM_OK;
}

template <typename RT>
void launch_explicit(mcm_heat_plan* P, const HeatArgs<RT>& a, int zr) {
  const unsigned ge = (unsigned)(((P->nx + kTileX - 1) / kTileX) * ((P->ny + kTileY - 1) / kTileY) * ((P->nz + zr - 1) / zr));
  if (P->nM * P->nM * 3 <= kLdsCoef) hipLaunchKernelGGL((heat_explicit_tile<RT, true>), dim3(ge), dim3(kTileX * kTileY), 0, P->stream, a, zr);
  else hipLaunchKernelGGL((heat_explicit_tile<RT, false>), dim3(ge), dim3(kTileX * kTileY), 0, P->stream, a, zr);
}

template <typename RT>
HeatArgs<RT> make_args(mcm_heat_plan* P, double dt, int lightsOn) {
  HeatArgs<RT> a{};
  a.nx = P->nx;
  a.ny = P->ny;
  a.nz = P->nz;
  a.nM = P->nM;
  a.sinked = P->sinked;
  a.lightsOn = lightsOn;
  a.calcDamage = P->calcDamage;
  a.dt = (RT)dt;
  a.Rgas = (RT)P->constR;
  a.czero = (RT)P->constC;
  a.M = (const uint8_t*)P->M.p;
  a.Q = (const float*)P->Q.p;
  a.c = (const RT*)P->c.p;
  a.logA = (const RT*)P->logA.p;
  a.E = (const RT*)P->E.p;
  a.T1 = (const RT*)P->T1.p;
  a.T2 = (RT*)P->T2.p;
  a.B = (RT*)P->B.p;
  a.Om = (RT*)P->Om.p;
  a.maxKey = (unsigned long long*)P->maxKey.p;
  return a;
}

// The explicit stencil's best run of z planes per workgroup depends on how
// the workgroup count fills the GPU (measured, ms/step at 100^3 / 160^3 /
// 200^3 / 256^3: 8 planes 0.0189 / 0.0446 / 0.0792 / 0.1272, 16 planes
// 0.0232 / 0.0375 / 0.0628 / 0.1095, 32 planes 0.0409 / 0.0501 / 0.0561 /
// 0.1175), so each explicit plan times a few candidate runs once at
// creation.  The trial steps read the initial state and write only the
// other state buffer (overwritten by the first real step), a scratch damage
// field (the DG-ADI work buffer, unused by this scheme) and scratch maxima;
// no sensor is recorded.
// The choice is kept per grid shape, precision and device for the process,
// so later plans of the same shape (a schedule cut into many calls) do not
// time again; a plan created for fewer than kTuneMinSteps steps does not
// time at all and takes kDefaultZRun, unless the shape is known.
constexpr int kDefaultZRun = 16;
constexpr int64_t kTuneMinSteps = 64;
std::mutex g_zrunMutex;
std::map<std::array<int, 6>, int> g_zrunCache;

bool valid_zrun(int zr) { return zr == 8 || zr == 12 || zr == 16 || zr == 24 || zr == 32; }

template <typename RT>
int tune_explicit(mcm_heat_plan* P, int64_t steps) {
  const std::array<int, 6> key = {P->device, P->nx, P->ny, P->nz, P->fp32, P->nM * P->nM * 3 <= kLdsCoef};
  if (const char* e = std::getenv("MCM_HEAT_ZRUN")) {
    const int zr = std::atoi(e);
    if (valid_zrun(zr)) {
      P->explicitZRun = zr;
      return MCM_OK;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_zrunMutex);
    auto it = g_zrunCache.find(key);
    if (it != g_zrunCache.end()) {
      P->explicitZRun = it->second;
      return MCM_OK;
    }
  }
  if (steps < kTuneMinSteps) {
    P->explicitZRun = kDefaultZRun;
    return MCM_OK;
  }
  HeatArgs<RT> a = make_args<RT>(P, 1e-9, 1);
  a.Om = (RT*)P->B.p;
  HeatBuf keys;
  if (int rc = halloc(keys, (size_t)P->nM * 8)) return rc;
  a.maxKey = (unsigned long long*)keys.p;
  hipEvent_t e0, e1;
  MCM_HIP(hipEventCreate(&e0));
  MCM_HIP(hipEventCreate(&e1));
  float best = 0.f;
  int bestZ = 16;
  for (int zr : {8, 12, 16, 24, 32}) {
    if (zr > 8 && zr >= P->nz + 8) break;  // longer runs than the grid change nothing
    launch_explicit<RT>(P, a, zr);
    (void)hipEventRecord(e0, P->stream);
    for (int r = 0; r < 3; ++r) launch_explicit<RT>(P, a, zr);
    (void)hipEventRecord(e1, P->stream);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
    if (best == 0.f || ms < best) best = ms, bestZ = zr;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MCM_HIP(hipGetLastError());
  MCM_HIP(hipStreamSynchronize(P->stream));
  P->explicitZRun = bestZ;
  std::lock_guard<std::mutex> lk(g_zrunMutex);
  g_zrunCache[key] = bestZ;
  return MCM_OK;
}

int ensure_sensor_capacity(mcm_heat_plan* P, int64_t cols) {
  if (P->nS == 0 || cols <= P->sensorCap) return MCM_OK;
  int64_t cap = std::max<int64_t>(cols, 2 * P->sensorCap);
  void* np = nullptr;
  MCM_HIP(hipMalloc(&np, (size_t)(cap * P->nS) * sizeof(double)));
  if (P->sensors.p) {
    MCM_HIP(hipMemcpyAsync(np, P->sensors.p, (size_t)(P->sensorCap * P->nS) * sizeof(double), hipMemcpyDeviceToDevice, P->stream));
    MCM_HIP(hipStreamSynchronize(P->stream));
    MCM_HIP(hipFree(P->sensors.p));
  }
  P->sensors.p = np;
  P->sensorCap = cap;
  return MCM_OK;
}

template <typename RT>
int record_sensors(mcm_heat_plan* P, int64_t col) {
  if (P->nS == 0) return MCM_OK;
  hipLaunchKernelGGL(heat_sensors<RT>, dim3(1), dim3(256), 0, P->stream, (const RT*)P->T1.p, P->nx, P->nx * P->ny,
                     (const long long*)P->sIdx.p, (const RT*)P->sW.p, (int)P->nS, (double*)P->sensors.p, (long long)col);
  MCM_HIP(hipGetLastError());
  return MCM_OK;
}

// Lane-per-line y (AXIS 1) or z (AXIS 2) sweep, coefficients in LDS when
// they fit.
template <typename RT, int AXIS>
void launch_lane_sweep(mcm_heat_plan* P, const HeatArgs<RT>& a, unsigned g) {
  const bool lc = P->nM * P->nM * 3 <= kLdsCoef;
  if (AXIS == 1) {
    if (lc) hipLaunchKernelGGL((heat_sweep_y<RT, true>), dim3(g), dim3(kHeatTPB), 0, P->stream, a);
    else hipLaunchKernelGGL((heat_sweep_y<RT, false>), dim3(g), dim3(kHeatTPB), 0, P->stream, a);
  } else {
    if (lc) hipLaunchKernelGGL((heat_sweep_z<RT, true>), dim3(g), dim3(kHeatTPB), 0, P->stream, a);
    else hipLaunchKernelGGL((heat_sweep_z<RT, false>), dim3(g), dim3(kHeatTPB), 0, P->stream, a);
  }
}

// x-sweep launch: the wavefront solver for FP64 lines of up to 512 voxels,
// else the sequential lane-per-line kernel (FP32 keeps the reference's
// sequential recurrence).
template <typename RT>
void launch_sweep_x(mcm_heat_plan* P, const HeatArgs<RT>& a, unsigned gx) {
  if constexpr (std::is_same<RT, double>::value) {
    if (wave_rows(P->nx)) {
      const bool lc = P->nM * P->nM * 3 <= kLdsCoef;
      with_wave_shape<true>(wave_shape(P->nx, true), [&](auto sTag, auto wTag) {
        constexpr int SS = decltype(sTag)::value, WW = decltype(wTag)::value;
        const int64_t lines = (int64_t)P->ny * P->nz, perBlock = kHeatTPB / WW;
        const unsigned gw = (unsigned)((lines + perBlock - 1) / perBlock);
        if (lc) hipLaunchKernelGGL((heat_sweep_x_wave<SS, WW, true>), dim3(gw), dim3(kHeatTPB), 0, P->stream, a);
        else hipLaunchKernelGGL((heat_sweep_x_wave<SS, WW, false>), dim3(gw), dim3(kHeatTPB), 0, P->stream, a);
      });
      return;
    }
  }
  if (P->nM * P->nM * 3 <= kLdsCoef) hipLaunchKernelGGL((heat_sweep_x<RT, true>), dim3(gx), dim3(kHeatTPB), 0, P->stream, a);
  else hipLaunchKernelGGL((heat_sweep_x<RT, false>), dim3(gx), dim3(kHeatTPB), 0, P->stream, a);
}

// y and z sweeps: the wavefront solver for FP64 lines of up to 512 voxels,
// else the sequential lane-per-line kernels.
template <int S, int W, int KIND>
void launch_yz_tile(mcm_heat_plan* P, const HeatArgs<double>& a) {
  constexpr int TX = YzTile<S, W>::TX;
  const unsigned g = (unsigned)(((P->nx + TX - 1) / TX) * (KIND == SWEEP_Y ? P->nz : P->ny));
  if (P->nM * P->nM * 3 <= kTileCoef) hipLaunchKernelGGL((heat_sweep_yz_tile<S, W, KIND, true>), dim3(g), dim3(kHeatTPB), 0, P->stream, a);
  else hipLaunchKernelGGL((heat_sweep_yz_tile<S, W, KIND, false>), dim3(g), dim3(kHeatTPB), 0, P->stream, a);
}

template <int KIND>
void launch_yz_wave(mcm_heat_plan* P, const HeatArgs<double>& a, int n) {
  with_wave_shape<false>(wave_shape(n, false), [&](auto sTag, auto wTag) {
    launch_yz_tile<decltype(sTag)::value, decltype(wTag)::value, KIND>(P, a);
  });
}

template <typename RT>
void launch_sweep_yz(mcm_heat_plan* P, const HeatArgs<RT>& a, unsigned gy, unsigned gz) {
  if constexpr (std::is_same<RT, double>::value) {
    if (wave_rows(P->ny)) launch_yz_wave<SWEEP_Y>(P, a, P->ny);
    else launch_lane_sweep<RT, 1>(P, a, gy);
    if (wave_rows(P->nz)) launch_yz_wave<SWEEP_Z>(P, a, P->nz);
    else launch_lane_sweep<RT, 2>(P, a, gz);
    return;
  }
  launch_lane_sweep<RT, 1>(P, a, gy);
  launch_lane_sweep<RT, 2>(P, a, gz);
}

// The kernels of one time step on the plan's stream; swaps the state buffers.
template <typename RT>
void enqueue_step(mcm_heat_plan* P, double dt, int lightsOn, const long long* colBase, long long col) {
  const int tpb = kHeatTPB;
  HeatArgs<RT> a = make_args<RT>(P, dt, lightsOn);
  if (P->nS) {
    // the step's first kernel records the state it starts from
    a.nS = (int)P->nS;
    a.sIdx = (const long long*)P->sIdx.p;
    a.sW = (const RT*)P->sW.p;
    a.sOut = (double*)P->sensors.p;
    a.sColBase = colBase;
    a.sCol = col;
  }
  if (P->explicitScheme) {
    launch_explicit<RT>(P, a, P->explicitZRun);
    (void)tpb;
  } else {
    launch_sweep_x<RT>(P, a, (unsigned)((P->ny * P->nz + tpb - 1) / tpb));
    launch_sweep_yz<RT>(P, a, (unsigned)((P->nx * P->nz + tpb - 1) / tpb), (unsigned)((P->nx * P->ny + tpb - 1) / tpb));
  }
  std::swap(P->T1.p, P->T2.p);
}

// K (even) steps as graph nodes on the plan's stream, which is capturing:
// step k records sensor column (device step counter) + k, and the last node
// advances the counter by K.  An even K leaves T1/T2 where they were.
template <typename RT>
void enqueue_graph_steps(mcm_heat_plan* P, int64_t K, double dt, int lightsOn) {
  for (int64_t k = 0; k < K; ++k) enqueue_step<RT>(P, dt, lightsOn, (const long long*)P->stepDev.p, (long long)k);
  if (P->nS) hipLaunchKernelGGL(heat_step_counter, dim3(1), dim3(64), 0, P->stream, (long long*)P->stepDev.p, (long long)K, 1);
}

// The cached graph of K steps at (dt, lightsOn) for the current buffers, or
// a new capture (at most four are kept).
template <typename RT>
int step_graph(mcm_heat_plan* P, int64_t K, double dt, int lightsOn, hipGraphExec_t* out) {
  for (HeatGraph& g : P->graphs)
    if (g.steps == K && g.dt == dt && g.lightsOn == lightsOn && g.T1 == P->T1.p && g.sensors == P->sensors.p) {
      *out = g.exec;
      return MCM_OK;
    }
  if (P->nS && !P->stepDev.p) MCM_HIP(hipMalloc(&P->stepDev.p, sizeof(long long)));
  MCM_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
  enqueue_graph_steps<RT>(P, K, dt, lightsOn);
  hipGraph_t graph = nullptr;
  const hipError_t launchErr = hipGetLastError();
  const hipError_t endErr = hipStreamEndCapture(P->stream, &graph);
  if (launchErr != hipSuccess || endErr != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(MCM_ERR_DEVICE, "capturing %lld heat steps: %s", (long long)K, hipGetErrorString(launchErr != hipSuccess ? launchErr : endErr));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t instErr = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (instErr != hipSuccess) return fail(MCM_ERR_DEVICE, "instantiating the heat step graph: %s", hipGetErrorString(instErr));
  if (P->graphs.size() >= 4) {
    // the oldest graph may still be queued or running on the stream: an
    // executable graph is destroyed only once its launches have finished
    MCM_HIP(hipStreamSynchronize(P->stream));
    (void)hipGraphExecDestroy(P->graphs.front().exec);
    P->graphs.erase(P->graphs.begin());
  }
  HeatGraph g;
  g.exec = exec;
  g.steps = K;
  g.dt = dt;
  g.lightsOn = lightsOn;
  g.T1 = P->T1.p;
  g.sensors = P->sensors.p;
  P->graphs.push_back(g);
  *out = exec;
  return MCM_OK;
}

template <typename RT>
int advance(mcm_heat_plan* P, int64_t steps, double dt, int lightsOn) {
  int rc = ensure_sensor_capacity(P, P->stepsDone + steps + 1);
  if (rc) return rc;
  int64_t n = 0;
  // runs of K steps replay one captured graph (one launch instead of 2-4
  // per step); the interrupt is polled between replays
  const int64_t K = P->graphSteps & ~(int64_t)1;
  if (K >= 2 && steps >= 2 * K) {
    hipGraphExec_t g = nullptr;
    if ((rc = step_graph<RT>(P, K, dt, lightsOn, &g))) return rc;
    if (P->nS) hipLaunchKernelGGL(heat_step_counter, dim3(1), dim3(64), 0, P->stream, (long long*)P->stepDev.p, (long long)P->stepsDone, 0);
    for (; steps - n >= K; n += K) {
      if (P->interrupt && P->interrupt()) {
        P->interrupted = true;
        return MCM_OK;
      }
      MCM_HIP(hipGraphLaunch(g, P->stream));
      P->stepsDone += K;
      P->sensorsRecorded = P->stepsDone;
    }
  }
  for (; n < steps; ++n) {
    if (P->interrupt && P->interrupt()) {
      P->interrupted = true;
      break;
    }
    enqueue_step<RT>(P, dt, lightsOn, nullptr, P->stepsDone);
    MCM_HIP(hipGetLastError());
    ++P->stepsDone;
    P->sensorsRecorded = P->stepsDone;
  }
  return MCM_OK;
}

template <typename RT>
int upload_state(mcm_heat_plan* P, const mcm_heat_input* in) {
  const int tpb = 256;
  const unsigned nb = (unsigned)((P->L + tpb - 1) / tpb);
  // stage the host arrays through device buffers, then convert on the device
  if (in->T_f64) {
    MCM_HIP(hipMemcpyAsync(P->stage64.p, in->T_f64, (size_t)P->L * 8, hipMemcpyHostToDevice, P->stream));
    hipLaunchKernelGGL(heat_load<RT>, dim3(nb), dim3(tpb), 0, P->stream, (RT*)P->T1.p, nullptr, (const double*)P->stage64.p, P->L);
  } else {
    MCM_HIP(hipMemcpyAsync(P->stage32.p, in->T, (size_t)P->L * 4, hipMemcpyHostToDevice, P->stream));
    hipLaunchKernelGGL(heat_load<RT>, dim3(nb), dim3(tpb), 0, P->stream, (RT*)P->T1.p, (const float*)P->stage32.p, nullptr, P->L);
  }
  MCM_HIP(hipStreamSynchronize(P->stream));
  if (P->calcDamage) {
    if (in->Omega_f64) {
      MCM_HIP(hipMemcpyAsync(P->stage64.p, in->Omega_f64, (size_t)P->L * 8, hipMemcpyHostToDevice, P->stream));
      hipLaunchKernelGGL(heat_load<RT>, dim3(nb), dim3(tpb), 0, P->stream, (RT*)P->Om.p, nullptr, (const double*)P->stage64.p, P->L);
    } else {
      MCM_HIP(hipMemcpyAsync(P->stage32.p, in->Omega, (size_t)P->L * 4, hipMemcpyHostToDevice, P->stream));
      hipLaunchKernelGGL(heat_load<RT>, dim3(nb), dim3(tpb), 0, P->stream, (RT*)P->Om.p, (const float*)P->stage32.p, nullptr, P->L);
    }
  }
  MCM_HIP(hipGetLastError());
  MCM_HIP(hipStreamSynchronize(P->stream));
  return MCM_OK;
}

template <typename RT>
int download(mcm_heat_plan* P, mcm_heat_output* out) {
  const int tpb = 256;
  const unsigned nb = (unsigned)((P->L + tpb - 1) / tpb);
  MCM_HIP(hipStreamSynchronize(P->stream));
  if (out->T || out->T_f64) {
    hipLaunchKernelGGL(heat_store<RT>, dim3(nb), dim3(tpb), 0, P->stream, (const RT*)P->T1.p, (float*)P->stage32.p,
                       (double*)(out->T_f64 ? P->stage64.p : nullptr), P->L);
    MCM_HIP(hipGetLastError());
    if (out->T) MCM_HIP(hipMemcpyAsync(out->T, P->stage32.p, (size_t)P->L * 4, hipMemcpyDeviceToHost, P->stream));
    if (out->T_f64) MCM_HIP(hipMemcpyAsync(out->T_f64, P->stage64.p, (size_t)P->L * 8, hipMemcpyDeviceToHost, P->stream));
    MCM_HIP(hipStreamSynchronize(P->stream));
  }
  if (P->calcDamage && (out->Omega || out->Omega_f64)) {
    hipLaunchKernelGGL(heat_store<RT>, dim3(nb), dim3(tpb), 0, P->stream, (const RT*)P->Om.p, (float*)P->stage32.p,
                       (double*)(out->Omega_f64 ? P->stage64.p : nullptr), P->L);
    MCM_HIP(hipGetLastError());
    if (out->Omega) MCM_HIP(hipMemcpyAsync(out->Omega, P->stage32.p, (size_t)P->L * 4, hipMemcpyDeviceToHost, P->stream));
    if (out->Omega_f64) MCM_HIP(hipMemcpyAsync(out->Omega_f64, P->stage64.p, (size_t)P->L * 8, hipMemcpyDeviceToHost, P->stream));
    MCM_HIP(hipStreamSynchronize(P->stream));
  } else if (!P->calcDamage && out->Omega) {
    out->Omega[0] = NAN;
  }
  if (P->nS && P->sensorsRecorded <= P->stepsDone) {
    // the current state's column (steps record the state they start from)
    int rc = ensure_sensor_capacity(P, P->stepsDone + 1);
    if (rc) return rc;
    if ((rc = record_sensors<RT>(P, P->stepsDone))) return rc;
    P->sensorsRecorded = P->stepsDone + 1;
  }
  if (out->sensorTemps && P->nS) {
    MCM_HIP(hipMemcpyAsync(out->sensorTemps, P->sensors.p, (size_t)((P->stepsDone + 1) * P->nS) * sizeof(double), hipMemcpyDeviceToHost,
                           P->stream));
  }
  if (out->maxMediaTemps) {
    std::vector<unsigned long long> keys((size_t)P->nM);
    MCM_HIP(hipMemcpyAsync(keys.data(), P->maxKey.p, keys.size() * 8, hipMemcpyDeviceToHost, P->stream));
    MCM_HIP(hipStreamSynchronize(P->stream));
    for (int m = 0; m < P->nM; ++m) out->maxMediaTemps[m] = (float)key_value(keys[m]);
  }
  MCM_HIP(hipStreamSynchronize(P->stream));
  out->stepsDone = P->stepsDone;
  return MCM_OK;
}

// A_d^T solve along axis AXIS: the wavefront solver for lines of up to 512
// voxels, else the lane-per-line recurrence.
template <int AXIS>
void adj_solve(mcm_heat_plan* P, const HeatArgs<double>& a, const double* r, double* y, double* beta, unsigned gLane) {
  const int n = AXIS == 0 ? P->nx : AXIS == 1 ? P->ny : P->nz;
  const int64_t lines = AXIS == 0 ? (int64_t)P->ny * P->nz : AXIS == 1 ? (int64_t)P->nx * P->nz : (int64_t)P->nx * P->ny;
  const bool lc = P->nM * P->nM * 3 <= kTileCoef;
  if (!wave_rows(n)) {
    hipLaunchKernelGGL(heat_adj_solve<AXIS>, dim3(gLane), dim3(kHeatTPB), 0, P->stream, a, r, y, beta);
    return;
  }
  if constexpr (AXIS != 0) {
    // y and z lines through LDS tiles, in the shapes of the forward sweeps
    with_wave_shape<false>(wave_shape(n, false), [&](auto sTag, auto wTag) {
      constexpr int S = decltype(sTag)::value, W = decltype(wTag)::value;
      constexpr int TX = YzTile<S, W>::TX;
      const unsigned g = (unsigned)(((P->nx + TX - 1) / TX) * (AXIS == 1 ? P->nz : P->ny));
      if (lc) hipLaunchKernelGGL((heat_adj_solve_tile<S, W, AXIS == 1 ? 1 : 2, true>), dim3(g), dim3(kHeatTPB), 0, P->stream, a, r, y);
      else hipLaunchKernelGGL((heat_adj_solve_tile<S, W, AXIS == 1 ? 1 : 2, false>), dim3(g), dim3(kHeatTPB), 0, P->stream, a, r, y);
    });
  } else {
    with_wave_shape<true>(wave_shape(n, true), [&](auto sTag, auto wTag) {
      constexpr int S = decltype(sTag)::value, W = decltype(wTag)::value;
      const unsigned gw = (unsigned)((lines + kHeatTPB / W - 1) / (kHeatTPB / W));
      if (lc) hipLaunchKernelGGL((heat_adj_solve_wave<S, W, AXIS, true>), dim3(gw), dim3(kHeatTPB), 0, P->stream, a, r, y);
      else hipLaunchKernelGGL((heat_adj_solve_wave<S, W, AXIS, false>), dim3(gw), dim3(kHeatTPB), 0, P->stream, a, r, y);
    });
  }
}

}  // namespace

extern "C" {

int mcm_heat_plan_create(const mcm_heat_input* in, mcm_heat_plan** out) {
  if (!in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (in->nx < 1 || in->ny < 1 || in->nz < 1) return fail(MCM_ERR_INVALID_ARGUMENT, "grid dimensions must be positive");
  if (in->nx * in->ny * in->nz >= ((int64_t)1 << 31)) return fail(MCM_ERR_INVALID_ARGUMENT, "grid too large for 32-bit voxel indices");
  if (in->nM < 1 || in->nM > 256) return fail(MCM_ERR_INVALID_ARGUMENT, "number of media must be 1..256");
  if (!in->T || !in->M || !in->dTdtperdeltaT || !in->dTdt_abs || !in->A || !in->E)
    return fail(MCM_ERR_INVALID_ARGUMENT, "missing T, M, dTdtperdeltaT, dTdt_abs, A or E");
  if (in->steps < 0) return fail(MCM_ERR_INVALID_ARGUMENT, "steps must be non-negative");
  if (in->precision != MCM_HEAT_FP64 && in->precision != MCM_HEAT_FP32) return fail(MCM_ERR_INVALID_ARGUMENT, "unknown precision");
  if (in->scheme != MCM_HEAT_DGADI && in->scheme != MCM_HEAT_EXPLICIT) return fail(MCM_ERR_INVALID_ARGUMENT, "unknown scheme");
  if (in->nSensors < 0 || (in->nSensors > 0 && (!in->tempSensorCornerIdxs || !in->tempSensorInterpWeights)))
    return fail(MCM_ERR_INVALID_ARGUMENT, "sensor arrays missing");
  const int64_t L = in->nx * in->ny * in->nz;
  for (int64_t k = 0; k < L; ++k)
    if (in->M[k] >= in->nM) return fail(MCM_ERR_INVALID_ARGUMENT, "P.M holds a medium index outside 0..nM-1");
  std::vector<long long> sidx((size_t)in->nSensors);
  for (int64_t s = 0; s < in->nSensors; ++s) {
    const double v = in->tempSensorCornerIdxs[s];
    const long long k = (long long)v;
    // the 8 interpolation corners must lie in the grid
    if (!(v >= 0) || k + 1 + in->nx + in->nx * in->ny >= L) return fail(MCM_ERR_INVALID_ARGUMENT, "sensor corner index out of range");
    sidx[s] = k;
  }
  int rc = use_device(in->GPUdevice);
  if (rc) return rc;
  std::unique_ptr<mcm_heat_plan> P(new mcm_heat_plan());
  P->device = in->GPUdevice;
  P->fp32 = in->precision == MCM_HEAT_FP32;
  P->explicitScheme = in->scheme == MCM_HEAT_EXPLICIT;
  P->nx = (int)in->nx;
  P->ny = (int)in->ny;
  P->nz = (int)in->nz;
  P->nM = (int)in->nM;
  P->L = L;
  P->sinked = in->heatBoundaryType != 0;
  P->calcDamage = in->Omega != nullptr || in->Omega_f64 != nullptr;
  P->nS = in->nSensors;
  P->interrupt = in->isInterruptPending;
  MCM_HIP(hipStreamCreate(&P->stream));
  // default steps per captured graph (0: launch every step, e.g. under a
  // kernel tracer that does not follow graph launches)
  if (const char* e = std::getenv("MCM_HEAT_GRAPH_STEPS")) P->graphSteps = std::max<int64_t>(0, std::atoll(e));
  const size_t rs = rt_size(P.get());
  if ((rc = halloc(P->M, (size_t)L)) || (rc = halloc(P->Q, (size_t)L * 4)) || (rc = halloc(P->c, (size_t)P->nM * P->nM * 3 * rs)) ||
      (rc = halloc(P->logA, (size_t)P->nM * rs)) || (rc = halloc(P->E, (size_t)P->nM * rs)) || (rc = halloc(P->T1, (size_t)L * rs)) ||
      (rc = halloc(P->T2, (size_t)L * rs)) || (rc = halloc(P->B, (size_t)L * rs)) ||
      (rc = halloc(P->Om, (size_t)(P->calcDamage ? L : 1) * rs)) || (rc = halloc(P->maxKey, (size_t)P->nM * 8)) ||
      (rc = halloc(P->sIdx, (size_t)std::max<int64_t>(P->nS, 1) * 8)) || (rc = halloc(P->sW, (size_t)std::max<int64_t>(P->nS, 1) * 3 * rs)) ||
      (rc = halloc(P->stage32, (size_t)L * 4)) || (rc = halloc(P->stage64, (size_t)L * 8)))
    return rc;
  hipStream_t s = P->stream;
  MCM_HIP(hipMemcpyAsync(P->M.p, in->M, (size_t)L, hipMemcpyHostToDevice, s));
  MCM_HIP(hipMemcpyAsync(P->Q.p, in->dTdt_abs, (size_t)L * 4, hipMemcpyHostToDevice, s));
  // per-medium constants in the working precision (the reference casts
  // log(A) and E to float, :711-718)
  const int nM = P->nM;
  std::vector<double> c64((size_t)nM * nM * 3), la64(nM), e64(nM), w64((size_t)std::max<int64_t>(P->nS, 1) * 3);
  std::vector<float> c32(c64.size()), la32(nM), e32(nM), w32(w64.size());
  for (size_t k = 0; k < c64.size(); ++k) {
    c64[k] = (double)in->dTdtperdeltaT[k];
    c32[k] = in->dTdtperdeltaT[k];
  }
  for (int m = 0; m < nM; ++m) {
    la64[m] = std::log(in->A[m]);
    la32[m] = (float)std::log(in->A[m]);
    e64[m] = in->E[m];
    e32[m] = (float)in->E[m];
  }
  for (int64_t k = 0; k < P->nS * 3; ++k) {
    w64[k] = in->tempSensorInterpWeights[k];
    w32[k] = (float)in->tempSensorInterpWeights[k];
  }
  if (P->fp32) {
    MCM_HIP(hipMemcpyAsync(P->c.p, c32.data(), c32.size() * 4, hipMemcpyHostToDevice, s));
    MCM_HIP(hipMemcpyAsync(P->logA.p, la32.data(), nM * 4, hipMemcpyHostToDevice, s));
    MCM_HIP(hipMemcpyAsync(P->E.p, e32.data(), nM * 4, hipMemcpyHostToDevice, s));
    MCM_HIP(hipMemcpyAsync(P->sW.p, w32.data(), w32.size() * 4, hipMemcpyHostToDevice, s));
    P->constR = (double)8.3144598f;   // finiteElementHeatPropagator.c:86-87
    P->constC = (double)273.15f;
  } else {
    MCM_HIP(hipMemcpyAsync(P->c.p, c64.data(), c64.size() * 8, hipMemcpyHostToDevice, s));
    MCM_HIP(hipMemcpyAsync(P->logA.p, la64.data(), nM * 8, hipMemcpyHostToDevice, s));
    MCM_HIP(hipMemcpyAsync(P->E.p, e64.data(), nM * 8, hipMemcpyHostToDevice, s));
    MCM_HIP(hipMemcpyAsync(P->sW.p, w64.data(), w64.size() * 8, hipMemcpyHostToDevice, s));
    P->constR = 8.3144598;
    P->constC = 273.15;
  }
  if (P->nS) MCM_HIP(hipMemcpyAsync(P->sIdx.p, sidx.data(), sidx.size() * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(heat_init_keys, dim3(1), dim3(256), 0, s, (unsigned long long*)P->maxKey.p, nM);
  MCM_HIP(hipGetLastError());
  MCM_HIP(hipStreamSynchronize(s));
  rc = P->fp32 ? upload_state<float>(P.get(), in) : upload_state<double>(P.get(), in);
  if (rc) return rc;
  if ((rc = ensure_sensor_capacity(P.get(), std::max<int64_t>(in->steps, 1) + 1))) return rc;
  if (P->explicitScheme && (rc = P->fp32 ? tune_explicit<float>(P.get(), in->steps) : tune_explicit<double>(P.get(), in->steps))) return rc;
  MCM_HIP(hipStreamSynchronize(s));
  *out = P.release();
  return MCM_OK;
}

int mcm_heat_plan_advance(mcm_heat_plan* P, int64_t steps, double dt, int32_t lightsOn) {
  if (!P) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  if (steps < 0) return fail(MCM_ERR_INVALID_ARGUMENT, "steps must be non-negative");
  MCM_HIP(hipSetDevice(P->device));
  return P->fp32 ? advance<float>(P, steps, dt, lightsOn) : advance<double>(P, steps, dt, lightsOn);
}

int mcm_heat_group_create(mcm_heat_plan* const* plans, int64_t nPlans, mcm_heat_group** out) {
  if (!plans || !out || nPlans < 1) return fail(MCM_ERR_INVALID_ARGUMENT, "a group needs at least one plan");
  *out = nullptr;
  for (int64_t i = 0; i < nPlans; ++i) {
    if (!plans[i]) return fail(MCM_ERR_INVALID_ARGUMENT, "plan %lld is null", (long long)i);
    if (plans[i]->device != plans[0]->device) return fail(MCM_ERR_INVALID_ARGUMENT, "the plans of a group share one device");
    for (int64_t j = 0; j < i; ++j)
      if (plans[j] == plans[i]) return fail(MCM_ERR_INVALID_ARGUMENT, "plan %lld appears twice", (long long)i);
  }
  MCM_HIP(hipSetDevice(plans[0]->device));
  std::unique_ptr<mcm_heat_group> G(new mcm_heat_group);
  G->device = plans[0]->device;
  G->plans.assign(plans, plans + nPlans);
  G->events.assign((size_t)nPlans + 1, nullptr);
  for (hipEvent_t& e : G->events) MCM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *out = G.release();
  return MCM_OK;
}

int mcm_heat_group_advance(mcm_heat_group* G, int64_t steps, const double* dt, const int32_t* lightsOn, int64_t stepsPerGraph) {
  if (!G || !dt || !lightsOn) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  if (steps < 0 || stepsPerGraph < 0) return fail(MCM_ERR_INVALID_ARGUMENT, "steps must be non-negative");
  const size_t n = G->plans.size();
  mcm_heat_plan* P0 = G->plans[0];
  MCM_HIP(hipSetDevice(P0->device));
  for (size_t i = 0; i < n; ++i) {
    int rc = ensure_sensor_capacity(G->plans[i], G->plans[i]->stepsDone + steps + 1);
    if (rc) return rc;
    if (G->plans[i]->nS && !G->plans[i]->stepDev.p) MCM_HIP(hipMalloc(&G->plans[i]->stepDev.p, sizeof(long long)));
  }
  const int64_t K = stepsPerGraph & ~(int64_t)1;
  int64_t done = 0;
  if (K >= 2 && steps >= K) {
    mcm_heat_group::Graph* g = nullptr;
    for (mcm_heat_group::Graph& c : G->graphs) {
      bool same = c.steps == K;
      for (size_t i = 0; same && i < n; ++i)
        same = c.dt[i] == dt[i] && c.lightsOn[i] == (lightsOn[i] != 0) && c.T1[i] == G->plans[i]->T1.p && c.sensors[i] == G->plans[i]->sensors.p;
      if (same) g = &c;
    }
    if (!g) {
      // fork from P0's stream into every plan's stream, K steps per plan, join
      MCM_HIP(hipStreamBeginCapture(P0->stream, hipStreamCaptureModeThreadLocal));
      hipError_t err = hipEventRecord(G->events[0], P0->stream);
      for (size_t i = 1; i < n && err == hipSuccess; ++i) err = hipStreamWaitEvent(G->plans[i]->stream, G->events[0], 0);
      for (size_t i = 0; i < n && err == hipSuccess; ++i) {
        mcm_heat_plan* P = G->plans[i];
        if (P->fp32) enqueue_graph_steps<float>(P, K, dt[i], lightsOn[i] != 0);
        else enqueue_graph_steps<double>(P, K, dt[i], lightsOn[i] != 0);
        err = hipGetLastError();
      }
      for (size_t i = 1; i < n && err == hipSuccess; ++i) {
        err = hipEventRecord(G->events[i + 1], G->plans[i]->stream);
        if (err == hipSuccess) err = hipStreamWaitEvent(P0->stream, G->events[i + 1], 0);
      }
      hipGraph_t graph = nullptr;
      const hipError_t endErr = hipStreamEndCapture(P0->stream, &graph);
      if (err != hipSuccess || endErr != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(MCM_ERR_DEVICE, "capturing the group's steps: %s", hipGetErrorString(err != hipSuccess ? err : endErr));
      }
      hipGraphExec_t exec = nullptr;
      err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (err != hipSuccess) return fail(MCM_ERR_DEVICE, "instantiating the group's step graph: %s", hipGetErrorString(err));
      if (G->graphs.size() >= 4) {
        // replays of the oldest graph end on P0's stream (every branch joins
        // it there): wait for them before the executable graph is destroyed
        MCM_HIP(hipStreamSynchronize(P0->stream));
        (void)hipGraphExecDestroy(G->graphs.front().exec);
        G->graphs.erase(G->graphs.begin());
      }
      mcm_heat_group::Graph c;
      c.exec = exec;
      c.steps = K;
      for (size_t i = 0; i < n; ++i) {
        c.dt.push_back(dt[i]);
        c.lightsOn.push_back(lightsOn[i] != 0);
        c.T1.push_back(G->plans[i]->T1.p);
        c.sensors.push_back(G->plans[i]->sensors.p);
      }
      G->graphs.push_back(std::move(c));
      g = &G->graphs.back();
    }
    // order the replays after every plan's earlier work; set the step counters
    for (size_t i = 1; i < n; ++i) {
      MCM_HIP(hipEventRecord(G->events[i + 1], G->plans[i]->stream));
      MCM_HIP(hipStreamWaitEvent(P0->stream, G->events[i + 1], 0));
    }
    for (size_t i = 0; i < n; ++i)
      if (G->plans[i]->nS)
        hipLaunchKernelGGL(heat_step_counter, dim3(1), dim3(64), 0, P0->stream, (long long*)G->plans[i]->stepDev.p,
                           (long long)G->plans[i]->stepsDone, 0);
    for (; steps - done >= K; done += K) {
      if (P0->interrupt && P0->interrupt()) {
        P0->interrupted = true;
        break;
      }
      MCM_HIP(hipGraphLaunch(g->exec, P0->stream));
      for (size_t i = 0; i < n; ++i) {
        G->plans[i]->stepsDone += K;
        G->plans[i]->sensorsRecorded = G->plans[i]->stepsDone;
      }
    }
    // later work on any plan's stream follows the replays
    MCM_HIP(hipEventRecord(G->events[0], P0->stream));
    for (size_t i = 1; i < n; ++i) MCM_HIP(hipStreamWaitEvent(G->plans[i]->stream, G->events[0], 0));
    if (P0->interrupted) return MCM_OK;
  }
  for (size_t i = 0; i < n && done < steps; ++i) {
    mcm_heat_plan* P = G->plans[i];
    const int rc = P->fp32 ? advance<float>(P, steps - done, dt[i], lightsOn[i] != 0) : advance<double>(P, steps - done, dt[i], lightsOn[i] != 0);
    if (rc) return rc;
  }
  return MCM_OK;
}

int mcm_heat_group_destroy(mcm_heat_group* G) {
  if (!G) return MCM_OK;
  // the plans may already be gone: wait for the device, not their streams
  (void)hipSetDevice(G->device);
  (void)hipDeviceSynchronize();
  delete G;
  return MCM_OK;
}

int mcm_heat_plan_set_explicit_zrun(mcm_heat_plan* P, int32_t zPlanes) {
  if (!P) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  if (!P->explicitScheme) return fail(MCM_ERR_INVALID_ARGUMENT, "the plan does not use the explicit scheme");
  if (!valid_zrun(zPlanes)) return fail(MCM_ERR_INVALID_ARGUMENT, "z planes per workgroup must be 8, 12, 16, 24 or 32");
  MCM_HIP(hipSetDevice(P->device));
  MCM_HIP(hipStreamSynchronize(P->stream));
  P->drop_graphs();  // captured steps hold the old launch shape
  P->explicitZRun = zPlanes;
  return MCM_OK;
}

int32_t mcm_heat_plan_explicit_zrun(const mcm_heat_plan* P) { return P ? P->explicitZRun : 0; }

int mcm_heat_plan_set_graph_steps(mcm_heat_plan* P, int64_t steps) {
  if (!P) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  if (steps < 0) return fail(MCM_ERR_INVALID_ARGUMENT, "steps per graph must be non-negative");
  MCM_HIP(hipSetDevice(P->device));
  MCM_HIP(hipStreamSynchronize(P->stream));
  P->drop_graphs();
  P->graphSteps = steps;
  return MCM_OK;
}

int mcm_heat_plan_download(mcm_heat_plan* P, mcm_heat_output* out) {
  if (!P || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  MCM_HIP(hipSetDevice(P->device));
  return P->fp32 ? download<float>(P, out) : download<double>(P, out);
}

int mcm_heat_plan_destroy(mcm_heat_plan* P) {
  if (!P) return MCM_OK;
  (void)hipSetDevice(P->device);
  (void)hipStreamSynchronize(P->stream);
  delete P;
  return MCM_OK;
}


int mcm_heat_adjoint_run(const mcm_heat_input* in, const mcm_heat_segment* segments, int64_t nSegments, const uint8_t* targetMask,
                         int64_t checkpointEvery, mcm_heat_adjoint_output* out) {
  if (!in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  if (in->precision != MCM_HEAT_FP64 || in->scheme != MCM_HEAT_DGADI)
    return fail(MCM_ERR_UNSUPPORTED, "the adjoint is implemented for the FP64 DG-ADI scheme");
  if (!in->Omega && !in->Omega_f64) return fail(MCM_ERR_INVALID_ARGUMENT, "the objective is the damage integral: Omega is required");
  if (nSegments < 0 || (nSegments > 0 && !segments)) return fail(MCM_ERR_INVALID_ARGUMENT, "segments missing");
  std::vector<double> dts;
  std::vector<int> lights;
  if (nSegments == 0) {
    dts.assign((size_t)std::max<int64_t>(in->steps, 0), in->dt);
    lights.assign(dts.size(), in->lightsOn != 0);
  }
  for (int64_t k = 0; k < nSegments; ++k) {
    if (segments[k].steps < 0) return fail(MCM_ERR_INVALID_ARGUMENT, "segment steps must be non-negative");
    dts.insert(dts.end(), (size_t)segments[k].steps, segments[k].dt);
    lights.insert(lights.end(), (size_t)segments[k].steps, segments[k].lightsOn != 0);
  }
  const int64_t N = (int64_t)dts.size();
  const int64_t C = checkpointEvery > 0 ? checkpointEvery : 100;
  mcm_heat_plan* raw = nullptr;
  int rc = mcm_heat_plan_create(in, &raw);
  if (rc) return rc;
  std::unique_ptr<mcm_heat_plan, int (*)(mcm_heat_plan*)> PP(raw, mcm_heat_plan_destroy);
  mcm_heat_plan* P = PP.get();
  const int64_t L = P->L;
  const size_t bytes = (size_t)L * sizeof(double);
  hipStream_t s = P->stream;
  // target voxels: the mask, or every voxel whose medium has E != 0
  std::vector<uint8_t> mask((size_t)L);
  for (int64_t i = 0; i < L; ++i) mask[i] = targetMask ? (targetMask[i] != 0) : (in->E[in->M[i]] != 0);
  const int64_t nChk = std::max<int64_t>(1, (N + C - 1) / C);
  HeatBuf dMask, chk;
  if ((rc = halloc(dMask, (size_t)L)) || (rc = halloc(chk, (size_t)nChk * bytes))) return rc;
  MCM_HIP(hipMemcpyAsync(dMask.p, mask.data(), (size_t)L, hipMemcpyHostToDevice, s));
  auto chkp = [&](int64_t k) { return (double*)chk.p + (size_t)k * L; };
  // forward pass with checkpoints every C steps
  for (int64_t n = 0; n < N; ++n) {
    if (n % C == 0) MCM_HIP(hipMemcpyAsync(chkp(n / C), P->T1.p, bytes, hipMemcpyDeviceToDevice, s));
    if ((rc = advance<double>(P, 1, dts[(size_t)n], lights[(size_t)n]))) return rc;
  }
  if (N == 0) MCM_HIP(hipMemcpyAsync(chkp(0), P->T1.p, bytes, hipMemcpyDeviceToDevice, s));
  std::vector<double> om((size_t)L);
  MCM_HIP(hipMemcpyAsync(om.data(), P->Om.p, bytes, hipMemcpyDeviceToHost, s));
  MCM_HIP(hipStreamSynchronize(s));
  double J = 0;
  for (int64_t i = 0; i < L; ++i)
    if (mask[i]) J += om[i];
  out->J = J;
  if (out->Omega_f64) std::memcpy(out->Omega_f64, om.data(), bytes);
  if (!out->dJ_dQ && !out->dJ_dT0 && !out->dJ_dc) return MCM_OK;
  if (P->nz > 65535) return fail(MCM_ERR_UNSUPPORTED, "heat adjoint: nz > 65535 z planes");
  // reverse sweep, one checkpoint segment at a time
  HeatBuf seg, lam, lamN, mu1, mu2, mu3, dQ, s1, s2, dC;
  const int nBins = P->nM * P->nM * 3;
  if ((rc = halloc(seg, (size_t)(C + 1) * bytes)) || (rc = halloc(lam, bytes)) || (rc = halloc(lamN, bytes)) || (rc = halloc(mu1, bytes)) ||
      (rc = halloc(mu2, bytes)) || (rc = halloc(mu3, bytes)) || (rc = halloc(dQ, bytes)))
    return rc;
  if (out->dJ_dc) {
    // the step's intermediate states and the coefficient sums
    if ((rc = halloc(s1, bytes)) || (rc = halloc(s2, bytes)) || (rc = halloc(dC, (size_t)nBins * sizeof(double)))) return rc;
    MCM_HIP(hipMemsetAsync(dC.p, 0, (size_t)nBins * sizeof(double), s));
  }
  MCM_HIP(hipMemsetAsync(lam.p, 0, bytes, s));
  MCM_HIP(hipMemsetAsync(dQ.p, 0, bytes, s));
  auto segp = [&](int64_t k) { return (double*)seg.p + (size_t)k * L; };
  const int tpb = kHeatTPB;
  const unsigned gv = (unsigned)((L + tpb - 1) / tpb);
  const unsigned gx = (unsigned)(((int64_t)P->ny * P->nz + tpb - 1) / tpb);
  const unsigned gy = (unsigned)(((int64_t)P->nx * P->nz + tpb - 1) / tpb);
  const unsigned gz = (unsigned)(((int64_t)P->nx * P->ny + tpb - 1) / tpb);
  for (int64_t k = (N + C - 1) / C - 1; k >= 0; --k) {
    const int64_t s0 = k * C, e0 = std::min(N, s0 + C);
    MCM_HIP(hipMemcpyAsync(P->T1.p, chkp(k), bytes, hipMemcpyDeviceToDevice, s));
    MCM_HIP(hipMemcpyAsync(segp(0), chkp(k), bytes, hipMemcpyDeviceToDevice, s));
    for (int64_t n = s0; n < e0; ++n) {
      if ((rc = advance<double>(P, 1, dts[(size_t)n], lights[(size_t)n]))) return rc;
      MCM_HIP(hipMemcpyAsync(segp(n - s0 + 1), P->T1.p, bytes, hipMemcpyDeviceToDevice, s));
    }
    for (int64_t n = e0 - 1; n >= s0; --n) {
      HeatArgs<double> a = make_args<double>(P, dts[(size_t)n], lights[(size_t)n]);
      const double* x0 = segp(n - s0);
      const double* x1 = segp(n - s0 + 1);
      double* B = (double*)P->B.p;
      hipLaunchKernelGGL(heat_adj_damage, dim3(gv), dim3(tpb), 0, s, a, x0, x1, (const uint8_t*)dMask.p, (double*)lam.p, L);
      adj_solve<2>(P, a, (const double*)lam.p, (double*)mu3.p, B, gz);
      adj_solve<1>(P, a, (const double*)mu3.p, (double*)mu2.p, B, gy);
      adj_solve<0>(P, a, (const double*)mu2.p, (double*)mu1.p, B, gx);
      const dim3 gc((unsigned)((P->nx * P->ny + tpb - 1) / tpb), (unsigned)P->nz);
      if (P->nM * P->nM * 3 <= kTileCoef)
        hipLaunchKernelGGL(heat_adj_combine<true>, gc, dim3(tpb), 0, s, a, (const double*)mu1.p, (const double*)mu2.p,
                           (const double*)mu3.p, x0, x1, (const uint8_t*)dMask.p, (double*)lamN.p, (double*)dQ.p, lights[(size_t)n]);
      else
        hipLaunchKernelGGL(heat_adj_combine<false>, gc, dim3(tpb), 0, s, a, (const double*)mu1.p, (const double*)mu2.p,
                           (const double*)mu3.p, x0, x1, (const uint8_t*)dMask.p, (double*)lamN.p, (double*)dQ.p, lights[(size_t)n]);
      MCM_HIP(hipGetLastError());
      if (out->dJ_dc) {
        const dim3 gp((unsigned)((P->nx * P->ny + tpb - 1) / tpb), (unsigned)P->nz);
        hipLaunchKernelGGL(heat_adj_undo<2>, gp, dim3(tpb), 0, s, a, x1, x0, (double*)s2.p);
        hipLaunchKernelGGL(heat_adj_undo<1>, gp, dim3(tpb), 0, s, a, (const double*)s2.p, x0, (double*)s1.p);
        const unsigned gcoef = (unsigned)std::min<int64_t>((L + tpb - 1) / tpb, 2048);
        if (nBins <= kCoefBins)
          hipLaunchKernelGGL(heat_adj_coef<true>, dim3(gcoef), dim3(tpb), 0, s, a, (const double*)mu1.p, (const double*)mu2.p,
                             (const double*)mu3.p, x0, (const double*)s1.p, (const double*)s2.p, x1, (double*)dC.p);
        else
          hipLaunchKernelGGL(heat_adj_coef<false>, dim3(gcoef), dim3(tpb), 0, s, a, (const double*)mu1.p, (const double*)mu2.p,
                             (const double*)mu3.p, x0, (const double*)s1.p, (const double*)s2.p, x1, (double*)dC.p);
        MCM_HIP(hipGetLastError());
      }
      std::swap(lam.p, lamN.p);
    }
  }
  if (out->dJ_dc) MCM_HIP(hipMemcpyAsync(out->dJ_dc, dC.p, (size_t)nBins * sizeof(double), hipMemcpyDeviceToHost, s));
  if (out->dJ_dQ) MCM_HIP(hipMemcpyAsync(out->dJ_dQ, dQ.p, bytes, hipMemcpyDeviceToHost, s));
  if (out->dJ_dT0) MCM_HIP(hipMemcpyAsync(out->dJ_dT0, lam.p, bytes, hipMemcpyDeviceToHost, s));
  MCM_HIP(hipStreamSynchronize(s));
  return MCM_OK;
}

// The whole MEX call (:688-821).  With steps == 0 the reference returns a
// zero-filled T (the step loop never writes its output array); here the input
// state is returned unchanged.
int mcm_heat_run(const mcm_heat_input* in, mcm_heat_output* out) {
  if (!in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  mcm_heat_plan* raw = nullptr;
  int rc = mcm_heat_plan_create(in, &raw);
  if (rc) return rc;
  std::unique_ptr<mcm_heat_plan, int (*)(mcm_heat_plan*)> P(raw, mcm_heat_plan_destroy);
  if ((rc = mcm_heat_plan_advance(P.get(), in->steps, in->dt, in->lightsOn))) return rc;
  return mcm_heat_plan_download(P.get(), out);
}

}  // extern "C"
