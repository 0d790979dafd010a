// Heat diffusion + Arrhenius damage on gfx950: kernels and the C ABI
// (mcm_heat_plan_*, mcm_heat_run).  Replaces finiteElementHeatPropagator.c
// (+MCmatlab/src/finiteElementHeatPropagator.c:688-821).
//
// One Douglas-Gunn ADI step (:767-782) is three line-sweep kernels instead of
// the reference's six launches:
//   sweep_x   substeps 1a + 1b: the explicit right-hand side (full 7-point
//             stencil, x half-weighted, plus the source) built on the fly and
//             Thomas-solved along x;
//   sweep_y   substeps 2a + 2b: the -dt/2 L_y correction and the y solve;
//   sweep_z   substeps 3a + 3b: the -dt/2 L_z correction, the z solve, the
//             Arrhenius damage update and the per-medium maximum temperature.
// Each lane owns one grid line.  For the y and z sweeps lanes run along x, so
// every load and store of a wavefront is one contiguous 64-voxel segment.
// The forward-eliminated diagonal of each line goes to a scratch array in the
// same layout (it is re-read by the same lane right after, from cache).
// The explicit (forward Euler) scheme is one kernel per step: one lane per
// (x, y) column walking z, keeping the z neighbours in registers.
//
// Arithmetic is the reference's expression for expression (built without
// FMA contraction), in FP64 (default) or in FP32 like the reference.
// Per-medium maxima use an order-preserving integer key of the FP64 value:
// each lane keeps a running maximum over its current run of one medium along
// its line and folds it into LDS when the medium changes; each workgroup
// folds LDS into the global keys at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
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
};

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

// substeps 1a + 1b (finiteElementHeatPropagator.c:288-312, :352-389): lane per (y, z) x-line
template <typename RT>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_x(HeatArgs<RT> a) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= a.ny * a.nz) return;
  const int y = line % a.ny, z = line / a.ny;
  const int nx = a.nx, nxy = a.nx * a.ny;
  const int base = y * nx + z * nxy;
  const RT dt = a.dt;
  RT tprev = 0, bprev = 1;
  for (int x = 0; x < nx; ++x) {
    const int i = base + x;
    const int m = a.M[i];
    const bool fz = frozen(a, x, y, z);
    RT dTdt = 0;
    if (!fz) {
      dTdt = a.lightsOn ? (RT)a.Q[i] : (RT)0;
      if (x != 0) dTdt += (a.T1[i - 1] - a.T1[i]) * cc(a, m, a.M[i - 1], 0) / 2;
      if (x != nx - 1) dTdt += (a.T1[i + 1] - a.T1[i]) * cc(a, m, a.M[i + 1], 0) / 2;
      if (y != 0) dTdt += (a.T1[i - nx] - a.T1[i]) * cc(a, m, a.M[i - nx], 1);
      if (y != a.ny - 1) dTdt += (a.T1[i + nx] - a.T1[i]) * cc(a, m, a.M[i + nx], 1);
      if (z != 0) dTdt += (a.T1[i - nxy] - a.T1[i]) * cc(a, m, a.M[i - nxy], 2);
      if (z != a.nz - 1) dTdt += (a.T1[i + nxy] - a.T1[i]) * cc(a, m, a.M[i + nxy], 2);
    }
    RT t = a.T1[i] + dt * dTdt;
    RT b = 1;
    if (!fz) {
      if (x < nx - 1) b += dt / 2 * cc(a, m, a.M[i + 1], 0);
      if (x > 0) {
        const int mp = a.M[i - 1];
        b += dt / 2 * cc(a, m, mp, 0);
        const RT w = -dt / 2 * cc(a, m, mp, 0) / bprev;
        if (x > 1 || !a.sinked) b += w * dt / 2 * cc(a, mp, m, 0);
        t -= w * tprev;
      }
    }
    a.T2[i] = t;
    a.B[i] = b;
    tprev = t;
    bprev = b;
  }
  RT xnext = 0;
  for (int x = nx - 1; x >= 0; --x) {
    const int i = base + x;
    RT t = a.T2[i];
    if (!frozen(a, x, y, z)) {
      t = (t + (x == nx - 1 ? (RT)0 : dt / 2 * cc(a, a.M[i], a.M[i + 1], 0) * xnext)) / a.B[i];
      a.T2[i] = t;
    }
    xnext = t;
  }
}

// substeps 2a + 2b (:429-448, :486-521): lane per (x, z) y-line
template <typename RT>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_y(HeatArgs<RT> a) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= a.nx * a.nz) return;
  const int x = line % a.nx, z = line / a.nx;
  const int nx = a.nx, ny = a.ny, nxy = a.nx * a.ny;
  const int base = x + z * nxy;
  const RT dt = a.dt;
  RT tprev = 0, bprev = 1;
  for (int y = 0; y < ny; ++y) {
    const int i = base + y * nx;
    const int m = a.M[i];
    const bool fz = frozen(a, x, y, z);
    RT t = a.T2[i];
    RT b = 1;
    if (!fz) {
      RT dTdt = 0;
      if (y != 0) dTdt -= (a.T1[i - nx] - a.T1[i]) * cc(a, m, a.M[i - nx], 1) / 2;
      if (y != ny - 1) dTdt -= (a.T1[i + nx] - a.T1[i]) * cc(a, m, a.M[i + nx], 1) / 2;
      t = t + dt * dTdt;
      if (y < ny - 1) b += dt / 2 * cc(a, m, a.M[i + nx], 1);
      if (y > 0) {
        const int mp = a.M[i - nx];
        b += dt / 2 * cc(a, m, mp, 1);
        const RT w = -dt / 2 * cc(a, m, mp, 1) / bprev;
        if (y > 1 || !a.sinked) b += w * dt / 2 * cc(a, mp, m, 1);
        t -= w * tprev;
      }
    }
    a.T2[i] = t;
    a.B[i] = b;
    tprev = t;
    bprev = b;
  }
  RT xnext = 0;
  for (int y = ny - 1; y >= 0; --y) {
    const int i = base + y * nx;
    RT t = a.T2[i];
    if (!frozen(a, x, y, z)) {
      t = (t + (y == ny - 1 ? (RT)0 : dt / 2 * cc(a, a.M[i], a.M[i + nx], 1) * xnext)) / a.B[i];
      a.T2[i] = t;
    }
    xnext = t;
  }
}

// substeps 3a + 3b (:546-565, :606-646) with the Arrhenius update and the
// per-medium maximum: lane per (x, y) z-line
template <typename RT>
__global__ void __launch_bounds__(kHeatTPB) heat_sweep_z(HeatArgs<RT> a) {
  __shared__ unsigned long long sMax[256];
  init_lds_max(sMax, a.nM);
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line < a.nx * a.ny) {
    const int x = line % a.nx, y = line / a.nx;
    const int nz = a.nz, nxy = a.nx * a.ny;
    const int base = line;
    const RT dt = a.dt;
    RT tprev = 0, bprev = 1;
    for (int z = 0; z < nz; ++z) {
      const int i = base + z * nxy;
      const int m = a.M[i];
      const bool fz = frozen(a, x, y, z);
      RT t = a.T2[i];
      RT b = 1;
      if (!fz) {
        RT dTdt = 0;
        if (z != 0) dTdt -= (a.T1[i - nxy] - a.T1[i]) * cc(a, m, a.M[i - nxy], 2) / 2;
        if (z != nz - 1) dTdt -= (a.T1[i + nxy] - a.T1[i]) * cc(a, m, a.M[i + nxy], 2) / 2;
        t = t + dt * dTdt;
        if (z < nz - 1) b += dt / 2 * cc(a, m, a.M[i + nxy], 2);
        if (z > 0) {
          const int mp = a.M[i - nxy];
          b += dt / 2 * cc(a, m, mp, 2);
          const RT w = -dt / 2 * cc(a, m, mp, 2) / bprev;
          if (z > 1 || !a.sinked) b += w * dt / 2 * cc(a, mp, m, 2);
          t -= w * tprev;
        }
      }
      a.T2[i] = t;
      a.B[i] = b;
      tprev = t;
      bprev = b;
    }
    RT xnext = 0;
    MaxRun run;
    for (int z = nz - 1; z >= 0; --z) {
      const int i = base + z * nxy;
      const int m = a.M[i];
      RT t = a.T2[i];
      if (!frozen(a, x, y, z)) {
        t = (t + (z == nz - 1 ? (RT)0 : dt / 2 * cc(a, m, a.M[i + nxy], 2) * xnext)) / a.B[i];
        a.T2[i] = t;
      }
      xnext = t;
      if (a.calcDamage && a.E[m] != 0) a.Om[i] += dt * (RT)exp(a.logA[m] - a.E[m] / (a.Rgas * ((t + a.T1[i]) / 2 + a.czero)));
      run.add(m, (double)t, sMax);
    }
    run.flush(sMax);
  }
  fold_lds_max(sMax, a.maxKey, a.nM);
}

// Forward Euler step (no reference counterpart; same neighbour order and
// damage/maximum update as the DG-ADI step): lane per (x, y) column, z
// neighbours carried in registers.
template <typename RT>
__global__ void __launch_bounds__(kHeatTPB) heat_explicit(HeatArgs<RT> a) {
  __shared__ unsigned long long sMax[256];
  init_lds_max(sMax, a.nM);
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line < a.nx * a.ny) {
    const int x = line % a.nx, y = line / a.nx;
    const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
    const RT dt = a.dt;
    MaxRun run;
    RT tBelow = 0, tHere = a.T1[line];
    int mBelow = 0, mHere = a.M[line];
    for (int z = 0; z < nz; ++z) {
      const int i = line + z * nxy;
      const RT tAbove = z != nz - 1 ? a.T1[i + nxy] : (RT)0;
      const int mAbove = z != nz - 1 ? a.M[i + nxy] : 0;
      RT dTdt = 0;
      if (!frozen(a, x, y, z)) {
        dTdt = a.lightsOn ? (RT)a.Q[i] : (RT)0;
        if (x != 0) dTdt += (a.T1[i - 1] - tHere) * cc(a, mHere, a.M[i - 1], 0);
        if (x != nx - 1) dTdt += (a.T1[i + 1] - tHere) * cc(a, mHere, a.M[i + 1], 0);
        if (y != 0) dTdt += (a.T1[i - nx] - tHere) * cc(a, mHere, a.M[i - nx], 1);
        if (y != ny - 1) dTdt += (a.T1[i + nx] - tHere) * cc(a, mHere, a.M[i + nx], 1);
        if (z != 0) dTdt += (tBelow - tHere) * cc(a, mHere, mBelow, 2);
        if (z != nz - 1) dTdt += (tAbove - tHere) * cc(a, mHere, mAbove, 2);
      }
      const RT t = tHere + dt * dTdt;
      a.T2[i] = t;
      if (a.calcDamage && a.E[mHere] != 0)
        a.Om[i] += dt * (RT)exp(a.logA[mHere] - a.E[mHere] / (a.Rgas * ((t + tHere) / 2 + a.czero)));
      run.add(mHere, (double)t, sMax);
      tBelow = tHere;
      mBelow = mHere;
      tHere = tAbove;
      mHere = mAbove;
    }
    run.flush(sMax);
  }
  fold_lds_max(sMax, a.maxKey, a.nM);
}

// recordTempSensors (:652-663): trilinear interpolation of the new T into
// column `col` of the sensor record
template <typename RT>
__global__ void heat_sensors(const RT* T, int nx, int nxy, const long long* idx, const RT* w, int nS, double* out, long long col) {
  for (int s = threadIdx.x; s < nS; s += blockDim.x) {
    const long long k = idx[s];
    const RT wx = w[s], wy = w[s + nS], wz = w[s + 2 * nS];
    out[s + col * nS] = (double)((1 - wx) * (1 - wy) * (1 - wz) * T[k] + (1 - wx) * (1 - wy) * wz * T[k + nxy] +
                                 (1 - wx) * wy * (1 - wz) * T[k + nx] + (1 - wx) * wy * wz * T[k + nx + nxy] +
                                 wx * (1 - wy) * (1 - wz) * T[k + 1] + wx * (1 - wy) * wz * T[k + 1 + nxy] +
                                 wx * wy * (1 - wz) * T[k + 1 + nx] + wx * wy * wz * T[k + 1 + nx + nxy]);
  }
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

}  // namespace mcm

using namespace mcm;

struct HeatBuf {
  void* p = nullptr;
  ~HeatBuf() {
    if (p) (void)hipFree(p);
  }
};

struct mcm_heat_plan {
  int device = 0;
  hipStream_t stream = nullptr;
  int fp32 = 0, explicitScheme = 0;
  int nx = 0, ny = 0, nz = 0, nM = 0;
  int64_t L = 0;
  int sinked = 0, calcDamage = 0;
  int64_t nS = 0;
  int64_t stepsDone = 0;
  int64_t sensorCap = 0;      // columns allocated in `sensors`
  bool interrupted = false;
  int (*interrupt)(void) = nullptr;
  HeatBuf M, Q, c, logA, E, T1, T2, B, Om, maxKey, sIdx, sW, sensors, stage32, stage64;
  double constR = 0, constC = 0;
  ~mcm_heat_plan() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

size_t rt_size(const mcm_heat_plan* p) { return p->fp32 ? sizeof(float) : sizeof(double); }

int halloc(HeatBuf& b, size_t bytes) {
  MCM_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 8)));
  return MCM_OK;
}

template <typename RT>
HeatArgs<RT> make_args(mcm_heat_plan* P, double dt, int lightsOn) {
  HeatArgs<RT> a;
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

template <typename RT>
int advance(mcm_heat_plan* P, int64_t steps, double dt, int lightsOn) {
  int rc = ensure_sensor_capacity(P, P->stepsDone + steps + 1);
  if (rc) return rc;
  const int tpb = kHeatTPB;
  const unsigned gx = (unsigned)((P->ny * P->nz + tpb - 1) / tpb);
  const unsigned gy = (unsigned)((P->nx * P->nz + tpb - 1) / tpb);
  const unsigned gz = (unsigned)((P->nx * P->ny + tpb - 1) / tpb);
  for (int64_t n = 0; n < steps; ++n) {
    if (P->interrupt && P->interrupt()) {
      P->interrupted = true;
      break;
    }
    HeatArgs<RT> a = make_args<RT>(P, dt, lightsOn);
    if (P->explicitScheme) {
      hipLaunchKernelGGL(heat_explicit<RT>, dim3(gz), dim3(tpb), 0, P->stream, a);
    } else {
      hipLaunchKernelGGL(heat_sweep_x<RT>, dim3(gx), dim3(tpb), 0, P->stream, a);
      hipLaunchKernelGGL(heat_sweep_y<RT>, dim3(gy), dim3(tpb), 0, P->stream, a);
      hipLaunchKernelGGL(heat_sweep_z<RT>, dim3(gz), dim3(tpb), 0, P->stream, a);
    }
    MCM_HIP(hipGetLastError());
    std::swap(P->T1.p, P->T2.p);
    ++P->stepsDone;
    if ((rc = record_sensors<RT>(P, P->stepsDone))) return rc;
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
  rc = P->fp32 ? record_sensors<float>(P.get(), 0) : record_sensors<double>(P.get(), 0);
  if (rc) return rc;
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
