// Media binning on the GPU: [downsampledParamVals, binidx] =
// getDownsampledParamVals(paramVals, N2) (+MCmatlab/src/getDownsampledParamVals.c:54-202).
//
// N1 parameter sets (mua, mus, g) of the voxels of one medium are reduced to
// N2 by farthest-point sampling in the range-normalised space of the rows
// that vary: the first pick is the set farthest from the mean, each further
// pick the set farthest from all picks so far, and every set is labelled
// with its nearest pick.  Each pass over the N1 sets is one kernel: a thread
// updates the running nearest distances and labels of its sets and keeps
// running maxima, and the workgroups publish theirs; a one-workgroup kernel
// reduces them and selects the next pick on the device, so the N2 passes
// run back to back on one stream with no host round trip.  (Selecting in
// the last workgroup to finish, by an atomic ticket, serialised thousands
// of same-address atomics per pass: 206 us per pass at 8e6 sets, against
// 27 us for the pass and the separate selection.)
//
// Selection.  The reference's single-threaded pass scans k = 0..N1-1 and
// keeps a running maximum whose index m starts at the previous pass's m
// (:164-166, kept in maxDistIdxs[0]); while k < m it compares against the
// value d[m] had before this pass (the distances only shrink), and after the
// pass the pick is m if d[m] > d[0], else 0 (:172-175).  With U the updated
// distances and S the old d[m0]:
//   * if some k < m0 has U[k] > S, the scan switches there and ends on the
//     first maximum of U over all k;
//   * else it ends on the first maximum of U over k >= m0.
// So a pass needs two first-maximum reductions (k < m0 and k >= m0), U[0]
// and S.  Distances are accumulated as in the reference: float += (double)x*x
// (exact square, one rounding per term).
//
// The running sums of the inputs (:73-83) are single-precision sums in input
// order, whose rounding depends on that order; they are taken on the host
// (one thread per row) in the pass that checks the input.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <type_traits>

#include "mcm_common.h"

namespace mcm {
namespace {

constexpr int kBinTPB = 256;
constexpr int kBinMaxBlocks = 8192;

struct BinArgs {
  const float* pv;       // [3 * n1] input, interleaved (mua, mus, g)
  float* norm;           // [rows][n1] normalised varying rows
  float* d;              // [n1] squared distance to the nearest pick
  uint8_t* bin;          // [n1] 1-based label of the nearest pick
  float4* part;          // [blocks] (lo value, lo index, hi value, hi index)
  int* picks;            // [n2] picked set of each bin
  int* st;               // [4]: pick, m0, unused, S (float bits)
  int64_t n1;
  int rows;
  int rowOf[3];          // input row -> normalised row (-1: constant)
  float range[3];
  float mean[3];
};

struct ArgMax {
  float v;
  int i;
};

// the larger value; equal values keep the lower index (first maximum)
__device__ inline ArgMax better(ArgMax a, ArgMax b) {
  if (b.v > a.v || (b.v == a.v && b.i >= 0 && (a.i < 0 || b.i < a.i))) return b;
  return a;
}

__device__ inline ArgMax wave_best(ArgMax a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ArgMax b;
    b.v = __shfl_xor(a.v, o, 64);
    b.i = __shfl_xor(a.i, o, 64);
    a = better(a, b);
  }
  return a;
}

// block-wide first maxima of two candidates (TPB threads); result valid in
// thread 0
template <int TPB = kBinTPB>
__device__ inline void block_best(ArgMax& lo, ArgMax& hi) {
  __shared__ ArgMax sLo[TPB / 64], sHi[TPB / 64];
  lo = wave_best(lo);
  hi = wave_best(hi);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sLo[w] = lo, sHi[w] = hi;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < TPB / 64; ++k) lo = better(lo, sLo[k]), hi = better(hi, sHi[k]);
}

constexpr int kSelTPB = 1024;

// Each workgroup publishes its two maxima; bin_select (one workgroup)
// reduces them and picks the next set.
__device__ inline void publish(const BinArgs& a, ArgMax lo, ArgMax hi) {
  block_best(lo, hi);
  if (threadIdx.x == 0) a.part[blockIdx.x] = make_float4(lo.v, __int_as_float(lo.i), hi.v, __int_as_float(hi.i));
}

// MODE 0: the first pick, farthest from the mean; 1: the second, farthest
// from the first; 2: later picks by the reference's scan rule.
template <int MODE>
__global__ void __launch_bounds__(kSelTPB) bin_select(BinArgs a, int nParts) {
  ArgMax L{-1.f, -1}, H{-1.f, -1};
  // up to 8 partials per thread, all loaded before they are compared
  for (int b0 = threadIdx.x; b0 < nParts; b0 += 8 * kSelTPB) {
    float4 q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int b = b0 + k * kSelTPB;
      q[k] = b < nParts ? a.part[b] : make_float4(-1.f, __int_as_float(-1), -1.f, __int_as_float(-1));
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      L = better(L, ArgMax{q[k].x, __float_as_int(q[k].y)});
      H = better(H, ArgMax{q[k].z, __float_as_int(q[k].w)});
    }
  }
  block_best<kSelTPB>(L, H);
  if (threadIdx.x != 0) return;
  if (MODE == 2) {
    const float S = __int_as_float(a.st[3]);
    int m = L.i >= 0 && L.v > S ? better(H, L).i : H.i;
    m = m < 0 ? 0 : m;
    const float um = m == L.i ? L.v : H.v;
    a.st[0] = um > a.d[0] ? m : 0;
    a.st[1] = m;
    a.st[3] = __float_as_int(um);
  } else {
    const int c = H.i < 0 ? 0 : H.i;
    a.st[0] = c;
    if (MODE == 0) a.picks[0] = c;
    a.st[1] = 0;  // the reference's running index starts at 0 (calloc)
    a.st[3] = 0;
  }
}

// pass 0: normalise the varying rows, distances to the mean, first pick
__global__ void __launch_bounds__(kBinTPB) bin_prepare(BinArgs a) {
  ArgMax best{-1.f, -1};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n1; i += stride) {
    float s = 0.f;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int q = a.rowOf[r];
      if (q < 0) continue;
      const float v = a.pv[3 * i + r] / a.range[r];
      a.norm[(int64_t)q * a.n1 + i] = v;
      const float x = v - a.mean[r];
      s = (float)((double)s + (double)x * (double)x);
    }
    if (s > best.v) best = ArgMax{s, (int)i};
  }
  publish(a, ArgMax{-1.f, -1}, best);
}

// Passes over the sets for ROWS varying rows: a thread takes kBinPer sets
// kBinTPB apart and issues all their loads before the arithmetic.
constexpr int kBinPer = 8;

template <int ROWS>
__device__ inline float sqr_dist_r(const float (&x)[ROWS > 0 ? ROWS : 1], const float (&p)[3]) {
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const float t = x[q] - p[q];
    s = (float)((double)s + (double)t * (double)t);
  }
  return s;
}

// pass 1: distances to the first pick, every label 1, second pick
template <int ROWS>
__global__ void __launch_bounds__(kBinTPB) bin_first(BinArgs a) {
  float p[3] = {0.f, 0.f, 0.f};
  const int c = a.picks[0];
#pragma unroll
  for (int q = 0; q < ROWS; ++q) p[q] = a.norm[(int64_t)q * a.n1 + c];
  ArgMax best{-1.f, -1};
  const int64_t stride = (int64_t)gridDim.x * kBinTPB * kBinPer;
  for (int64_t i0 = (int64_t)blockIdx.x * kBinTPB * kBinPer + threadIdx.x; i0 < a.n1; i0 += stride) {
    float x[kBinPer][ROWS > 0 ? ROWS : 1];
#pragma unroll
    for (int k = 0; k < kBinPer; ++k) {
      const int64_t i = i0 + k * kBinTPB;
#pragma unroll
      for (int q = 0; q < ROWS; ++q) x[k][q] = i < a.n1 ? a.norm[(int64_t)q * a.n1 + i] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < kBinPer; ++k) {
      const int64_t i = i0 + k * kBinTPB;
      if (i >= a.n1) break;
      const float s = sqr_dist_r<ROWS>(x[k], p);
      a.d[i] = s;
      a.bin[i] = 1;
      if (s > best.v) best = ArgMax{s, (int)i};
    }
  }
  publish(a, ArgMax{-1.f, -1}, best);
}

// pass j >= 1 of the reference loop: pick j is the current selection
template <int ROWS>
__global__ void __launch_bounds__(kBinTPB) bin_step(BinArgs a, int j) {
  const int c = a.st[0], m0 = a.st[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.picks[j] = c;
  float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < ROWS; ++q) p[q] = a.norm[(int64_t)q * a.n1 + c];
  const uint8_t label = (uint8_t)(j + 1);
  ArgMax lo{-1.f, -1}, hi{-1.f, -1};
  const int64_t stride = (int64_t)gridDim.x * kBinTPB * kBinPer;
  for (int64_t i0 = (int64_t)blockIdx.x * kBinTPB * kBinPer + threadIdx.x; i0 < a.n1; i0 += stride) {
    float x[kBinPer][ROWS > 0 ? ROWS : 1], u[kBinPer];
#pragma unroll
    for (int k = 0; k < kBinPer; ++k) {
      const int64_t i = i0 + k * kBinTPB;
      const bool ok = i < a.n1;
#pragma unroll
      for (int q = 0; q < ROWS; ++q) x[k][q] = ok ? a.norm[(int64_t)q * a.n1 + i] : 0.f;
      u[k] = ok ? a.d[i] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < kBinPer; ++k) {
      const int64_t i = i0 + k * kBinTPB;
      if (i >= a.n1) break;
      const float s = sqr_dist_r<ROWS>(x[k], p);
      float v = u[k];
      if (s < v) {
        v = s;
        a.d[i] = s;
        a.bin[i] = label;
      }
      if (i < m0) {
        if (v > lo.v) lo = ArgMax{v, (int)i};
      } else if (v > hi.v) {
        hi = ArgMax{v, (int)i};
      }
    }
  }
  publish(a, lo, hi);
}

// rescale the picks (:182-186)
__global__ void bin_output(BinArgs a, int64_t n2, float* out, float pv0, float pv1, float pv2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const int c = a.picks[i];
  const float first[3] = {pv0, pv1, pv2};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int q = a.rowOf[r];
    out[3 * i + r] = q >= 0 ? a.norm[(int64_t)q * a.n1 + c] * a.range[r] : first[r];
  }
}

struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

__global__ void bin_clear_state(int* st) {
  if (threadIdx.x < 4) st[threadIdx.x] = 0;
}

// Device buffers of one call (see BinArgs) and the call's stream.
struct BinBuffers {
  float *pv, *norm, *d, *out;
  uint8_t* bin;
  float4* part;
  int *picks, *st;
};

int blocks_for(int64_t n1) { return (int)std::min<int64_t>(kBinMaxBlocks, (n1 + kBinTPB * kBinPer - 1) / (kBinTPB * kBinPer)); }
int prep_blocks_for(int64_t n1) { return (int)std::min<int64_t>(kBinMaxBlocks, (n1 + kBinTPB - 1) / kBinTPB); }

int check_args(const float* paramVals, int64_t n1, int64_t n2, float* downsampled, uint8_t* binidx) {
  if (!paramVals || !downsampled || !binidx) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  if (n1 < 1) return fail(MCM_ERR_INVALID_ARGUMENT, "getDownsampledParamVals needs at least one parameter set");
  if (n2 < 1) return fail(MCM_ERR_INVALID_ARGUMENT, "getDownsampledParamVals needs at least one bin");
  if (n1 > INT32_MAX || n2 > INT32_MAX) return fail(MCM_ERR_INVALID_ARGUMENT, "more than 2^31 - 1 parameter sets or bins");
  return MCM_OK;
}

// ranges and running sums of input row r in input order (:64-83)
void row_pass(const float* paramVals, int64_t n1, int r, float* mx, float* mn, float* sum, bool* finite) {
  float hi = paramVals[r], lo = hi, sm = hi;
  bool ok = std::isfinite(hi);
  for (int64_t i = 1; i < n1; ++i) {
    const float v = paramVals[3 * i + r];
    ok = ok && std::isfinite(v);
    hi = v > hi ? v : hi;
    lo = v < lo ? v : lo;
    sm += v;
  }
  mx[r] = hi, mn[r] = lo, sum[r] = sm, finite[r] = ok;
}

// Everything after the upload of paramVals into B.pv and the row passes:
// the N2 passes, the output and the downloads, on stream s (synchronised
// before returning).
int run_passes(const float* paramVals, int64_t n1, int64_t n2, float* downsampled, uint8_t* binidx, const BinBuffers& B,
               const float* mx, const float* mn, const float* sum, hipStream_t s) {
  const int blocks = blocks_for(n1), blocksPrep = prep_blocks_for(n1);
  BinArgs a{};
  a.n1 = n1;
  a.rows = 0;
  for (int r = 0; r < 3; ++r) {
    a.rowOf[r] = mx[r] != mn[r] ? a.rows++ : -1;
    a.range[r] = mx[r] - mn[r];
    a.mean[r] = a.rowOf[r] >= 0 ? sum[r] / a.range[r] / (float)n1 : 0.f;  // :101-103
  }
  a.pv = B.pv;
  a.norm = B.norm;
  a.d = B.d;
  a.bin = B.bin;
  a.part = B.part;
  a.picks = B.picks;
  a.st = B.st;
  hipLaunchKernelGGL(bin_clear_state, dim3(1), dim3(64), 0, s, B.st);
  hipLaunchKernelGGL(bin_prepare, dim3(blocksPrep), dim3(kBinTPB), 0, s, a);
  hipLaunchKernelGGL(bin_select<0>, dim3(1), dim3(kSelTPB), 0, s, a, blocksPrep);
  auto passes = [&](auto rowsTag) {
    constexpr int R = decltype(rowsTag)::value;
    hipLaunchKernelGGL(bin_first<R>, dim3(blocks), dim3(kBinTPB), 0, s, a);
    hipLaunchKernelGGL(bin_select<1>, dim3(1), dim3(kSelTPB), 0, s, a, blocks);
    for (int64_t j = 1; j < n2; ++j) {
      hipLaunchKernelGGL(bin_step<R>, dim3(blocks), dim3(kBinTPB), 0, s, a, (int)j);
      hipLaunchKernelGGL(bin_select<2>, dim3(1), dim3(kSelTPB), 0, s, a, blocks);
    }
  };
  switch (a.rows) {
    case 0: passes(std::integral_constant<int, 0>()); break;
    case 1: passes(std::integral_constant<int, 1>()); break;
    case 2: passes(std::integral_constant<int, 2>()); break;
    default: passes(std::integral_constant<int, 3>()); break;
  }
  hipLaunchKernelGGL(bin_output, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, a, n2, B.out, paramVals[0], paramVals[1],
                     paramVals[2]);
  MCM_HIP(hipGetLastError());
  MCM_HIP(hipMemcpyAsync(downsampled, B.out, (size_t)n2 * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
  MCM_HIP(hipMemcpyAsync(binidx, B.bin, (size_t)n1, hipMemcpyDeviceToHost, s));
  MCM_HIP(hipStreamSynchronize(s));
  return MCM_OK;
}

// Scratch of the allocation-free entry point: every device buffer of a call
// with up to kScratchSets parameter sets and kScratchBins bins, laid out by
// scratch_layout.  Each GPU holds its own copy for the life of the process.
constexpr int64_t kScratchSets = 1 << 20;
constexpr int64_t kScratchBins = 4096;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScratchBytes = align256(kScratchSets * 12) * 2 + align256(kScratchSets * 4) + align256(kScratchSets) +
                                 align256(kBinMaxBlocks * 16) + align256(kScratchBins * 4) + align256(16) + align256(kScratchBins * 12);
__device__ __align__(256) unsigned char scratch[kScratchBytes];

BinBuffers scratch_layout(unsigned char* base, int64_t n1, int64_t n2) {
  BinBuffers B;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + o;
    o += align256(bytes);
    return p;
  };
  B.pv = (float*)take((size_t)n1 * 12);
  B.norm = (float*)take((size_t)n1 * 12);
  B.d = (float*)take((size_t)n1 * 4);
  B.bin = take((size_t)n1);
  B.part = (float4*)take((size_t)kBinMaxBlocks * 16);
  B.picks = (int*)take((size_t)n2 * 4);
  B.st = (int*)take(16);
  B.out = (float*)take((size_t)n2 * 12);
  return B;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratch[kMaxDevices] = {};

}  // namespace
}  // namespace mcm

using namespace mcm;

extern "C" int mcm_downsample_params(const float* paramVals, int64_t n1, int64_t n2, float* downsampled, uint8_t* binidx,
                                     int32_t device) {
  if (int rc = check_args(paramVals, n1, n2, downsampled, binidx)) return rc;
  if (int rc = use_device(device)) return rc;
  // the row passes run on host threads while the device buffers are set up
  // and the input is uploaded; the media functions' values are checked
  // finite by their callers, and a NaN here would leave the selection
  // undefined, so it is an error
  float mx[3], mn[3], sum[3];
  bool finite[3];
  struct Rows {  // joined on every return path
    std::thread t[3];
    ~Rows() {
      for (auto& x : t)
        if (x.joinable()) x.join();
    }
  } rowThreads;
  for (int r = 0; r < 3; ++r) rowThreads.t[r] = std::thread(row_pass, paramVals, n1, r, mx, mn, sum, finite);
  const int blocks = blocks_for(n1), blocksPrep = prep_blocks_for(n1);
  DevMem pv, norm, d, bin, part, picks, st, out;
  MCM_HIP(hipMalloc(&pv.p, (size_t)n1 * 3 * sizeof(float)));
  MCM_HIP(hipMalloc(&norm.p, (size_t)n1 * 3 * sizeof(float)));
  MCM_HIP(hipMalloc(&d.p, (size_t)n1 * sizeof(float)));
  MCM_HIP(hipMalloc(&bin.p, (size_t)n1));
  MCM_HIP(hipMalloc(&part.p, (size_t)std::max(blocks, blocksPrep) * sizeof(float4)));
  MCM_HIP(hipMalloc(&picks.p, (size_t)n2 * sizeof(int)));
  MCM_HIP(hipMalloc(&st.p, 4 * sizeof(int)));
  MCM_HIP(hipMalloc(&out.p, (size_t)n2 * 3 * sizeof(float)));
  hipStream_t s = nullptr;
  MCM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  MCM_HIP(hipMemcpyAsync(pv.p, paramVals, (size_t)n1 * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  for (auto& x : rowThreads.t) x.join();
  if (!(finite[0] && finite[1] && finite[2])) {
    (void)hipStreamSynchronize(s);
    return fail(MCM_ERR_INVALID_ARGUMENT, "getDownsampledParamVals: paramVals must be finite");
  }
  const BinBuffers B{(float*)pv.p, (float*)norm.p, (float*)d.p, (float*)out.p, (uint8_t*)bin.p, (float4*)part.p, (int*)picks.p, (int*)st.p};
  return run_passes(paramVals, n1, n2, downsampled, binidx, B, mx, mn, sum, s);
}

extern "C" int mcm_downsample_params_noalloc(const float* paramVals, int64_t n1, int64_t n2, float* downsampled, uint8_t* binidx,
                                             int32_t device, void* stream) {
  if (int rc = check_args(paramVals, n1, n2, downsampled, binidx)) return rc;
  if (n1 > kScratchSets || n2 > kScratchBins)
    return fail(MCM_ERR_INVALID_ARGUMENT, "mcm_downsample_params_noalloc: at most %lld parameter sets and %lld bins (got %lld, %lld)",
                (long long)kScratchSets, (long long)kScratchBins, (long long)n1, (long long)n2);
  if (device < 0 || device >= kMaxDevices) return fail(MCM_ERR_INVALID_ARGUMENT, "device %d out of range", device);
  if (int rc = use_device(device)) return rc;
  // the device's copy of the scratch array, looked up once per device with
  // that device current (the array's name is not a device address on the host)
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    if (!g_scratch[device]) {
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr)
        return fail(MCM_ERR_DEVICE, "hipGetSymbolAddress of the binning scratch: %s", hipGetErrorString(e));
      g_scratch[device] = ptr;
    }
    base = g_scratch[device];
  }
  hipStream_t s = (hipStream_t)stream;
  const BinBuffers B = scratch_layout((unsigned char*)base, n1, n2);
  MCM_HIP(hipMemcpyAsync(B.pv, paramVals, (size_t)n1 * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  float mx[3], mn[3], sum[3];
  bool finite[3];
  for (int r = 0; r < 3; ++r) row_pass(paramVals, n1, r, mx, mn, sum, finite);
  if (!(finite[0] && finite[1] && finite[2])) {
    (void)hipStreamSynchronize(s);
    return fail(MCM_ERR_INVALID_ARGUMENT, "getDownsampledParamVals: paramVals must be finite");
  }
  return run_passes(paramVals, n1, n2, downsampled, binidx, B, mx, mn, sum, s);
}
