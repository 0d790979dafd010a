// The device-resident Monte Carlo plan shared by the transport
// (mc_kernel.hip) and the gradient (mc_grad.hip) translation units: device
// buffers of the model, tallies, counters and launch parameters.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mc_photon.h"
#include "mc_setup.h"
#include "mcm_common.h"

namespace mcm {

struct McCounters {
  unsigned long long nextPacket;  // reset before every launch
  unsigned long long launched;
  unsigned long long collected;
  unsigned long long substeps;
  unsigned long long scatters;
  unsigned long long interfaces;
  unsigned int aborted;
  unsigned int runaway;
};

struct McLaunch {
  uint64_t firstPacket;
  uint64_t nPackets;
  uint32_t key0, key1, stream;
  float moveFrac;
  McCounters* ctr;
};

}  // namespace mcm

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const { return (T*)p; }
};

struct mcm_mc_plan {
  int device = 0;
  hipStream_t stream = nullptr;
  mcm::McTables T;
  mcm::McParams prm;
  int blocks = 0, blocksBatch = 0, blocksMulti = 0, threads = 256;
  float moveFrac = 0.5f;
  DevBuf M, normals, media, cdfIdx, cdfs, fpid1, aid1, fpid2, aid2, S, muaNorm;
  DevBuf NFR, image, FF, NI[6], counters, nfrOut;
  DevBuf launchRec, escapeRec;  // per-launch packet records (no deposition criteria)
  DevBuf gradMap, gradOut, gradSmall;  // gradient runs: mua map [L], dJ/dmua [L], sums/flags/counter
  uint64_t recCap = 0;
  std::vector<double> Shost;   // cumulative 3D source table of the current wavelength
  const float* sourceDistribution = nullptr;
  double power = 0;            // current wavelength's source power
  int curL = -1;
  uint64_t lastLaunched = 0, lastCollected = 0;
  mcm::McCounters stats{};
  ~mcm_mc_plan() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace mcm {

template <typename T>
inline int upload(DevBuf& b, const T* src, size_t n, hipStream_t s) {
  if (b.p) {
    MCM_HIP(hipFree(b.p));
    b.p = nullptr;
  }
  MCM_HIP(hipMalloc(&b.p, std::max<size_t>(n, 1) * sizeof(T)));
  if (src && n) MCM_HIP(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return MCM_OK;
}

inline int zeros(DevBuf& b, size_t bytes, hipStream_t s) {
  if (b.p) {
    MCM_HIP(hipFree(b.p));
    b.p = nullptr;
  }
  MCM_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 8)));
  MCM_HIP(hipMemsetAsync(b.p, 0, std::max<size_t>(bytes, 8), s));
  return MCM_OK;
}

// Select medium table, source table and refractive indices of wavelength iL
// (mc_kernel.hip).
int select_wavelength(mcm_mc_plan* pl, int64_t iL);

}  // namespace mcm
