// Differentiable absorbed power (no reference counterpart; the photon
// physics is the transport kernel's, mc_photon.h).
//
// J = absorbed power inside a box of voxels, per watt of source power, with
// the per-voxel absorption coefficients mua_v (an optional map overriding the
// media table) and the scattering coefficients of every medium scaled by s:
//   dJ/dmua_v   pathwise: mua only changes packet weights, never paths.  A
//               sub-step k of length s_k in voxel v_k (weight w_k before,
//               w_k+1 after, deposit d_k) contributes
//                 -s_k * (box deposits of the packet after sub-step k)
//                 + [v_k in box] * w_k+1 * s_k
//               to dJ/dmua_(v_k).  Russian roulette keeps the expectation for
//               every weight, so the estimator stays unbiased.
//   dJ/dlog s   likelihood ratio: every flight ends in an interaction with
//               density mus exp(-tau) or is cut off with probability
//               exp(-tau), so d log p(path)/d log s = K - tau_total (K
//               interactions, tau_total the optical depth traversed); the
//               estimator is J_packet * (K - tau_total).
// Each packet runs twice with the same Philox stream: pass 1 finds the
// packet's box deposit D, pass 2 replays the path and adds the dJ/dmua terms
// with FP64 atomics (only packets with D > 0 are replayed).  Both passes run
// the same instructions (the pass is a runtime flag), so pass 2 retraces
// pass 1's path exactly; the file is built without FMA contraction.
//
// Lanes are persistent and each packet is a small state machine like the
// transport's: FETCH (take the next packet number, one atomic per wavefront,
// and launch it), MOVE (one voxel sub-step), SCATTER (roulette +
// deflection).  Each iteration the wavefront runs one state for the lanes in
// it (MOVE while at least half of the live lanes want to move, otherwise the
// state with more waiting lanes), so a wavefront never waits for its
// longest-lived packet before starting others.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "mc_engine.h"
#include "mc_plan.h"

namespace mcm {

struct GradArgs {
  const float* muaMap;  // [L] or null
  int box[6];           // x0, x1, y0, y1, z0, z1
  uint64_t nPackets;
  uint32_t key0, key1, stream;
  double* grad;         // [L] sum over packets of dJ_packet/dmua
  double* sums;         // [4] sum D, sum D^2, sum D*S, sum (D*S)^2
  unsigned* aborted;
  unsigned long long* next;  // next packet number to hand out
  int replay;                // run pass 2 (dJ/dmua requested)
};

// Launch packet k of the stream (both passes start here).
MCM_FN bool grad_begin(Photon& P, const McParams& p, const GradArgs& g, const Medium* med, const uint8_t* cdf, PacketRng& R,
                       uint64_t k) {
  R.start(k);
  if (!launch(P, p, med, R)) return false;
  new_voxel_properties(P, p, med, cdf);
  if (g.muaMap) P.mua = g.muaMap[P.j];
  return true;
}

__device__ inline unsigned long long grad_bcast64(unsigned long long v, int src) {
  const unsigned lo = __shfl((unsigned)v, src), hi = __shfl((unsigned)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

template <bool FRESNEL>
__global__ void __launch_bounds__(256) mc_grad(McParams p, GradArgs g) {
  __shared__ Medium sMed[256];
  __shared__ uint8_t sCdf[256];
  __shared__ double sRed[4][4];
  for (int k = threadIdx.x; k < 256; k += blockDim.x) {
    sMed[k] = p.media[k];
    sCdf[k] = p.cdfIdx[k];
  }
  __syncthreads();
  constexpr int G_FETCH = 0, G_MOVE = 1, G_SCATTER = 2, G_DONE = 3;
  const int lane = __lane_id();
  PacketRng R;
  R.key0 = g.key0;
  R.key1 = g.key1;
  R.stream = g.stream;
  Photon P;
  int st = G_FETCH;
  bool pass2 = false;
  uint64_t k = 0;
  double D = 0.0, tau = 0.0, rest = 0.0, acc = 0.0;
  uint32_t K = 0;
  int accJ = -1;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (;;) {
    const unsigned long long bFetch = __ballot(st == G_FETCH);
    const unsigned long long bMove = __ballot(st == G_MOVE);
    const unsigned long long bScat = __ballot(st == G_SCATTER);
    const int cF = __popcll(bFetch), cM = __popcll(bMove), cS = __popcll(bScat);
    const int active = cF + cM + cS;
    if (active == 0) break;
    int phase = G_MOVE;
    if (2 * cM < active) phase = cS >= cF ? G_SCATTER : G_FETCH;
    bool ended = false;  // the lane's current pass ended this iteration
    if (phase == G_FETCH) {
      const int leader = __ffsll((long long)bFetch) - 1;
      unsigned long long base = 0;
      if (lane == leader) {
        base = atomicAdd(g.next, (unsigned long long)cF);
        if (atomicAdd(g.aborted, 0u)) base = g.nPackets;  // a failed launch stops the run
      }
      base = grad_bcast64(base, leader);
      if (st == G_FETCH) {
        k = base + (unsigned long long)__popcll(bFetch & ((1ull << lane) - 1ull));
        if (k >= g.nPackets) {
          st = G_DONE;
        } else if (!grad_begin(P, p, g, sMed, sCdf, R, k)) {
          atomicOr(g.aborted, 1u);
          st = G_DONE;
        } else {
          st = G_MOVE;
          pass2 = false;
          D = 0.0;
          tau = 0.0;
          K = 0;
        }
      }
    } else if (phase == G_MOVE) {
      if (st == G_MOVE) {
        if (P.alive && P.stepLeft > 0.f) {
          if (P.substeps >= kMaxSubsteps) {
            P.alive = false;
          } else {
            const int j = P.j;
            const bool inside = P.insideVolume;
            const float mus = P.mus;
            const int ix = clamp_axis(P.i[0], p.n[0]), iy = clamp_axis(P.i[1], p.n[1]), iz = clamp_axis(P.i[2], p.n[2]);
            const bool boxed = inside && ix >= g.box[0] && ix < g.box[1] && iy >= g.box[2] && iy < g.box[3] && iz >= g.box[4] &&
                               iz < g.box[5];
            const MoveResult mr = move<FRESNEL>(P, p, sMed);
            if (FRESNEL && mr.fresnel) fresnel(P, p, sMed, mr.jNew, mr.mNew, R);
            tau += (double)mr.s * (double)mus;
            if (!pass2) {
              if (boxed) D += (double)mr.absorb;
            } else {
              if (boxed) rest -= (double)mr.absorb;
              double gv = -(double)mr.s * rest;
              if (boxed) gv += (double)P.weight * (double)mr.s;
              if (j == accJ) {
                acc += gv;
              } else {
                if (accJ >= 0) unsafeAtomicAdd(&g.grad[accJ], acc);
                accJ = j;
                acc = gv;
              }
            }
            if (!P.sameVoxel) {
              check_escape(P, p);
              if (P.alive) {
                new_voxel_properties(P, p, sMed, sCdf);
                if (g.muaMap) P.mua = g.muaMap[P.j];
              }
            }
          }
        }
        if (!P.alive) ended = true;
        else if (!(P.stepLeft > 0.f)) st = G_SCATTER;
      }
    } else {
      if (st == G_SCATTER) {
        ++K;  // the flight ended in an interaction
        roulette(P, R);
        if (P.alive) scatter(P, p, p.cdfs, 0, R);
        if (P.alive) st = G_MOVE;
        else ended = true;
      }
    }
    if (ended) {
      if (!pass2) {
        const double S = (double)K - tau;
        s0 += D;
        s1 += D * D;
        s2 += D * S;
        s3 += (D * S) * (D * S);
        st = G_FETCH;
        if (D > 0.0 && g.replay && grad_begin(P, p, g, sMed, sCdf, R, k)) {
          pass2 = true;
          rest = D;
          tau = 0.0;
          K = 0;
          acc = 0.0;
          accJ = -1;
          st = G_MOVE;
        }
      } else {
        if (accJ >= 0) unsafeAtomicAdd(&g.grad[accJ], acc);
        accJ = -1;
        st = G_FETCH;
      }
    }
  }
  // wavefront then workgroup reduction of the four sums
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
    s2 += __shfl_down(s2, off, 64);
    s3 += __shfl_down(s3, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sRed[w][0] = s0;
    sRed[w][1] = s1;
    sRed[w][2] = s2;
    sRed[w][3] = s3;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sRed[q][threadIdx.x];
    unsafeAtomicAdd(&g.sums[threadIdx.x], t);
  }
}

}  // namespace mcm

using namespace mcm;

extern "C" int mcm_mc_grad_run(mcm_mc_plan* pl, const mcm_mc_grad_input* in, mcm_mc_grad_output* out) {
  if (!pl || !in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  McTables& T = pl->T;
  if (pl->prm.criteriaActive) return fail(MCM_ERR_UNSUPPORTED, "gradients run without deposition criteria");
  if (in->iL < 0 || in->iL >= T.nL) return fail(MCM_ERR_INVALID_ARGUMENT, "wavelength index out of range");
  if (in->nPhotons < 2) return fail(MCM_ERR_INVALID_ARGUMENT, "nPhotons must be at least 2");
  const int64_t n3[3] = {T.nx, T.ny, T.nz};
  for (int a = 0; a < 3; ++a)
    if (in->box[2 * a] < 0 || in->box[2 * a + 1] > n3[a] || in->box[2 * a] >= in->box[2 * a + 1])
      return fail(MCM_ERR_INVALID_ARGUMENT, "box must satisfy 0 <= lo < hi <= grid size on every axis");
  if (in->muaMap)
    for (int64_t k = 0; k < T.L; ++k)
      if (!(in->muaMap[k] >= 0.f)) return fail(MCM_ERR_INVALID_ARGUMENT, "muaMap must be non-negative");
  MCM_HIP(hipSetDevice(pl->device));
  int rc = select_wavelength(pl, in->iL);
  if (rc) return rc;
  hipStream_t s = pl->stream;
  // device buffers live in the plan across calls: [L] floats for the map,
  // [L] doubles for the gradient, and 64 bytes for the four sums, the abort
  // flag and the packet counter
  if (!pl->gradOut.p) {
    MCM_HIP(hipMalloc(&pl->gradMap.p, (size_t)T.L * sizeof(float)));
    MCM_HIP(hipMalloc(&pl->gradOut.p, (size_t)T.L * sizeof(double)));
    MCM_HIP(hipMalloc(&pl->gradSmall.p, 64));
  }
  if (in->muaMap) MCM_HIP(hipMemcpyAsync(pl->gradMap.p, in->muaMap, (size_t)T.L * sizeof(float), hipMemcpyHostToDevice, s));
  MCM_HIP(hipMemsetAsync(pl->gradOut.p, 0, (size_t)T.L * sizeof(double), s));
  MCM_HIP(hipMemsetAsync(pl->gradSmall.p, 0, 64, s));
  char* small = (char*)pl->gradSmall.p;
  GradArgs g;
  std::memset(&g, 0, sizeof g);
  g.muaMap = in->muaMap ? pl->gradMap.as<float>() : nullptr;
  for (int a = 0; a < 6; ++a) g.box[a] = (int)in->box[a];
  g.nPackets = in->nPhotons;
  g.key0 = (uint32_t)in->seed;
  g.key1 = (uint32_t)(in->seed >> 32);
  g.stream = (uint32_t)in->iL;
  g.grad = pl->gradOut.as<double>();
  g.sums = (double*)small;                          // bytes 0..31
  g.next = (unsigned long long*)(small + 32);       // bytes 32..39
  g.aborted = (unsigned*)(small + 40);              // bytes 40..43
  g.replay = out->dJ_dmua != nullptr;              // J and dJ/dlog s need pass 1 only
  int per = 0;
  const void* kern = T.fresnelPossible ? (const void*)mc_grad<true> : (const void*)mc_grad<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0) != hipSuccess || per < 1) per = 2;
  const uint64_t lanes = (uint64_t)per * cu_count(pl->device) * 256;
  // persistent lanes; no more lanes than packets
  const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(lanes, in->nPhotons + 255) / 256);
  if (T.fresnelPossible) hipLaunchKernelGGL(mc_grad<true>, dim3(blocks), dim3(256), 0, s, pl->prm, g);
  else hipLaunchKernelGGL(mc_grad<false>, dim3(blocks), dim3(256), 0, s, pl->prm, g);
  MCM_HIP(hipGetLastError());
  unsigned char hs[64];
  MCM_HIP(hipMemcpyAsync(hs, small, 64, hipMemcpyDeviceToHost, s));
  if (out->dJ_dmua) MCM_HIP(hipMemcpyAsync(out->dJ_dmua, pl->gradOut.p, (size_t)T.L * sizeof(double), hipMemcpyDeviceToHost, s));
  MCM_HIP(hipStreamSynchronize(s));
  double sums[4];
  unsigned aborted = 0;
  std::memcpy(sums, hs, sizeof sums);
  std::memcpy(&aborted, hs + 40, sizeof aborted);
  if (aborted) return fail(MCM_ERR_INVALID_ARGUMENT, "all packets launch outside the simulation cuboid");
  const double N = (double)in->nPhotons;
  const double power = T.spectrum[(size_t)in->iL];
  auto se = [&](double sx, double sx2) { return power * std::sqrt(std::max(0.0, sx2 - sx * sx / N) / (N * (N - 1))); };
  out->J = power * sums[0] / N;
  out->J_se = se(sums[0], sums[1]);
  out->dJ_dlogmus = power * sums[2] / N;
  out->dJ_dlogmus_se = se(sums[2], sums[3]);
  if (out->dJ_dmua)
    for (int64_t k = 0; k < T.L; ++k) out->dJ_dmua[k] = power * out->dJ_dmua[k] / N;
  return MCM_OK;
}
