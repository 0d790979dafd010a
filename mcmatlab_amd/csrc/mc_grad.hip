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
// Each lane runs its packets twice with the same Philox stream: pass 1 finds
// the packet's box deposit D, pass 2 replays the path and adds the dJ/dmua
// terms with FP64 atomics (only packets with D > 0 are replayed).  The file
// is built without FMA contraction so that both passes, compiled as separate
// instantiations, evaluate identical operation sequences and retrace the
// same path.
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
};

// One packet; pass 1 returns its box deposit and score, pass 2 (given the
// deposit D) adds the packet's dJ/dmua terms.
template <bool FRESNEL, bool PASS2>
__device__ inline bool grad_packet(const McParams& p, const GradArgs& g, const Medium* med, const uint8_t* cdf, PacketRng& R, uint64_t k,
                                   double& D, double& score) {
  R.start(k);
  Photon P;
  if (!launch(P, p, med, R)) return false;
  new_voxel_properties(P, p, med, cdf);
  if (g.muaMap) P.mua = g.muaMap[P.j];
  double tau = 0.0;
  uint32_t K = 0;
  double rest = D;
  double acc = 0.0;
  int accJ = -1;
  while (P.alive) {
    while (P.alive && P.stepLeft > 0.f) {
      if (P.substeps >= kMaxSubsteps) {
        P.alive = false;
        break;
      }
      const int j = P.j;
      const bool inside = P.insideVolume;
      const float mus = P.mus;
      const int ix = clamp_axis(P.i[0], p.n[0]), iy = clamp_axis(P.i[1], p.n[1]), iz = clamp_axis(P.i[2], p.n[2]);
      const bool boxed = inside && ix >= g.box[0] && ix < g.box[1] && iy >= g.box[2] && iy < g.box[3] && iz >= g.box[4] && iz < g.box[5];
      const MoveResult mr = move<FRESNEL>(P, p, med);
      if (FRESNEL && mr.fresnel) fresnel(P, p, med, mr.jNew, mr.mNew, R);
      tau += (double)mr.s * (double)mus;
      if (!PASS2) {
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
          new_voxel_properties(P, p, med, cdf);
          if (g.muaMap) P.mua = g.muaMap[P.j];
        }
      }
    }
    if (!P.alive) break;
    ++K;  // the flight ended in an interaction
    roulette(P, R);
    if (P.alive) scatter(P, p, p.cdfs, 0, R);
  }
  if (PASS2 && accJ >= 0) unsafeAtomicAdd(&g.grad[accJ], acc);
  score = (double)K - tau;
  return true;
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
  PacketRng R;
  R.key0 = g.key0;
  R.key1 = g.key1;
  R.stream = g.stream;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < g.nPackets; k += stride) {
    double D = 0.0, S = 0.0;
    if (!grad_packet<FRESNEL, false>(p, g, sMed, sCdf, R, k, D, S)) {
      atomicOr(g.aborted, 1u);
      break;
    }
    s0 += D;
    s1 += D * D;
    s2 += D * S;
    s3 += (D * S) * (D * S);
    if (D > 0.0) grad_packet<FRESNEL, true>(p, g, sMed, sCdf, R, k, D, S);
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
  DevBuf dMap, dGrad, dSums, dAbort;
  if (in->muaMap && (rc = upload(dMap, in->muaMap, (size_t)T.L, s))) return rc;
  if ((rc = zeros(dGrad, (size_t)T.L * sizeof(double), s)) || (rc = zeros(dSums, 4 * sizeof(double), s)) ||
      (rc = zeros(dAbort, sizeof(unsigned), s)))
    return rc;
  GradArgs g;
  std::memset(&g, 0, sizeof g);
  g.muaMap = in->muaMap ? dMap.as<float>() : nullptr;
  for (int a = 0; a < 6; ++a) g.box[a] = (int)in->box[a];
  g.nPackets = in->nPhotons;
  g.key0 = (uint32_t)in->seed;
  g.key1 = (uint32_t)(in->seed >> 32);
  g.stream = (uint32_t)in->iL;
  g.grad = dGrad.as<double>();
  g.sums = dSums.as<double>();
  g.aborted = dAbort.as<unsigned>();
  int per = 0;
  const void* kern = T.fresnelPossible ? (const void*)mc_grad<true> : (const void*)mc_grad<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0) != hipSuccess || per < 1) per = 2;
  const uint64_t lanes = (uint64_t)per * cu_count(pl->device) * 256;
  // no more lanes than packets / 4, so the static packet assignment stays balanced
  const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(lanes, (in->nPhotons + 1023) / 4) / 256);
  if (T.fresnelPossible) hipLaunchKernelGGL(mc_grad<true>, dim3(blocks), dim3(256), 0, s, pl->prm, g);
  else hipLaunchKernelGGL(mc_grad<false>, dim3(blocks), dim3(256), 0, s, pl->prm, g);
  MCM_HIP(hipGetLastError());
  double sums[4];
  unsigned aborted = 0;
  MCM_HIP(hipMemcpyAsync(sums, dSums.p, sizeof sums, hipMemcpyDeviceToHost, s));
  MCM_HIP(hipMemcpyAsync(&aborted, dAbort.p, sizeof aborted, hipMemcpyDeviceToHost, s));
  MCM_HIP(hipStreamSynchronize(s));
  if (aborted) return fail(MCM_ERR_INVALID_ARGUMENT, "all packets launch outside the simulation cuboid");
  const double N = (double)in->nPhotons;
  const double power = T.spectrum[(size_t)in->iL];
  auto se = [&](double sx, double sx2) { return power * std::sqrt(std::max(0.0, sx2 - sx * sx / N) / (N * (N - 1))); };
  out->J = power * sums[0] / N;
  out->J_se = se(sums[0], sums[1]);
  out->dJ_dlogmus = power * sums[2] / N;
  out->dJ_dlogmus_se = se(sums[2], sums[3]);
  if (out->dJ_dmua) {
    std::vector<double> h((size_t)T.L);
    MCM_HIP(hipMemcpy(h.data(), dGrad.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < T.L; ++k) out->dJ_dmua[k] = power * h[k] / N;
  }
  return MCM_OK;
}
