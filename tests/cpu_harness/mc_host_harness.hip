// Host build of the GPU kernel's per-packet engine (mcmatlab_amd/csrc/
// mc_engine.h + mc_photon.h), run packet by packet with plain adds.
// TEST INFRASTRUCTURE ONLY: it lets CPU tests check the FP32 physics and the
// tally/normalisation logic of the GPU path against the oracle without a GPU.
// Differences to the GPU run: no wavefront scheduling, host libm instead of
// the device math library, no FMA contraction differences controlled.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mc_engine.h"
#include "mc_setup.h"

using namespace mcm;

namespace {

struct HostSink {
  static constexpr bool kDeferEscapes = false;
  const McParams* p;
  uint64_t collectedN = 0, runawayN = 0;
  void deposit(int j, double v) { p->NFR[j] += v; }
  void flush() {}
  void image_add(int b, double w) { p->image[b] += w; }
  void farField(int b, double w) { p->FF[b] += w; }
  void edge(int f, int i, double w) { p->NI[f][i] += w; }
  void collected() { ++collectedN; }
  void runaway() { ++runawayN; }
};

// 1: launch through LaunchRecord and tally escapes through EscapeRecord, as
// the GPU kernels do when no deposition criteria are active
int g_mode = 0;

struct HostDeferSink : HostSink {
  static constexpr bool kDeferEscapes = true;
  std::vector<EscapeRecord>* rec;
  void escape(const Photon& P, bool escaped, bool edgeFlux, uint32_t) {
    EscapeRecord E;
    pack_escape(P, escaped, edgeFlux, 0, E);
    rec->push_back(E);
  }
};

template <bool FR, bool CR>
void run_packets_records(const McParams& p, uint64_t n, uint32_t k0, uint32_t k1, uint32_t stream, uint64_t* launched,
                         uint64_t* collected, double* stats) {
  PacketEngine<FR, CR> E;
  E.init(k0, k1, stream);
  std::vector<EscapeRecord> rec;
  HostDeferSink sink;
  sink.p = &p;
  sink.rec = &rec;
  for (uint64_t k = 0; k < n; ++k) {
    PacketRng R;
    R.key0 = k0;
    R.key1 = k1;
    R.stream = stream;
    R.start(k);
    Photon P0;
    const bool ok = launch(P0, p, p.media, R);
    LaunchRecord L;
    pack_launch(P0, R, ok, L);
    E.replay = false;
    if (!E.start_from(k, L, p, p.media, p.cdfIdx)) break;
    ++*launched;
    while (E.state != ST_FETCH) {
      if (E.state == ST_MOVE) {
        stats[0] += 1;
        E.do_move(p, p.media, p.cdfIdx, sink);
      } else if (E.state == ST_SCATTER) {
        stats[1] += 1;
        E.do_scatter(p, p.media, p.cdfIdx, sink);
      } else {
        stats[2] += 1;
        E.do_fresnel(p, p.media, p.cdfIdx, sink);
      }
    }
  }
  HostSink tally{&p};
  for (const EscapeRecord& r : rec) tally_escape(r, p, tally);
  *collected += tally.collectedN;
}

template <bool FR, bool CR>
void run_packets(const McParams& p, uint64_t n, uint32_t k0, uint32_t k1, uint32_t stream, uint64_t* launched, uint64_t* collected,
                 double* stats) {
  if (!CR && g_mode == 1) {
    run_packets_records<FR, CR>(p, n, k0, k1, stream, launched, collected, stats);
    return;
  }
  PacketEngine<FR, CR> E;
  E.init(k0, k1, stream);
  HostSink sink{&p};
  for (uint64_t k = 0; k < n; ++k) {
    E.replay = false;
    if (!E.start(k, p, p.media, p.cdfIdx)) break;
    ++*launched;
    while (E.state != ST_FETCH) {
      if (E.state == ST_MOVE) {
        stats[0] += 1;
        E.do_move(p, p.media, p.cdfIdx, sink);
      } else if (E.state == ST_SCATTER) {
        stats[1] += 1;
        E.do_scatter(p, p.media, p.cdfIdx, sink);
      } else {
        stats[2] += 1;
        E.do_fresnel(p, p.media, p.cdfIdx, sink);
      }
    }
  }
  *collected += sink.collectedN;
}

}  // namespace

extern "C" void mc_host_set_mode(int mode) { g_mode = mode; }

extern "C" int mc_host_run(const mcm_mc_input* in, mcm_mc_output* out, uint64_t seed, double* stats /* [3] */) {
  McTables T;
  const std::string err = build_mc_tables(in, T);
  if (!err.empty()) return MCM_ERR_INVALID_ARGUMENT;
  McParams p = T.base;
  std::vector<double> nfr((size_t)T.L), image((size_t)std::max<int64_t>(T.nImage, 1)), ff((size_t)std::max<int64_t>(T.nFF, 1));
  std::vector<double> ni[6];
  for (int f = 0; f < 6; ++f) ni[f].assign((size_t)std::max<int64_t>(T.nNI[f], 1), 0.0);
  p.M = T.M0.data();
  p.normalRI = T.fresnelPossible ? T.normals.data() : nullptr;
  p.cdfs = T.cdfs.data();
  p.FPIDdist1 = T.fpid1.data();
  p.AIDdist1 = T.aid1.data();
  p.FPIDdist2 = T.fpid2.data();
  p.AIDdist2 = T.aid2.data();
  p.NFR = nfr.data();
  p.image = image.data();
  p.FF = ff.data();
  for (int f = 0; f < 6; ++f) p.NI[f] = ni[f].data();
  std::vector<double> S;
  double tmp[3] = {0, 0, 0};
  if (!stats) stats = tmp;
  const unsigned long long nReq = (unsigned long long)T.nPhotonsRequested;
  double cum = 0, cumColl = 0;
  for (int iL = 0; iL < T.nL; ++iL) {
    p.media = T.media.data() + (size_t)iL * 256;
    p.cdfIdx = T.cdfIdx.data() + (size_t)iL * 256;
    if (T.fresnelPossible)
      for (int64_t j = 0; j < T.L; ++j) T.normals[j].w = p.media[T.M0[j]].ri;
    double power;
    if (in->sourceDistribution) {
      S.assign((size_t)T.L + 1, 0.0);
      for (int64_t k = 1; k <= T.L; ++k) S[k] = S[k - 1] + (double)in->sourceDistribution[(size_t)iL * T.L + k - 1];
      const double tot = S[T.L];
      power = tot * T.V;
      for (int64_t k = 1; k <= T.L; ++k) S[k] /= tot;
      S[T.L] = 1.0;
      p.S = S.data();
    } else {
      power = T.spectrum[iL];
    }
    const unsigned long long reqL = iL == T.nL - 1 ? nReq - (unsigned long long)cum : nReq / (unsigned long long)T.nL;
    uint64_t launched = 0, collected = 0;
    const bool fr = T.fresnelPossible, cr = p.criteriaActive != 0;
    if (fr && cr) run_packets<true, true>(p, reqL, (uint32_t)seed, (uint32_t)(seed >> 32), iL, &launched, &collected, stats);
    else if (fr) run_packets<true, false>(p, reqL, (uint32_t)seed, (uint32_t)(seed >> 32), iL, &launched, &collected, stats);
    else if (cr) run_packets<false, true>(p, reqL, (uint32_t)seed, (uint32_t)(seed >> 32), iL, &launched, &collected, stats);
    else run_packets<false, false>(p, reqL, (uint32_t)seed, (uint32_t)(seed >> 32), iL, &launched, &collected, stats);
    cum += (double)launched;
    cumColl += (double)collected;
    double nf = (double)launched / power;
    if (in->sourceDistribution) nf /= power;
    else if (p.beamType == 2 && (p.boundaryType == 0 || p.boundaryType == 2)) nf /= kKillRange * kKillRange;
    if (T.calcNFR && out->NFR)
      for (int64_t j = 0; j < T.L; ++j) {
        out->NFR[(size_t)iL * T.L + j] = (float)(nfr[j] / (T.V * nf * T.muaNorm[(size_t)iL * T.nM + T.M0[j]]));
        nfr[j] = 0;
      }
    if (p.useLC && out->image) {
      const int64_t LLC = (int64_t)T.lcRes0 * T.lcRes0;
      const double div = LLC > 1 ? (double)p.lcFSorNA * (double)p.lcFSorNA / LLC * nf : nf;
      for (int64_t k = 0; k < T.nImage; ++k) {
        out->image[(size_t)iL * T.nImage + k] = (float)(image[k] / div);
        image[k] = 0;
      }
    }
    if (T.nFF && out->farField)
      for (int64_t k = 0; k < T.nFF; ++k) {
        out->farField[(size_t)iL * T.nFF + k] = (float)(ff[k] / nf);
        ff[k] = 0;
      }
    float* dst[6] = {out->NI_xpos, out->NI_xneg, out->NI_ypos, out->NI_yneg, out->NI_zpos, out->NI_zneg};
    const double area[6] = {T.dA[0], T.dA[0], T.dA[1], T.dA[1], T.dA[2], T.dA[2]};
    for (int f = 0; f < 6; ++f)
      if (T.nNI[f] && dst[f])
        for (int64_t k = 0; k < T.nNI[f]; ++k) {
          dst[f][(size_t)iL * T.nNI[f] + k] = (float)(ni[f][k] / (area[f] * nf));
          ni[f][k] = 0;
        }
  }
  out->nPhotons = cum;
  out->nPhotonsCollected = cumColl;
  out->nThreads = 1;
  out->simulationTime = 0;
  out->examplePaths = nullptr;
  out->examplePathsElems = 0;
  return MCM_OK;
}

// The FP32 Lambert W lower branch of the device code (LG01 source sampling).
extern "C" float mc_host_lambert_wm1(float x) { return lambert_wm1(x); }

// Philox words of the device header (for the RNG known-answer test).
extern "C" void mcm_test_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox4x32_10(c, key[0], key[1]);
  for (int k = 0; k < 4; ++k) out[k] = c[k];
}
