// Monte Carlo photon transport on gfx950: kernels and the C ABI
// (mcm_mc_plan_*, mcm_mc_run).  Replaces the reference's photon loop
// (+MCmatlab/src/MCmatlab.c:178-315) and MEX body (:317-791).
//
// Kernel design
//   One packet per lane, persistent lanes: a lane whose packet dies fetches
//   the next packet number from a device counter (one atomic per wavefront
//   fetch) and keeps going until the launch's packets are used up.  Every
//   packet draws from its own Philox stream, so the set of simulated packets
//   and each packet's trajectory are independent of scheduling.
//
//   Each packet is a small state machine: FETCH (launch), MOVE (one voxel
//   sub-step with continuous absorption), SCATTER (roulette + deflection) and
//   FRESNEL (reflection/refraction at a refractive-index change).  Lanes of a
//   wavefront are usually in different states; instead of executing every
//   state's code under a divergent mask each iteration, the wavefront votes
//   (ballot) and runs one state per iteration: MOVE while at least `moveFrac`
//   of the live lanes want to move, otherwise whichever state has the most
//   lanes waiting.  Lanes waiting for a rarer state idle briefly and are then
//   served together, which keeps SIMD occupancy high for the expensive
//   scatter and interface code.
//
//   Per-medium optical constants live in LDS.  Deposits accumulate in a
//   per-lane FP64 register while the packet stays in one voxel and go to the
//   FP64 fluence grid with a hardware FP64 atomic add when it leaves (or
//   dies).  Boundary, far-field and detector tallies are FP64 atomics too.
//
//   Deposition criteria evaluated at end of life (MCmatlab.c:262-264) are
//   handled by replay instead of per-packet deposit records: when a packet
//   dies meeting the criteria, the lane re-launches the same packet number,
//   whose stream reproduces the trajectory exactly, and deposits the
//   absorbed weights scaled by the final weight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "mc_engine.h"
#include "mc_photon.h"
#include "mc_plan.h"
#include "mc_setup.h"
#include "mcm_common.h"

namespace mcm {

// Work items of a batched launch (nItems == 0: a plain launch).  Item i
// owns the global packet indices cum[i] .. cum[i+1]-1; packet g of item i is
// packet g - cum[i] of Philox stream (key[2i], key[2i+1], stream[i]), so an
// item reproduces a plain launch of the same wavelength, seed and stream.
// Media tables are per item ([nItems * nM]); tallies are per item slices.
// Kernels take the kind of launch as a template argument: kPlain, kBatch
// (items on the launch's grid) or kMulti (items on different grids: each
// item names a model, whose grid, tables and tally layout it uses).
constexpr int kBatchMaxItems = 256;
constexpr int kBatchMaxMedia = 2048;  // nItems * nM entries staged in LDS
constexpr int kPlain = 0, kBatch = 1, kMulti = 2;

// What the transport reads of a model that differs between the models of a
// multi-grid launch (copied into the lane's parameters at each fetch).
struct McGeom {
  int n[3];
  float d[3], inv_d[3];
  float tooOld;
  int boundaryType, calcNFR, needTime, useLC, farFieldRes;
  const uint8_t* M;
  const float4* normalRI;
  const float* cdfs;
};

struct McBatch {
  int nItems, nM;
  const unsigned long long* cum;  // [nItems + 1]
  const uint32_t* key;            // [2 * nItems]
  const uint32_t* stream;         // [nItems]
  const Medium* media;            // [nItems * nM]
  const uint8_t* cdfIdx;          // [nItems * nM]
  int64_t L, nImage, nFF, nNI[6];
  double* NFR;                    // [nItems * L]
  double* image;                  // [nItems * nImage]
  double* FF;                     // [nItems * nFF]
  double* NI[6];                  // [nItems * nNI[f]]
  unsigned long long* collected;  // [nItems]
  // multi-grid launches
  const int* model;               // [nItems] model of each item
  const McParams* params;         // [nModels] (launch and escape tallies)
  const McGeom* geom;             // [nModels]
  const int64_t* off;             // [nItems * 9] offsets of the item's NFR, image, FF, NI[0..5] slices
  unsigned long long* next;       // [nItems] next global packet index of each item in this launch
  const unsigned long long* end;  // [nItems] end of each item's packets in this launch
};

// Per-launch packet ranges of a multi-grid launch: item i's packets in
// [first, first + n) are [next[i], end[i]).
__global__ void mc_multi_ranges(McBatch bt, unsigned long long first, unsigned long long n, unsigned long long* next,
                                unsigned long long* end) {
  for (int i = threadIdx.x; i < bt.nItems; i += blockDim.x) {
    const unsigned long long lo = bt.cum[i] > first ? bt.cum[i] : first;
    const unsigned long long hi = bt.cum[i + 1] < first + n ? bt.cum[i + 1] : first + n;
    next[i] = lo;
    end[i] = hi > lo ? hi : lo;
  }
}

// The parameters a wavefront uses: the launch's, or in a multi-grid launch
// the wavefront's own copy, pointed at the model of its current item.
template <int KIND>
struct LaneParams {
  __device__ static const McParams& get(const McParams& p, McParams&) { return p; }
};
template <>
struct LaneParams<kMulti> {
  __device__ static const McParams& get(const McParams&, McParams& q) { return q; }
};

// Point a multi-grid wavefront at item `it` (wavefront-uniform, so all of
// this lives in scalar registers).
template <class Engine, class Sink>
__device__ inline __attribute__((always_inline)) void multi_item(const McBatch& bt, int it, McParams& q, const Medium*& med,
                                                                 const uint8_t*& cdf, const Medium* sMed, const uint8_t* sCdf,
                                                                 Engine& E, Sink& sink) {
  const McGeom& G = bt.geom[bt.model[it]];
  for (int a = 0; a < 3; ++a) {
    q.n[a] = G.n[a];
    q.d[a] = G.d[a];
    q.inv_d[a] = G.inv_d[a];
  }
  q.tooOld = G.tooOld;
  q.boundaryType = G.boundaryType;
  q.calcNFR = G.calcNFR;
  q.needTime = G.needTime;
  q.useLC = G.useLC;
  q.farFieldRes = G.farFieldRes;
  q.M = G.M;
  q.normalRI = G.normalRI;
  q.cdfs = G.cdfs;
  med = sMed + it * bt.nM;
  cdf = sCdf + it * bt.nM;
  E.R.key0 = bt.key[2 * it];
  E.R.key1 = bt.key[2 * it + 1];
  E.R.stream = bt.stream[it];
  sink.NFR = bt.NFR + bt.off[9 * (size_t)it];
  sink.item = it;
}

// item owning global packet index g: the last i with cum[i] <= g
__device__ inline int batch_item(unsigned long long g, const unsigned long long* cum, int n) {
  int lo = 0, hi = n;  // cum[lo] <= g < cum[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cum[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ inline unsigned long long bcast64(unsigned long long v, int src) {
  const unsigned lo = __shfl((unsigned)v, src), hi = __shfl((unsigned)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ inline void fadd(double* p, double v) { unsafeAtomicAdd(p, v); }

// lanes per transport workgroup: the stride of a lane's deposit ring slots
constexpr int kRingStride = 256;

// Largest number of packets per transport launch (sizes the launch and
// escape record buffers: 64 + 48 bytes per packet).
constexpr uint64_t kMaxLaunchPackets = 1ull << 24;

// Tally sink of the GPU kernel: FP64 hardware atomics; consecutive deposits
// into the same voxel are merged in a register first.  Holds copies of the
// tally pointers (not a pointer to the launch parameters, whose address must
// not escape or the parameters are copied to scratch).
struct DeviceSink {
  static constexpr bool kDeferEscapes = false;
  double* NFR;
  double* image;
  double* FF;
  double *ni0, *ni1, *ni2, *ni3, *ni4, *ni5;
  double acc;
  int accJ;
  unsigned long long nCollected;
  unsigned nRunaway;
  // the accelerator complex die (XCD) the wavefront runs on; each XCD has
  // its own L2, so XCD-private fluence copies keep hot voxels from being
  // contended across XCDs
  __device__ static int xcd() {
    unsigned r;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(r));
    return (int)(r & 7u);
  }
  __device__ explicit DeviceSink(const McParams& p)
      : NFR(p.NFR + (size_t)(xcd() % p.nfrReplicas) * p.L), image(p.image), FF(p.FF), ni0(p.NI[0]), ni1(p.NI[1]), ni2(p.NI[2]), ni3(p.NI[3]), ni4(p.NI[4]),
        ni5(p.NI[5]), acc(0.0), accJ(-1), nCollected(0), nRunaway(0) {}
  __device__ void deposit(int j, double v) {
    if (j == accJ) {
      acc += v;
    } else {
      if (accJ >= 0) fadd(&NFR[accJ], acc);
      accJ = j;
      acc = v;
    }
  }
  __device__ void flush() {
    if (accJ >= 0) fadd(&NFR[accJ], acc);
    accJ = -1;
    acc = 0.0;
  }
  __device__ void image_add(int b, double w) { fadd(&image[b], w); }
  __device__ void farField(int b, double w) { fadd(&FF[b], w); }
  __device__ void edge(int f, int i, double w) {
    double* base = f == 0 ? ni0 : f == 1 ? ni1 : f == 2 ? ni2 : f == 3 ? ni3 : f == 4 ? ni4 : ni5;
    fadd(&base[i], w);
  }
  __device__ void collected() { ++nCollected; }
  __device__ void runaway() { ++nRunaway; }
};

// Deposits waiting in LDS per lane (see DeferSink), drained by the whole
// wavefront together.
#ifndef MCM_DEPOSIT_RING
#define MCM_DEPOSIT_RING 8
#endif
constexpr int kRing = MCM_DEPOSIT_RING;

// Sink of the transport kernel without deposition criteria: a packet dies
// once, so its escape is written to the record slot of its own launch index
// (no shared counter) and tallied by mc_tally_escapes after the launch.
//
// Merged deposits do not go to the fluence one at a time: each lane parks
// (address, value) in its slots of an LDS ring and the wavefront drains the
// rings together when one of them is nearly full.  Global loads and atomics
// share one in-order completion counter, so a load issued after an atomic
// (the next voxel's medium index) waits for the atomic's round trip to the
// memory side; draining in batches pays that wait once per batch instead of
// once per voxel crossing.
struct DeferSink : DeviceSink {
  static constexpr bool kDeferEscapes = true;
  EscapeRecord* rec;
  int item;
  unsigned long long* ringA;  // LDS: slot k of this lane at [k * stride]
  double* ringV;
  int nRing;
  __device__ DeferSink(const McParams& p, EscapeRecord* r) : DeviceSink(p), rec(r), item(0), ringA(nullptr), ringV(nullptr), nRing(0) {}
  __device__ void park(double* a, double v) {
    if (!ringA) {
      fadd(a, v);
      return;
    }
    ringA[nRing * kRingStride] = (unsigned long long)a;
    ringV[nRing * kRingStride] = v;
    if (++nRing == kRing) drain();  // a full ring drains this lane alone
  }
  __device__ void drain() {
    for (int k = 0; k < nRing; ++k) fadd((double*)ringA[k * kRingStride], ringV[k * kRingStride]);
    nRing = 0;
  }
  __device__ void deposit(int j, double v) {
    if (j == accJ) {
      acc += v;
    } else {
      if (accJ >= 0) park(&NFR[accJ], acc);
      accJ = j;
      acc = v;
    }
  }
  __device__ void flush() {
    if (accJ >= 0) park(&NFR[accJ], acc);
    accJ = -1;
    acc = 0.0;
  }
  __device__ void escape(const Photon& P, bool escaped, bool edgeFlux, uint32_t slot) {
    EscapeRecord E;
    pack_escape(P, escaped, edgeFlux, item, E);
    rec[slot] = E;
  }
};

template <class Sink>
__device__ inline Sink make_sink(const McParams& p, EscapeRecord* r);
template <>
__device__ inline DeviceSink make_sink<DeviceSink>(const McParams& p, EscapeRecord*) {
  return DeviceSink(p);
}
template <>
__device__ inline DeferSink make_sink<DeferSink>(const McParams& p, EscapeRecord* r) {
  return DeferSink(p, r);
}

// launchPhoton for packets firstPacket .. firstPacket + nPackets - 1 into
// rec[0 .. nPackets-1]; clears the packets' escape record slots.  A launch
// that fails 10^6 times flags the run as aborted and the remaining packets
// are skipped (the transport kernel does not start any packet of an aborted
// run).
template <int BATCH>
__global__ void __launch_bounds__(256) mc_launch_packets(McParams p, McLaunch lp, McBatch bt, LaunchRecord* __restrict__ rec,
                                                         EscapeRecord* __restrict__ esc) {
  PacketRng R;
  R.key0 = lp.key0;
  R.key1 = lp.key1;
  R.stream = lp.stream;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < lp.nPackets; k += stride) {
    if (__hip_atomic_load(&lp.ctr->aborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    const Medium* media = p.media;
    int it = 0;
    if constexpr (BATCH != kPlain) {
      const unsigned long long g = lp.firstPacket + k;
      it = batch_item(g, bt.cum, bt.nItems);
      R.key0 = bt.key[2 * it];
      R.key1 = bt.key[2 * it + 1];
      R.stream = bt.stream[it];
      R.start(g - bt.cum[it]);
      media = bt.media + (size_t)it * bt.nM;
    } else {
      R.start(lp.firstPacket + k);
    }
    Photon P;
    bool ok;
    if constexpr (BATCH == kMulti) ok = launch(P, bt.params[bt.model[it]], media, R);
    else ok = launch(P, p, media, R);
    LaunchRecord L;
    pack_launch(P, R, ok, L);
    rec[k] = L;
    esc[k].c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!ok) atomicOr(&lp.ctr->aborted, 1u);
  }
}

// Light collector, far field and edge-flux tallies of the recorded escapes
// (slots whose flags are zero belong to packets absorbed inside).
template <int BATCH>
__global__ void __launch_bounds__(256) mc_tally_escapes(McParams p, McLaunch lp, McBatch bt, const EscapeRecord* __restrict__ rec) {
  DeviceSink sink(p);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < lp.nPackets; k += stride) {
    const float4 c = rec[k].c;
    if (c.y == 0.f) continue;
    EscapeRecord E;
    E.a = rec[k].a;
    E.b = rec[k].b;
    E.c = c;
    if constexpr (BATCH == kMulti) {
      const int it = (int)c.z;
      const int64_t* o = bt.off + 9 * (size_t)it;
      DeviceSink s(p);
      s.image = bt.image + o[1];
      s.FF = bt.FF + o[2];
      s.ni0 = bt.NI[0] + o[3];
      s.ni1 = bt.NI[1] + o[4];
      s.ni2 = bt.NI[2] + o[5];
      s.ni3 = bt.NI[3] + o[6];
      s.ni4 = bt.NI[4] + o[7];
      s.ni5 = bt.NI[5] + o[8];
      tally_escape(E, bt.params[bt.model[it]], s);
      if (s.nCollected) atomicAdd(&bt.collected[it], s.nCollected);
    } else if constexpr (BATCH == kBatch) {
      const int it = (int)c.z;
      DeviceSink s(p);
      s.image = bt.image + (size_t)it * bt.nImage;
      s.FF = bt.FF + (size_t)it * bt.nFF;
      s.ni0 = bt.NI[0] + (size_t)it * bt.nNI[0];
      s.ni1 = bt.NI[1] + (size_t)it * bt.nNI[1];
      s.ni2 = bt.NI[2] + (size_t)it * bt.nNI[2];
      s.ni3 = bt.NI[3] + (size_t)it * bt.nNI[3];
      s.ni4 = bt.NI[4] + (size_t)it * bt.nNI[4];
      s.ni5 = bt.NI[5] + (size_t)it * bt.nNI[5];
      tally_escape(E, p, s);
      if (s.nCollected) atomicAdd(&bt.collected[it], s.nCollected);
    } else {
      tally_escape(E, p, sink);
    }
  }
  if (sink.nCollected) atomicAdd(&lp.ctr->collected, sink.nCollected);
}

template <bool FRESNEL, bool CRITERIA, int BATCH>
__global__ void __launch_bounds__(256, BATCH == kMulti ? 4 : 1) mc_transport(McParams p, McLaunch lp, McBatch bt, const LaunchRecord* __restrict__ launches,
                                                    EscapeRecord* __restrict__ escapes) {
  static_assert(!(BATCH != kPlain && CRITERIA), "batched launches run without deposition criteria");
  constexpr int kMed = BATCH != kPlain ? kBatchMaxMedia : 256;
  __shared__ Medium sMed[kMed];
  __shared__ uint8_t sCdf[kMed];
  __shared__ unsigned long long sCum[BATCH != kPlain ? kBatchMaxItems + 1 : 1];
  if constexpr (BATCH != kPlain) {
    for (int k = threadIdx.x; k < bt.nItems * bt.nM; k += blockDim.x) {
      sMed[k] = bt.media[k];
      sCdf[k] = bt.cdfIdx[k];
    }
    for (int k = threadIdx.x; k <= bt.nItems; k += blockDim.x) sCum[k] = bt.cum[k];
  } else {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) {
      sMed[k] = p.media[k];
      sCdf[k] = p.cdfIdx[k];
    }
  }
  __syncthreads();

  const int lane = __lane_id();
  PacketEngine<FRESNEL, CRITERIA> E;
  E.init(lp.key0, lp.key1, lp.stream);
  using Sink = typename std::conditional<CRITERIA, DeviceSink, DeferSink>::type;
  Sink sink = make_sink<Sink>(p, escapes);
  // the deposit rings of the plain launch (batched and multi-grid launches
  // keep every item's tables in LDS and deposit directly)
  constexpr int RING = (!CRITERIA && BATCH == kPlain) ? kRing : 0;
  __shared__ unsigned long long sRingA[RING > 0 ? RING * kRingStride : 1];
  __shared__ double sRingV[RING > 0 ? RING * kRingStride : 1];
  if constexpr (RING > 0) {
    sink.ringA = sRingA + threadIdx.x;
    sink.ringV = sRingV + threadIdx.x;
  }
  // the media tables of the lane's work item (the launch's in a plain launch)
  const Medium* med = sMed;
  const uint8_t* cdf = sCdf;
  unsigned long long nLaunched = 0, nSub = 0, nScat = 0, nIf = 0;
  McParams q;
  const McParams& pr = LaneParams<BATCH>::get(p, q);
  // Multi-grid launches: every lane of a wavefront holds packets of one
  // item (the wavefront's current item), so the item's grid and tables are
  // wavefront-uniform.  Wavefronts start spread over the launch's packets
  // (so items share the GPU in proportion to their packets and their tails
  // overlap); once their item has no packets left they drain and move on to
  // the next item that has.
  int cur = 0;
  bool curDone = false;
  if constexpr (BATCH == kMulti) {
    q = p;
    const unsigned long long nW = (unsigned long long)gridDim.x * (blockDim.x / 64);
    const unsigned long long w = (unsigned long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    cur = __builtin_amdgcn_readfirstlane(batch_item(lp.firstPacket + lp.nPackets * w / nW, sCum, bt.nItems));
    multi_item(bt, cur, q, med, cdf, sMed, sCdf, E, sink);
  }

  for (;;) {
    if constexpr (RING > 0) {
      // a lane parks at most two deposits per iteration (its sub-step's and
      // its death's): drain every ring once one of them could overflow
      if (__ballot(sink.nRing >= RING - 2)) sink.drain();
    }
    const int st = E.state;
    const unsigned long long bFetch = __ballot(st == ST_FETCH);
    const unsigned long long bMove = __ballot(st == ST_MOVE);
    const unsigned long long bScat = __ballot(st == ST_SCATTER);
    const unsigned long long bFres = FRESNEL ? __ballot(st == ST_FRESNEL) : 0ull;
    const int cM = __popcll(bMove), cS = __popcll(bScat), cR = __popcll(bFres);
    // lanes waiting on an exhausted item are idle, not fetching
    const int cF = (BATCH == kMulti && curDone) ? 0 : __popcll(bFetch);
    const int active = cF + cM + cS + cR;
    if (active == 0) {
      if constexpr (BATCH == kMulti) {
        if (curDone && !__hip_atomic_load(&lp.ctr->aborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          int nxt = -1;
          for (int j = 1; j <= bt.nItems && nxt < 0; ++j) {
            const int c = cur + j < bt.nItems ? cur + j : cur + j - bt.nItems;
            const unsigned long long v = __hip_atomic_load(&bt.next[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__builtin_amdgcn_readfirstlane((unsigned)(v < bt.end[c]))) nxt = c;
          }
          if (nxt >= 0) {
            cur = nxt;
            curDone = false;
            multi_item(bt, cur, q, med, cdf, sMed, sCdf, E, sink);
            continue;
          }
        }
      }
      break;
    }
    // one state per iteration, chosen by the wavefront (see file comment)
    int phase = ST_MOVE;
    if ((float)cM < lp.moveFrac * (float)active) {
      int best = cM;
      if (cS > best) { best = cS; phase = ST_SCATTER; }
      if (FRESNEL && cR > best) { best = cR; phase = ST_FRESNEL; }
      if (cF > best) phase = ST_FETCH;
    }
    if (phase == ST_MOVE) {
      if (st == ST_MOVE) {
        ++nSub;
        E.do_move(pr, med, cdf, sink);
      }
    } else if (phase == ST_SCATTER) {
      if (st == ST_SCATTER) {
        ++nScat;
        E.do_scatter(pr, med, cdf, sink);
      }
    } else if (FRESNEL && phase == ST_FRESNEL) {
      if (st == ST_FRESNEL) {
        ++nIf;
        E.do_fresnel(pr, med, cdf, sink);
      }
    } else {
      // consecutive packet numbers for the waiting lanes, one atomic per wavefront
      const int leader = __ffsll((long long)bFetch) - 1;
      unsigned long long base = 0;
      if constexpr (BATCH == kMulti) {
        const unsigned long long end = bt.end[cur];
        if (lane == leader) {
          base = atomicAdd(&bt.next[cur], (unsigned long long)cF);
          if (atomicAdd(&lp.ctr->aborted, 0u)) base = end;
        }
        base = bcast64(base, leader);
        curDone = base + (unsigned long long)cF >= end;
        if (st == ST_FETCH) {
          const unsigned long long g = base + (unsigned long long)__popcll(bFetch & ((1ull << lane) - 1ull));
          if (g < end) {
            const unsigned long long k = g - lp.firstPacket;
            E.replay = false;
            E.slot = (uint32_t)k;
            const LaunchRecord L = launches[k];
            if (E.start_from(g - sCum[cur], L, pr, med, cdf)) {
              ++nLaunched;
            } else {
              atomicOr(&lp.ctr->aborted, 1u);
              E.state = ST_DONE;
            }
          }
        }
        continue;
      }
      if (lane == leader) {
        base = atomicAdd(&lp.ctr->nextPacket, (unsigned long long)cF);
        if (atomicAdd(&lp.ctr->aborted, 0u)) base = lp.nPackets;  // a failed launch stops the run
      }
      base = bcast64(base, leader);
      if (st == ST_FETCH) {
        const unsigned long long k = base + (unsigned long long)__popcll(bFetch & ((1ull << lane) - 1ull));
        if (k >= lp.nPackets) {
          E.state = ST_DONE;
        } else {
          E.replay = false;
          bool ok;
          E.slot = (uint32_t)k;
          if constexpr (CRITERIA) {
            ok = E.start(lp.firstPacket + k, pr, med, cdf);
          } else if constexpr (BATCH != kPlain) {
            const unsigned long long g = lp.firstPacket + k;
            const int it = batch_item(g, sCum, bt.nItems);
            med = sMed + it * bt.nM;
            cdf = sCdf + it * bt.nM;
            E.R.key0 = bt.key[2 * it];
            E.R.key1 = bt.key[2 * it + 1];
            E.R.stream = bt.stream[it];
            sink.NFR = bt.NFR + (size_t)it * bt.L;
            sink.item = it;
            const LaunchRecord L = launches[k];
            ok = E.start_from(g - sCum[it], L, pr, med, cdf);
          } else {
            const LaunchRecord L = launches[k];
            ok = E.start_from(lp.firstPacket + k, L, pr, med, cdf);
          }
          if (ok) {
            ++nLaunched;
          } else {
            atomicOr(&lp.ctr->aborted, 1u);
            E.state = ST_DONE;
          }
        }
      }
    }
  }
  sink.flush();
  if constexpr (RING > 0) sink.drain();
  if (nLaunched) atomicAdd(&lp.ctr->launched, nLaunched);
  if (sink.nCollected) atomicAdd(&lp.ctr->collected, sink.nCollected);
  if (nSub) atomicAdd(&lp.ctr->substeps, nSub);
  if (nScat) atomicAdd(&lp.ctr->scatters, nScat);
  if (nIf) atomicAdd(&lp.ctr->interfaces, nIf);
  if (sink.nRunaway) atomicAdd(&lp.ctr->runaway, (unsigned)sink.nRunaway);
}

// NFR normalisation (normalizeDepositionAndResetO, MCmatlablib.c:1204-1207):
// out[j] = dep[j] / (V * normfactor * mua[M[j]]) as single; clears dep.
__global__ void mc_normalize_nfr(double* dep, const uint8_t* M, const double* mua, double scale, float* out, int64_t L) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  out[j] = (float)(dep[j] / (scale * mua[M[j]]));
  dep[j] = 0.0;
}

// Sum the fluence copies 1..R-1 into copy 0 and clear them.
__global__ void mc_fold_replicas(double* dep, int64_t L, int R) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  double s = dep[j];
  for (int r = 1; r < R; ++r) {
    s += dep[(size_t)r * L + j];
    dep[(size_t)r * L + j] = 0.0;
  }
  dep[j] = s;
}

__global__ void mc_fill_normal_ri(float4* nri, const uint8_t* M, const Medium* media, int64_t L) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  nri[j].w = media[M[j]].ri;
}

// Example paths (updatePaths/addToPhotonPath/deletePhotonPath,
// MCmatlablib.c:465-514, and the master-thread bookkeeping of
// MCmatlab.c:266-279): one thread replays packets 0, 1, 2, ... of the
// wavelength, recording (x, y, z, weight) after every launch, sub-step and
// scattering, until nPaths packets have finished a recorded path.  Paths are
// separated by NaN columns as in the reference.  Writes at most `cap` columns;
// *nOut receives the number needed.
__global__ void mc_example_paths(McParams p, McLaunch lp, int64_t nPaths, double* out, int64_t cap, int64_t* nOut) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  __shared__ Medium sMed[256];
  __shared__ uint8_t sCdf[256];
  for (int k = 0; k < 256; ++k) {
    sMed[k] = p.media[k];
    sCdf[k] = p.cdfIdx[k];
  }
  int64_t n = 0, finished = 0;
  PacketRng R;
  R.key0 = lp.key0;
  R.key1 = lp.key1;
  R.stream = lp.stream;
  Photon P;
  const float hx = p.n[0] * 0.5f, hy = p.n[1] * 0.5f;
  bool started = false;
  auto put = [&](double x, double y, double z, double w) {
    if (n < cap) {
      out[4 * n] = x;
      out[4 * n + 1] = y;
      out[4 * n + 2] = z;
      out[4 * n + 3] = w;
    }
    ++n;
  };
  auto record = [&](bool teleported) {
    if (!(p.evalAtEnd || criteria_met(P, p))) return;
    const int nan = started ? (teleported ? 1 : 0) : 2;
    for (int k = 0; k < nan; ++k) put(NAN, NAN, NAN, NAN);
    put((P.i[0] - hx) * p.d[0], (P.i[1] - hy) * p.d[1], P.i[2] * p.d[2], P.weight);
    started = true;
  };
  for (uint64_t k = 0; k < lp.nPackets && finished < nPaths; ++k) {
    R.start(lp.firstPacket + k);
    started = false;
    if (!launch(P, p, sMed, R)) break;
    record(false);
    new_voxel_properties(P, p, sMed, sCdf);
    while (P.alive) {
      while (P.alive && P.stepLeft > 0.f) {
        if (P.substeps >= kMaxSubsteps) {
          P.alive = false;
          break;
        }
        const MoveResult mr = move<true>(P, p, sMed);
        if (mr.fresnel) fresnel(P, p, sMed, mr.jNew, mr.mNew, R);
        record(false);
        if (!P.sameVoxel) {
          const float i0 = P.i[0], i1 = P.i[1];
          check_escape(P, p);
          if (p.boundaryType == 3 && (P.i[0] != i0 || P.i[1] != i1)) record(true);
          if (P.alive) new_voxel_properties(P, p, sMed, sCdf);
        }
      }
      if (P.alive) roulette(P, R);
      if (P.alive) {
        scatter(P, p, p.cdfs, p.M[P.j], R);
        record(false);
      }
    }
    if (started) {
      if (p.evalAtEnd && !criteria_met(P, p)) {
        // drop the unfinished path back to its leading NaN pair
        for (; n > 1; --n) {
          if (n - 1 < cap && n - 2 < cap && isnan(out[4 * (n - 1)]) && isnan(out[4 * (n - 2)])) {
            n -= 2;
            break;
          }
        }
      } else {
        ++finished;
      }
    }
  }
  *nOut = n;
}

}  // namespace mcm

using namespace mcm;

namespace mcm {

// Select medium table, source table and refractive indices of wavelength iL.
int select_wavelength(mcm_mc_plan* pl, int64_t iL) {
  if (pl->curL == iL) return MCM_OK;
  McTables& T = pl->T;
  McParams& p = pl->prm;
  p.media = pl->media.as<Medium>() + (size_t)iL * 256;
  p.cdfIdx = pl->cdfIdx.as<uint8_t>() + (size_t)iL * 256;
  if (T.fresnelPossible) {
    const int tpb = 256;
    const int64_t nb = (T.L + tpb - 1) / tpb;
    hipLaunchKernelGGL(mc_fill_normal_ri, dim3((unsigned)nb), dim3(tpb), 0, pl->stream, pl->normals.as<float4>(), pl->M.as<uint8_t>(),
                       p.media, T.L);
    MCM_HIP(hipGetLastError());
  }
  if (pl->sourceDistribution) {
    // cumulative emission table in FP64 (MCmatlab.c:626-629)
    std::vector<double>& S = pl->Shost;
    S.assign((size_t)T.L + 1, 0.0);
    const float* pdf = pl->sourceDistribution + (size_t)iL * T.L;
    for (int64_t k = 1; k <= T.L; ++k) S[k] = S[k - 1] + (double)pdf[k - 1];
    const double total = S[T.L];
    if (!(total > 0)) return fail(MCM_ERR_INVALID_ARGUMENT, "sourceDistribution of wavelength %lld sums to zero", (long long)iL);
    pl->power = total * T.V;
    for (int64_t k = 1; k <= T.L; ++k) S[k] /= total;
    S[T.L] = 1.0;
    MCM_HIP(hipMemcpyAsync(pl->S.p, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, pl->stream));
  } else {
    pl->power = T.spectrum[(size_t)iL];
  }
  pl->curL = (int)iL;
  return MCM_OK;
}

}  // namespace mcm

namespace {

// Fold the fluence copies into copy 0 (on the plan's stream).
int fold_replicas(mcm_mc_plan* pl) {
  const int R = pl->prm.nfrReplicas;
  if (R <= 1 || !pl->T.calcNFR) return MCM_OK;
  const int tpb = 256;
  const int64_t nb = (pl->T.L + tpb - 1) / tpb;
  MCM_HIP(hipDeviceSynchronize());  // launches on other streams write the copies
  hipLaunchKernelGGL(mc_fold_replicas, dim3((unsigned)nb), dim3(tpb), 0, pl->stream, pl->NFR.as<double>(), pl->T.L, R);
  MCM_HIP(hipGetLastError());
  return MCM_OK;
}

// Number of fluence copies: folds the current tallies, then reallocates.
int set_replicas(mcm_mc_plan* pl, int R) {
  if (R == pl->prm.nfrReplicas || !pl->T.calcNFR) return MCM_OK;
  int rc = fold_replicas(pl);
  if (rc) return rc;
  MCM_HIP(hipStreamSynchronize(pl->stream));
  DevBuf nb;
  MCM_HIP(hipMalloc(&nb.p, (size_t)R * pl->T.L * sizeof(double)));
  MCM_HIP(hipMemsetAsync(nb.p, 0, (size_t)R * pl->T.L * sizeof(double), pl->stream));
  MCM_HIP(hipMemcpyAsync(nb.p, pl->NFR.p, (size_t)pl->T.L * sizeof(double), hipMemcpyDeviceToDevice, pl->stream));
  MCM_HIP(hipStreamSynchronize(pl->stream));
  std::swap(pl->NFR.p, nb.p);
  pl->prm.nfrReplicas = R;
  pl->prm.NFR = pl->NFR.as<double>();
  return MCM_OK;
}

int ensure_records(mcm_mc_plan* pl, uint64_t n) {
  if (n <= pl->recCap) return MCM_OK;
  MCM_HIP(hipDeviceSynchronize());  // earlier launches may use the old buffers on any stream
  for (DevBuf* b : {&pl->launchRec, &pl->escapeRec})
    if (b->p) {
      MCM_HIP(hipFree(b->p));
      b->p = nullptr;
    }
  if (hipMalloc(&pl->launchRec.p, (size_t)n * sizeof(LaunchRecord)) != hipSuccess ||
      hipMalloc(&pl->escapeRec.p, (size_t)n * sizeof(EscapeRecord)) != hipSuccess) {
    (void)hipGetLastError();
    pl->recCap = 0;
    return fail(MCM_ERR_OUT_OF_MEMORY, "Error: Out of memory (packet records for %llu packets)", (unsigned long long)n);
  }
  pl->recCap = n;
  return MCM_OK;
}

// One launch of n <= kMaxLaunchPackets packets.  Without deposition
// criteria: launch kernel -> transport -> escape tallies; with criteria the
// transport launches and tallies inline.  `bt` (batched launches) maps the
// global packet indices first .. first+n-1 to work items.
int launch_transport(mcm_mc_plan* pl, uint64_t first, uint64_t n, uint64_t seed, uint32_t stream, hipStream_t s,
                     const McBatch* bt = nullptr, int kind = kPlain, bool fresnel = false, bool tallies = false) {
  McLaunch lp;
  lp.firstPacket = first;
  lp.nPackets = n;
  lp.key0 = (uint32_t)seed;
  lp.key1 = (uint32_t)(seed >> 32);
  lp.stream = stream;
  lp.moveFrac = pl->moveFrac;
  lp.ctr = pl->counters.as<McCounters>();
  const bool fr = kind == kMulti ? fresnel : pl->T.fresnelPossible;
  const bool cr = kind == kPlain && pl->prm.criteriaActive != 0;
  McBatch none;
  std::memset(&none, 0, sizeof none);
  const McBatch& B = bt ? *bt : none;
  const dim3 b(pl->threads);
  MCM_HIP(hipMemsetAsync(pl->counters.p, 0, sizeof(unsigned long long), s));  // nextPacket
  if (cr) {
    const dim3 g(pl->blocks);
    if (fr) hipLaunchKernelGGL((mc_transport<true, true, kPlain>), g, b, 0, s, pl->prm, lp, B, nullptr, nullptr);
    else hipLaunchKernelGGL((mc_transport<false, true, kPlain>), g, b, 0, s, pl->prm, lp, B, nullptr, nullptr);
    MCM_HIP(hipGetLastError());
    return MCM_OK;
  }
  int rc = ensure_records(pl, n);
  if (rc) return rc;
  LaunchRecord* L = pl->launchRec.as<LaunchRecord>();
  EscapeRecord* E = pl->escapeRec.as<EscapeRecord>();
  const McTables& T = pl->T;
  bool anyNI = false;
  for (int f = 0; f < 6; ++f) anyNI = anyNI || T.nNI[f] > 0;
  const bool tally = kind == kMulti ? tallies : (T.base.useLC || T.base.farFieldRes || anyNI);
  if (kind == kMulti) {
    const dim3 g(pl->blocksMulti);
    hipLaunchKernelGGL(mc_multi_ranges, dim3(1), dim3(256), 0, s, B, (unsigned long long)first, (unsigned long long)n, B.next,
                       const_cast<unsigned long long*>(B.end));
    hipLaunchKernelGGL(mc_launch_packets<kMulti>, g, b, 0, s, pl->prm, lp, B, L, E);
    if (fr) hipLaunchKernelGGL((mc_transport<true, false, kMulti>), g, b, 0, s, pl->prm, lp, B, L, E);
    else hipLaunchKernelGGL((mc_transport<false, false, kMulti>), g, b, 0, s, pl->prm, lp, B, L, E);
    if (tally) hipLaunchKernelGGL(mc_tally_escapes<kMulti>, g, b, 0, s, pl->prm, lp, B, E);
  } else if (bt) {
    const dim3 g(pl->blocksBatch);
    hipLaunchKernelGGL(mc_launch_packets<kBatch>, g, b, 0, s, pl->prm, lp, B, L, E);
    if (fr) hipLaunchKernelGGL((mc_transport<true, false, kBatch>), g, b, 0, s, pl->prm, lp, B, L, E);
    else hipLaunchKernelGGL((mc_transport<false, false, kBatch>), g, b, 0, s, pl->prm, lp, B, L, E);
    if (tally) hipLaunchKernelGGL(mc_tally_escapes<kBatch>, g, b, 0, s, pl->prm, lp, B, E);
  } else {
    const dim3 g(pl->blocks);
    hipLaunchKernelGGL(mc_launch_packets<kPlain>, g, b, 0, s, pl->prm, lp, B, L, E);
    if (fr) hipLaunchKernelGGL((mc_transport<true, false, kPlain>), g, b, 0, s, pl->prm, lp, B, L, E);
    else hipLaunchKernelGGL((mc_transport<false, false, kPlain>), g, b, 0, s, pl->prm, lp, B, L, E);
    if (tally) hipLaunchKernelGGL(mc_tally_escapes<kPlain>, g, b, 0, s, pl->prm, lp, B, E);
  }
  MCM_HIP(hipGetLastError());
  return MCM_OK;
}

// Normalised outputs of one work item from its FP64 tally slices
// (normalizeDepositionAndResetO, MCmatlablib.c:1186-1250); null slices are
// skipped.
int finish_item(mcm_mc_plan* pl, int64_t iL, uint64_t nPhotons, const double* nfr, const double* img, const double* ff,
                const double* const ni[6], unsigned long long collected, double minutes, int64_t lanes, mcm_mc_output& o) {
  const McTables& T = pl->T;
  hipStream_t s = pl->stream;
  std::vector<double> h;
  auto slice = [&](const double* src, int64_t n, float* dst, double div) -> int {
    if (n <= 0 || !dst || !src) return MCM_OK;
    h.resize((size_t)n);
    MCM_HIP(hipMemcpy(h.data(), src, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < n; ++k) dst[k] = (float)(h[k] / div);
    return MCM_OK;
  };
  int rc;
  double normfactor = (double)nPhotons / T.spectrum[(size_t)iL];
  if (T.base.beamType == 2 && (T.base.boundaryType == 0 || T.base.boundaryType == 2)) normfactor /= kKillRange * kKillRange;
  if (T.calcNFR && o.NFR && nfr && nPhotons) {
    if (!pl->nfrOut.p) MCM_HIP(hipMalloc(&pl->nfrOut.p, (size_t)T.L * sizeof(float)));
    const int tpb = 256;
    const int64_t nb = (T.L + tpb - 1) / tpb;
    hipLaunchKernelGGL(mc_normalize_nfr, dim3((unsigned)nb), dim3(tpb), 0, s, const_cast<double*>(nfr), pl->M.as<uint8_t>(),
                       pl->muaNorm.as<double>() + (size_t)iL * T.nM, T.V * normfactor, pl->nfrOut.as<float>(), T.L);
    MCM_HIP(hipGetLastError());
    MCM_HIP(hipMemcpyAsync(o.NFR, pl->nfrOut.p, (size_t)T.L * sizeof(float), hipMemcpyDeviceToHost, s));
    MCM_HIP(hipStreamSynchronize(s));
  }
  if (T.base.useLC) {
    const int64_t LLC = (int64_t)T.lcRes0 * T.lcRes0;
    const double div = LLC > 1 ? (double)T.base.lcFSorNA * (double)T.base.lcFSorNA / LLC * normfactor : normfactor;
    if ((rc = slice(img, T.nImage, o.image, div))) return rc;
  }
  if (T.nFF && (rc = slice(ff, T.nFF, o.farField, normfactor))) return rc;
  float* nio[6] = {o.NI_xpos, o.NI_xneg, o.NI_ypos, o.NI_yneg, o.NI_zpos, o.NI_zneg};
  const double area[6] = {T.dA[0], T.dA[0], T.dA[1], T.dA[1], T.dA[2], T.dA[2]};
  for (int f = 0; f < 6; ++f)
    if (T.nNI[f] && (rc = slice(ni[f], T.nNI[f], nio[f], area[f] * normfactor))) return rc;
  o.nPhotons = (double)nPhotons;
  o.nPhotonsCollected = (double)collected;
  o.nThreads = (double)lanes;
  o.simulationTime = minutes;
  o.examplePaths = nullptr;
  o.examplePathsElems = 0;
  return MCM_OK;
}

int occupancy_blocks(int device, bool fr, bool cr, int kind = kPlain) {
  int per = 0;
  const void* k;
  if (kind == kMulti) k = fr ? (const void*)mc_transport<true, false, kMulti> : (const void*)mc_transport<false, false, kMulti>;
  else if (kind == kBatch) k = fr ? (const void*)mc_transport<true, false, kBatch> : (const void*)mc_transport<false, false, kBatch>;
  else if (fr && cr) k = (const void*)mc_transport<true, true, kPlain>;
  else if (fr) k = (const void*)mc_transport<true, false, kPlain>;
  else if (cr) k = (const void*)mc_transport<false, true, kPlain>;
  else k = (const void*)mc_transport<false, false, kPlain>;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 256, 0);
  if (e != hipSuccess || per < 1) per = 4;
  return per * cu_count(device);
}

}  // namespace

extern "C" {

int mcm_mc_output_sizes(const mcm_mc_input* in, mcm_mc_sizes* sz) {
  if (!in || !sz) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  McTables T;
  const std::string err = build_mc_tables(in, T);
  if (!err.empty()) return fail(MCM_ERR_INVALID_ARGUMENT, "%s", err.c_str());
  std::memset(sz, 0, sizeof *sz);
  sz->NFR = T.calcNFR ? T.L * T.nL : 0;
  sz->image = T.base.useLC ? T.nImage * T.nL : 0;
  sz->farField = T.nFF * T.nL;
  int64_t* ni[6] = {&sz->NI_xpos, &sz->NI_xneg, &sz->NI_ypos, &sz->NI_yneg, &sz->NI_zpos, &sz->NI_zneg};
  for (int f = 0; f < 6; ++f) *ni[f] = T.nNI[f] * T.nL;
  sz->image_dims[0] = T.lcRes0;
  sz->image_dims[1] = T.lcRes0;
  sz->image_dims[2] = T.lcRes1;
  sz->image_dims[3] = T.nL;
  const int64_t fd[6][2] = {{T.ny, T.nz}, {T.ny, T.nz}, {T.nx, T.nz}, {T.nx, T.nz}, {T.nx, T.ny}, {T.nx, T.ny}};
  for (int f = 0; f < 6; ++f) {
    sz->NI_dims[f][0] = T.nNI[f] ? (T.base.boundaryType == 2 ? kKillRange * T.nx : fd[f][0]) : 0;
    sz->NI_dims[f][1] = T.nNI[f] ? (T.base.boundaryType == 2 ? kKillRange * T.ny : fd[f][1]) : 0;
    sz->NI_dims[f][2] = T.nNI[f] ? T.nL : 0;
  }
  return MCM_OK;
}

int mcm_mc_plan_create(const mcm_mc_input* in, mcm_mc_plan** out) {
  if (!in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::unique_ptr<mcm_mc_plan> pl(new mcm_mc_plan());
  const std::string err = build_mc_tables(in, pl->T);
  if (!err.empty()) return fail(MCM_ERR_INVALID_ARGUMENT, "%s", err.c_str());
  int rc = use_device(in->GPUdevice);
  if (rc) return rc;
  pl->device = in->GPUdevice;
  MCM_HIP(hipStreamCreate(&pl->stream));
  McTables& T = pl->T;
  hipStream_t s = pl->stream;
  if ((rc = upload(pl->M, T.M0.data(), T.M0.size(), s))) return rc;
  if (T.fresnelPossible && (rc = upload(pl->normals, T.normals.data(), T.normals.size(), s))) return rc;
  if ((rc = upload(pl->media, T.media.data(), T.media.size(), s))) return rc;
  if ((rc = upload(pl->cdfIdx, T.cdfIdx.data(), T.cdfIdx.size(), s))) return rc;
  if ((rc = upload(pl->cdfs, T.cdfs.data(), T.cdfs.size(), s))) return rc;
  if ((rc = upload(pl->fpid1, T.fpid1.data(), T.fpid1.size(), s))) return rc;
  if ((rc = upload(pl->aid1, T.aid1.data(), T.aid1.size(), s))) return rc;
  if ((rc = upload(pl->fpid2, T.fpid2.data(), T.fpid2.size(), s))) return rc;
  if ((rc = upload(pl->aid2, T.aid2.data(), T.aid2.size(), s))) return rc;
  if ((rc = upload(pl->muaNorm, T.muaNorm.data(), T.muaNorm.size(), s))) return rc;
  pl->sourceDistribution = in->sourceDistribution;
  if (in->sourceDistribution && (rc = upload<double>(pl->S, nullptr, (size_t)T.L + 1, s))) return rc;
  if ((rc = zeros(pl->NFR, (T.calcNFR ? (size_t)T.L : 1) * sizeof(double), s))) return rc;
  if ((rc = zeros(pl->image, (size_t)std::max<int64_t>(T.nImage, 1) * sizeof(double), s))) return rc;
  if ((rc = zeros(pl->FF, (size_t)std::max<int64_t>(T.nFF, 1) * sizeof(double), s))) return rc;
  for (int f = 0; f < 6; ++f)
    if ((rc = zeros(pl->NI[f], (size_t)std::max<int64_t>(T.nNI[f], 1) * sizeof(double), s))) return rc;
  if ((rc = zeros(pl->counters, sizeof(McCounters), s))) return rc;
  McParams& p = pl->prm;
  p = T.base;
  p.M = pl->M.as<uint8_t>();
  p.normalRI = T.fresnelPossible ? pl->normals.as<float4>() : nullptr;
  p.cdfs = pl->cdfs.as<float>();
  p.FPIDdist1 = pl->fpid1.as<float>();
  p.AIDdist1 = pl->aid1.as<float>();
  p.FPIDdist2 = pl->fpid2.as<float>();
  p.AIDdist2 = pl->aid2.as<float>();
  p.S = in->sourceDistribution ? pl->S.as<double>() : nullptr;
  p.NFR = pl->NFR.as<double>();
  p.image = pl->image.as<double>();
  p.FF = pl->FF.as<double>();
  for (int f = 0; f < 6; ++f) p.NI[f] = pl->NI[f].as<double>();
  pl->threads = 256;
  pl->blocks = occupancy_blocks(pl->device, T.fresnelPossible, p.criteriaActive != 0);
  pl->blocksBatch = occupancy_blocks(pl->device, T.fresnelPossible, false, kBatch);
  pl->blocksMulti = std::min(occupancy_blocks(pl->device, true, false, kMulti), occupancy_blocks(pl->device, false, false, kMulti));
  if (const char* e = std::getenv("MCM_MC_MOVE_FRAC")) pl->moveFrac = (float)std::atof(e);
  if (const char* e = std::getenv("MCM_MC_BLOCKS")) pl->blocks = pl->blocksBatch = pl->blocksMulti = std::max(1, std::atoi(e));
  if ((rc = select_wavelength(pl.get(), 0))) return rc;
  MCM_HIP(hipStreamSynchronize(s));
  *out = pl.release();
  return MCM_OK;
}

int mcm_mc_plan_simulate(mcm_mc_plan* pl, int64_t iL, uint64_t nPhotons, uint64_t seed, uint64_t firstPhoton, void* stream) {
  if (!pl) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  if (iL < 0 || iL >= pl->T.nL) return fail(MCM_ERR_INVALID_ARGUMENT, "wavelength index out of range");
  MCM_HIP(hipSetDevice(pl->device));
  hipStream_t s = stream ? (hipStream_t)stream : pl->stream;
  if (s != pl->stream) MCM_HIP(hipStreamSynchronize(pl->stream));
  int rc = select_wavelength(pl, iL);
  if (rc) return rc;
  if (s != pl->stream) MCM_HIP(hipStreamSynchronize(pl->stream));
  // launched/collected/... accumulate over launches until normalize
  for (uint64_t done = 0; done < nPhotons;) {
    const uint64_t n = std::min<uint64_t>(nPhotons - done, kMaxLaunchPackets);
    if ((rc = launch_transport(pl, firstPhoton + done, n, seed, (uint32_t)iL, s))) return rc;
    done += n;
  }
  return MCM_OK;
}

int mcm_mc_plan_tallies(mcm_mc_plan* pl, const double** d_deposit, uint64_t* nPhotons) {
  if (!pl) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  MCM_HIP(hipSetDevice(pl->device));
  MCM_HIP(hipStreamSynchronize(pl->stream));
  McCounters c;
  MCM_HIP(hipMemcpy(&c, pl->counters.p, sizeof c, hipMemcpyDeviceToHost));
  int rc = fold_replicas(pl);
  if (rc) return rc;
  MCM_HIP(hipStreamSynchronize(pl->stream));
  if (d_deposit) *d_deposit = pl->NFR.as<double>();
  if (nPhotons) *nPhotons = c.launched;
  return MCM_OK;
}

int mcm_mc_plan_normalize(mcm_mc_plan* pl, int64_t iL, mcm_mc_output* out) {
  if (!pl || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  MCM_HIP(hipSetDevice(pl->device));
  McTables& T = pl->T;
  hipStream_t s = pl->stream;
  MCM_HIP(hipDeviceSynchronize());
  McCounters c;
  MCM_HIP(hipMemcpy(&c, pl->counters.p, sizeof c, hipMemcpyDeviceToHost));
  pl->stats = c;
  // normfactor (MCmatlablib.c:1195-1200)
  double normfactor = (double)c.launched / pl->power;
  if (pl->sourceDistribution) normfactor /= pl->power;
  else if (T.base.beamType == 2 && (T.base.boundaryType == 0 || T.base.boundaryType == 2)) normfactor /= kKillRange * kKillRange;
  int rc0 = fold_replicas(pl);
  if (rc0) return rc0;
  if (T.calcNFR && out->NFR) {
    if (!pl->nfrOut.p) MCM_HIP(hipMalloc(&pl->nfrOut.p, (size_t)T.L * sizeof(float)));
    const int tpb = 256;
    const int64_t nb = (T.L + tpb - 1) / tpb;
    hipLaunchKernelGGL(mc_normalize_nfr, dim3((unsigned)nb), dim3(tpb), 0, s, pl->NFR.as<double>(), pl->M.as<uint8_t>(),
                       pl->muaNorm.as<double>() + (size_t)iL * T.nM, T.V * normfactor, pl->nfrOut.as<float>(), T.L);
    MCM_HIP(hipGetLastError());
    MCM_HIP(hipMemcpyAsync(out->NFR + (size_t)iL * T.L, pl->nfrOut.p, (size_t)T.L * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  std::vector<double> h;
  auto small = [&](DevBuf& b, int64_t n, float* dst, double div) -> int {
    if (n <= 0) return MCM_OK;
    h.resize((size_t)n);
    MCM_HIP(hipMemcpyAsync(h.data(), b.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
    MCM_HIP(hipStreamSynchronize(s));
    if (dst)
      for (int64_t k = 0; k < n; ++k) dst[k] = (float)(h[k] / div);
    MCM_HIP(hipMemsetAsync(b.p, 0, (size_t)n * sizeof(double), s));
    return MCM_OK;
  };
  int rc;
  if (T.base.useLC) {
    const int64_t LLC = (int64_t)T.lcRes0 * T.lcRes0;
    const double div = LLC > 1 ? (double)T.base.lcFSorNA * (double)T.base.lcFSorNA / LLC * normfactor : normfactor;
    if ((rc = small(pl->image, T.nImage, out->image ? out->image + (size_t)iL * T.nImage : nullptr, div))) return rc;
  }
  if (T.nFF && (rc = small(pl->FF, T.nFF, out->farField ? out->farField + (size_t)iL * T.nFF : nullptr, normfactor))) return rc;
  float* ni[6] = {out->NI_xpos, out->NI_xneg, out->NI_ypos, out->NI_yneg, out->NI_zpos, out->NI_zneg};
  const double area[6] = {T.dA[0], T.dA[0], T.dA[1], T.dA[1], T.dA[2], T.dA[2]};
  for (int f = 0; f < 6; ++f)
    if (T.nNI[f] && (rc = small(pl->NI[f], T.nNI[f], ni[f] ? ni[f] + (size_t)iL * T.nNI[f] : nullptr, area[f] * normfactor))) return rc;
  MCM_HIP(hipMemsetAsync(pl->counters.p, 0, sizeof(McCounters), s));
  MCM_HIP(hipStreamSynchronize(s));
  pl->lastLaunched = c.launched;
  pl->lastCollected = c.collected;
  return MCM_OK;
}

int mcm_mc_plan_autotune(mcm_mc_plan* pl, int64_t iL, uint64_t nPackets, double* ms) {
  if (!pl) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  if (iL < 0 || iL >= pl->T.nL) return fail(MCM_ERR_INVALID_ARGUMENT, "wavelength index out of range");
  MCM_HIP(hipSetDevice(pl->device));
  const McTables& T = pl->T;
  if (nPackets == 0) nPackets = 1ull << 21;
  const int cand[4] = {1, 2, 4, 8};
  double t[4] = {0, 0, 0, 0};
  int best = pl->prm.nfrReplicas;
  double bestMs = INFINITY;
  hipEvent_t e0, e1;
  MCM_HIP(hipEventCreate(&e0));
  MCM_HIP(hipEventCreate(&e1));
  int rc = MCM_OK;
  for (int c = 0; c < 4 && rc == MCM_OK; ++c) {
    const int R = cand[c];
    if (R > 1 && (!T.calcNFR || (double)R * T.L * 8 > 2147483648.0)) continue;
    if ((rc = set_replicas(pl, R))) break;
    if ((rc = mcm_mc_plan_simulate(pl, iL, std::min<uint64_t>(nPackets, 1u << 16), 0x7E57ull, 0, nullptr))) break;  // warm-up
    (void)hipEventRecord(e0, pl->stream);
    if ((rc = mcm_mc_plan_simulate(pl, iL, nPackets, 0x7E57ull, 1ull << 40, nullptr))) break;
    (void)hipEventRecord(e1, pl->stream);
    (void)hipEventSynchronize(e1);
    float f = 0.f;
    (void)hipEventElapsedTime(&f, e0, e1);
    t[c] = f;
    bestMs = std::min<double>(bestMs, f);
  }
  // the fewest copies within 10 % of the fastest: more copies only pay off
  // clearly (contended narrow beams); otherwise keep the working set small
  for (int c = 3; c >= 0; --c)
    if (t[c] > 0 && t[c] <= 1.10 * bestMs) best = cand[c];
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = set_replicas(pl, best))) return rc;
  // discard the trial tallies
  hipStream_t s = pl->stream;
  MCM_HIP(hipDeviceSynchronize());
  if (T.calcNFR) MCM_HIP(hipMemsetAsync(pl->NFR.p, 0, (size_t)pl->prm.nfrReplicas * T.L * sizeof(double), s));
  MCM_HIP(hipMemsetAsync(pl->image.p, 0, (size_t)std::max<int64_t>(T.nImage, 1) * sizeof(double), s));
  MCM_HIP(hipMemsetAsync(pl->FF.p, 0, (size_t)std::max<int64_t>(T.nFF, 1) * sizeof(double), s));
  for (int f = 0; f < 6; ++f) MCM_HIP(hipMemsetAsync(pl->NI[f].p, 0, (size_t)std::max<int64_t>(T.nNI[f], 1) * sizeof(double), s));
  MCM_HIP(hipMemsetAsync(pl->counters.p, 0, sizeof(McCounters), s));
  MCM_HIP(hipStreamSynchronize(s));
  if (ms)
    for (int c = 0; c < 4; ++c) ms[c] = t[c];
  return MCM_OK;
}

int mcm_mc_plan_set_fluence_copies(mcm_mc_plan* pl, int32_t copies) {
  if (!pl) return fail(MCM_ERR_INVALID_ARGUMENT, "null plan");
  if (copies != 1 && copies != 2 && copies != 4 && copies != 8) return fail(MCM_ERR_INVALID_ARGUMENT, "fluence copies must be 1, 2, 4 or 8");
  if (copies > 1 && !pl->T.calcNFR) return fail(MCM_ERR_INVALID_ARGUMENT, "fluence copies need calcNFR");
  MCM_HIP(hipSetDevice(pl->device));
  return set_replicas(pl, copies);
}

int mcm_mc_plan_counters(mcm_mc_plan* pl, double* out) {
  if (!pl || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  MCM_HIP(hipSetDevice(pl->device));
  MCM_HIP(hipDeviceSynchronize());
  McCounters c;
  MCM_HIP(hipMemcpy(&c, pl->counters.p, sizeof c, hipMemcpyDeviceToHost));
  out[0] = (double)c.launched;
  out[1] = (double)c.collected;
  out[2] = (double)c.substeps;
  out[3] = (double)c.scatters;
  out[4] = (double)c.interfaces;
  out[5] = (double)c.runaway;
  out[6] = (double)c.aborted;
  out[7] = (double)pl->blocks * pl->threads;
  return MCM_OK;
}

int mcm_mc_batch_run(mcm_mc_plan* pl, const mcm_mc_item* items, int64_t nItems, mcm_mc_output* outs) {
  if (!pl || !items || !outs) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  if (nItems < 1 || nItems > kBatchMaxItems) return fail(MCM_ERR_INVALID_ARGUMENT, "nItems must be 1..%d", kBatchMaxItems);
  McTables& T = pl->T;
  if (nItems * T.nM > kBatchMaxMedia) return fail(MCM_ERR_UNSUPPORTED, "nItems * nM exceeds %d", kBatchMaxMedia);
  if (pl->prm.criteriaActive) return fail(MCM_ERR_UNSUPPORTED, "batched launches run without deposition criteria");
  if (pl->sourceDistribution) return fail(MCM_ERR_UNSUPPORTED, "batched launches need a beam source");
  for (int64_t i = 0; i < nItems; ++i) {
    if (items[i].iL < 0 || items[i].iL >= T.nL) return fail(MCM_ERR_INVALID_ARGUMENT, "item %lld: wavelength out of range", (long long)i);
    if (T.fresnelPossible)
      for (int m = 0; m < T.nM; ++m)
        if (T.media[(size_t)items[i].iL * 256 + m].ri != T.media[(size_t)items[0].iL * 256 + m].ri)
          return fail(MCM_ERR_UNSUPPORTED, "batched items need the same refractive indices");
  }
  MCM_HIP(hipSetDevice(pl->device));
  int rc = select_wavelength(pl, items[0].iL);  // interface normals and their indices
  if (rc) return rc;
  hipStream_t s = pl->stream;
  const int nI = (int)nItems, nM = T.nM;
  std::vector<unsigned long long> cum((size_t)nI + 1, 0);
  std::vector<uint32_t> key(2 * (size_t)nI), strm((size_t)nI);
  std::vector<Medium> med((size_t)nI * nM);
  std::vector<uint8_t> cdf((size_t)nI * nM);
  for (int i = 0; i < nI; ++i) {
    cum[i + 1] = cum[i] + items[i].nPhotons;
    key[2 * i] = (uint32_t)items[i].seed;
    key[2 * i + 1] = (uint32_t)(items[i].seed >> 32);
    strm[i] = items[i].stream;
    for (int m = 0; m < nM; ++m) {
      med[(size_t)i * nM + m] = T.media[(size_t)items[i].iL * 256 + m];
      cdf[(size_t)i * nM + m] = T.cdfIdx[(size_t)items[i].iL * 256 + m];
    }
  }
  DevBuf dCum, dKey, dStrm, dMed, dCdf, dNFR, dImg, dFF, dNI[6], dColl;
  if ((rc = upload(dCum, cum.data(), cum.size(), s)) || (rc = upload(dKey, key.data(), key.size(), s)) ||
      (rc = upload(dStrm, strm.data(), strm.size(), s)) || (rc = upload(dMed, med.data(), med.size(), s)) ||
      (rc = upload(dCdf, cdf.data(), cdf.size(), s)))
    return rc;
  McBatch B;
  std::memset(&B, 0, sizeof B);
  B.nItems = nI;
  B.nM = nM;
  B.cum = dCum.as<unsigned long long>();
  B.key = dKey.as<uint32_t>();
  B.stream = dStrm.as<uint32_t>();
  B.media = dMed.as<Medium>();
  B.cdfIdx = dCdf.as<uint8_t>();
  B.L = T.L;
  B.nImage = T.nImage;
  B.nFF = T.nFF;
  if ((rc = zeros(dNFR, (T.calcNFR ? (size_t)nI * T.L : 1) * sizeof(double), s)) ||
      (rc = zeros(dImg, (size_t)std::max<int64_t>(nI * T.nImage, 1) * sizeof(double), s)) ||
      (rc = zeros(dFF, (size_t)std::max<int64_t>(nI * T.nFF, 1) * sizeof(double), s)) ||
      (rc = zeros(dColl, (size_t)nI * sizeof(unsigned long long), s)))
    return rc;
  for (int f = 0; f < 6; ++f) {
    B.nNI[f] = T.nNI[f];
    if ((rc = zeros(dNI[f], (size_t)std::max<int64_t>(nI * T.nNI[f], 1) * sizeof(double), s))) return rc;
    B.NI[f] = dNI[f].as<double>();
  }
  B.NFR = dNFR.as<double>();
  B.image = dImg.as<double>();
  B.FF = dFF.as<double>();
  B.collected = dColl.as<unsigned long long>();
  MCM_HIP(hipMemsetAsync(pl->counters.p, 0, sizeof(McCounters), s));
  const auto t0 = std::chrono::steady_clock::now();
  const unsigned long long total = cum[nI];
  for (unsigned long long done = 0; done < total;) {
    const uint64_t n = std::min<unsigned long long>(total - done, kMaxLaunchPackets);
    if ((rc = launch_transport(pl, done, n, 0, 0, s, &B))) return rc;
    done += n;
  }
  MCM_HIP(hipStreamSynchronize(s));
  const double minutes = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 60.0;
  McCounters c;
  MCM_HIP(hipMemcpy(&c, pl->counters.p, sizeof c, hipMemcpyDeviceToHost));
  pl->stats = c;
  MCM_HIP(hipMemsetAsync(pl->counters.p, 0, sizeof(McCounters), s));
  if (c.aborted) return fail(MCM_ERR_INVALID_ARGUMENT, "all packets launch outside the simulation cuboid");
  std::vector<unsigned long long> coll((size_t)nI);
  MCM_HIP(hipMemcpy(coll.data(), dColl.p, coll.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int i = 0; i < nI; ++i) {
    const double* ni[6];
    for (int f = 0; f < 6; ++f) ni[f] = dNI[f].as<double>() + (size_t)i * T.nNI[f];
    if ((rc = finish_item(pl, items[i].iL, items[i].nPhotons, T.calcNFR ? dNFR.as<double>() + (size_t)i * T.L : nullptr,
                          dImg.as<double>() + (size_t)i * T.nImage, dFF.as<double>() + (size_t)i * T.nFF, ni, coll[i], minutes,
                          (int64_t)pl->blocksBatch * pl->threads, outs[i])))
      return rc;
  }
  return MCM_OK;
}

int mcm_mc_multi_run(mcm_mc_plan* const* plans, int64_t nPlans, const mcm_mc_multi_item* items, int64_t nItems, mcm_mc_output* outs) {
  if (!plans || !items || !outs || nPlans < 1) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  if (nItems < 1 || nItems > kBatchMaxItems) return fail(MCM_ERR_INVALID_ARGUMENT, "nItems must be 1..%d", kBatchMaxItems);
  for (int64_t k = 0; k < nPlans; ++k) {
    if (!plans[k]) return fail(MCM_ERR_INVALID_ARGUMENT, "plan %lld is null", (long long)k);
    if (plans[k]->device != plans[0]->device) return fail(MCM_ERR_INVALID_ARGUMENT, "the plans of a multi-grid launch share one device");
  }
  const int nI = (int)nItems;
  int nMmax = 1;
  std::vector<int64_t> firstL((size_t)nPlans, -1);
  for (int i = 0; i < nI; ++i) {
    const mcm_mc_multi_item& it = items[i];
    if (it.plan < 0 || it.plan >= nPlans) return fail(MCM_ERR_INVALID_ARGUMENT, "item %d: plan index out of range", i);
    mcm_mc_plan* pl = plans[it.plan];
    const McTables& T = pl->T;
    if (it.iL < 0 || it.iL >= T.nL) return fail(MCM_ERR_INVALID_ARGUMENT, "item %d: wavelength out of range", i);
    if (pl->prm.criteriaActive) return fail(MCM_ERR_UNSUPPORTED, "multi-grid launches run without deposition criteria");
    if (pl->sourceDistribution) return fail(MCM_ERR_UNSUPPORTED, "multi-grid launches need beam sources");
    int64_t& l0 = firstL[(size_t)it.plan];
    if (l0 < 0) l0 = it.iL;
    if (T.fresnelPossible)
      for (int m = 0; m < T.nM; ++m)
        if (T.media[(size_t)it.iL * 256 + m].ri != T.media[(size_t)l0 * 256 + m].ri)
          return fail(MCM_ERR_UNSUPPORTED, "items of one plan need the same refractive indices");
    nMmax = std::max(nMmax, T.nM);
  }
  if (nI * nMmax > kBatchMaxMedia) return fail(MCM_ERR_UNSUPPORTED, "nItems * (largest nM) exceeds %d", kBatchMaxMedia);
  mcm_mc_plan* pl0 = plans[0];
  MCM_HIP(hipSetDevice(pl0->device));
  int rc;
  // every used plan's tables at its items' refractive indices
  for (int64_t k = 0; k < nPlans; ++k)
    if (firstL[(size_t)k] >= 0) {
      if ((rc = select_wavelength(plans[k], firstL[(size_t)k]))) return rc;
      MCM_HIP(hipStreamSynchronize(plans[k]->stream));
    }
  hipStream_t s = pl0->stream;
  // tally slices of every item in one buffer per kind
  std::vector<int64_t> off((size_t)nI * 9);
  int64_t tot[9] = {0};
  bool tallies = false;
  for (int i = 0; i < nI; ++i) {
    const McTables& T = plans[items[i].plan]->T;
    const int64_t sz[9] = {T.calcNFR ? T.L : 0, T.nImage, T.nFF, T.nNI[0], T.nNI[1], T.nNI[2], T.nNI[3], T.nNI[4], T.nNI[5]};
    for (int k = 0; k < 9; ++k) {
      off[(size_t)i * 9 + k] = tot[k];
      tot[k] += sz[k];
      if (k > 0 && sz[k] > 0) tallies = true;
    }
    tallies = tallies || T.base.useLC || T.base.farFieldRes;
  }
  DevBuf dTal[9], dColl, dOff, dParams, dGeom;
  for (int k = 0; k < 9; ++k)
    if ((rc = zeros(dTal[k], (size_t)std::max<int64_t>(tot[k], 1) * sizeof(double), s))) return rc;
  if ((rc = zeros(dColl, (size_t)nI * sizeof(unsigned long long), s)) || (rc = upload(dOff, off.data(), off.size(), s))) return rc;
  std::vector<McParams> params((size_t)nPlans);
  std::vector<McGeom> geom((size_t)nPlans);
  for (int64_t k = 0; k < nPlans; ++k) {
    const McParams& p = plans[k]->prm;
    params[(size_t)k] = p;
    McGeom& g = geom[(size_t)k];
    for (int a = 0; a < 3; ++a) {
      g.n[a] = p.n[a];
      g.d[a] = p.d[a];
      g.inv_d[a] = p.inv_d[a];
    }
    g.tooOld = p.tooOld;
    g.boundaryType = p.boundaryType;
    g.calcNFR = p.calcNFR;
    g.needTime = p.needTime;
    g.useLC = p.useLC;
    g.farFieldRes = p.farFieldRes;
    g.M = p.M;
    g.normalRI = p.normalRI;
    g.cdfs = p.cdfs;
  }
  if ((rc = upload(dParams, params.data(), params.size(), s)) || (rc = upload(dGeom, geom.data(), geom.size(), s))) return rc;
  MCM_HIP(hipMemsetAsync(pl0->counters.p, 0, sizeof(McCounters), s));
  const auto t0 = std::chrono::steady_clock::now();
  // one sequence of launches per interface class (Fresnel or matched), so
  // that every item runs the kernel its plan's own launches use
  for (int cls = 0; cls < 2; ++cls) {
    std::vector<int> idx;
    for (int i = 0; i < nI; ++i)
      if (plans[items[i].plan]->T.fresnelPossible == (cls == 0)) idx.push_back(i);
    if (idx.empty()) continue;
    const int n = (int)idx.size();
    std::vector<unsigned long long> cum((size_t)n + 1, 0);
    std::vector<uint32_t> key(2 * (size_t)n), strm((size_t)n);
    std::vector<int> model((size_t)n);
    std::vector<int64_t> offc((size_t)n * 9);
    std::vector<Medium> med((size_t)n * nMmax);
    std::vector<uint8_t> cdf((size_t)n * nMmax, 0);
    for (int j = 0; j < n; ++j) {
      const mcm_mc_multi_item& it = items[idx[(size_t)j]];
      const McTables& T = plans[it.plan]->T;
      cum[(size_t)j + 1] = cum[(size_t)j] + it.nPhotons;
      key[2 * (size_t)j] = (uint32_t)it.seed;
      key[2 * (size_t)j + 1] = (uint32_t)(it.seed >> 32);
      strm[(size_t)j] = it.stream;
      model[(size_t)j] = (int)it.plan;
      for (int k = 0; k < 9; ++k) offc[(size_t)j * 9 + k] = off[(size_t)idx[(size_t)j] * 9 + k];
      for (int m = 0; m < T.nM; ++m) {
        med[(size_t)j * nMmax + m] = T.media[(size_t)it.iL * 256 + m];
        cdf[(size_t)j * nMmax + m] = T.cdfIdx[(size_t)it.iL * 256 + m];
      }
    }
    DevBuf dCum, dKey, dStrm, dModel, dOffc, dMed, dCdf, dCollc, dNext, dEnd;
    if ((rc = upload(dCum, cum.data(), cum.size(), s)) || (rc = upload(dKey, key.data(), key.size(), s)) ||
        (rc = upload(dStrm, strm.data(), strm.size(), s)) || (rc = upload(dModel, model.data(), model.size(), s)) ||
        (rc = upload(dOffc, offc.data(), offc.size(), s)) || (rc = upload(dMed, med.data(), med.size(), s)) ||
        (rc = upload(dCdf, cdf.data(), cdf.size(), s)) || (rc = zeros(dCollc, (size_t)n * sizeof(unsigned long long), s)) ||
        (rc = zeros(dNext, (size_t)n * sizeof(unsigned long long), s)) || (rc = zeros(dEnd, (size_t)n * sizeof(unsigned long long), s)))
      return rc;
    McBatch B;
    std::memset(&B, 0, sizeof B);
    B.nItems = n;
    B.nM = nMmax;
    B.cum = dCum.as<unsigned long long>();
    B.key = dKey.as<uint32_t>();
    B.stream = dStrm.as<uint32_t>();
    B.media = dMed.as<Medium>();
    B.cdfIdx = dCdf.as<uint8_t>();
    B.NFR = dTal[0].as<double>();
    B.image = dTal[1].as<double>();
    B.FF = dTal[2].as<double>();
    for (int f = 0; f < 6; ++f) B.NI[f] = dTal[3 + f].as<double>();
    B.collected = dCollc.as<unsigned long long>();
    B.model = dModel.as<int>();
    B.params = dParams.as<McParams>();
    B.geom = dGeom.as<McGeom>();
    B.off = dOffc.as<int64_t>();
    B.next = dNext.as<unsigned long long>();
    B.end = dEnd.as<unsigned long long>();
    // launches of whole items where they fit (an item split over two
    // launches pays the tail of its longest-lived packets twice)
    const unsigned long long total = cum[(size_t)n];
    for (unsigned long long done = 0; done < total;) {
      unsigned long long stop = done + std::min<unsigned long long>(total - done, kMaxLaunchPackets);
      for (int j = n; j > 0; --j)
        if (cum[(size_t)j] <= stop && cum[(size_t)j] > done) {
          stop = cum[(size_t)j] - done >= kMaxLaunchPackets / 4 || cum[(size_t)j] == total ? cum[(size_t)j] : stop;
          break;
        }
      if ((rc = launch_transport(pl0, done, stop - done, 0, 0, s, &B, kMulti, cls == 0, tallies))) return rc;
      done = stop;
    }
    std::vector<unsigned long long> cc((size_t)n);
    MCM_HIP(hipMemcpyAsync(cc.data(), dCollc.p, cc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MCM_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < n; ++j) outs[idx[(size_t)j]].nPhotonsCollected = (double)cc[(size_t)j];
  }
  MCM_HIP(hipStreamSynchronize(s));
  const double minutes = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 60.0;
  McCounters c;
  MCM_HIP(hipMemcpy(&c, pl0->counters.p, sizeof c, hipMemcpyDeviceToHost));
  pl0->stats = c;
  MCM_HIP(hipMemsetAsync(pl0->counters.p, 0, sizeof(McCounters), s));
  if (c.aborted) return fail(MCM_ERR_INVALID_ARGUMENT, "all packets launch outside the simulation cuboid");
  for (int i = 0; i < nI; ++i) {
    mcm_mc_plan* pl = plans[items[i].plan];
    const int64_t* o = &off[(size_t)i * 9];
    const double* ni[6];
    for (int f = 0; f < 6; ++f) ni[f] = dTal[3 + f].as<double>() + o[3 + f];
    const unsigned long long coll = (unsigned long long)outs[i].nPhotonsCollected;
    if ((rc = finish_item(pl, items[i].iL, items[i].nPhotons, pl->T.calcNFR ? dTal[0].as<double>() + o[0] : nullptr,
                          dTal[1].as<double>() + o[1], dTal[2].as<double>() + o[2], ni, coll, minutes,
                          (int64_t)pl0->blocksMulti * pl0->threads, outs[i])))
      return rc;
  }
  return MCM_OK;
}

int mcm_mc_plan_destroy(mcm_mc_plan* pl) {
  if (!pl) return MCM_OK;
  (void)hipSetDevice(pl->device);
  (void)hipStreamSynchronize(pl->stream);
  delete pl;
  return MCM_OK;
}

// The whole MEX call: every wavelength, host buffers in and out.
int mcm_mc_run(const mcm_mc_input* in, mcm_mc_output* out) {
  if (!in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  mcm_mc_plan* raw = nullptr;
  int rc = mcm_mc_plan_create(in, &raw);
  if (rc) return rc;
  std::unique_ptr<mcm_mc_plan, int (*)(mcm_mc_plan*)> pl(raw, mcm_mc_plan_destroy);
  McTables& T = pl->T;
  const unsigned long long nReq = T.timed ? ~0ull : (unsigned long long)T.nPhotonsRequested;
  // large runs first pick the fastest number of fluence copies for this model
  if ((T.timed || nReq >= 20000000ull) && !pl->prm.criteriaActive && T.calcNFR) {
    if ((rc = mcm_mc_plan_autotune(pl.get(), 0, 0, nullptr))) return rc;
  }
  double cumPhotons = 0, cumCollected = 0, cumMinutes = 0;
  bool aborted = false, interrupted = false;
  const int nL = T.nL;
  for (int iL = 0; iL < nL && !aborted && !interrupted; ++iL) {
    unsigned long long reqL = 0;
    if (!T.timed)
      reqL = iL == nL - 1 ? nReq - (unsigned long long)(T.requestCollected ? cumCollected : cumPhotons) : nReq / (unsigned long long)nL;
    const double budgetUs = T.timed ? T.simulationTimeRequested * 60e6 / nL : INFINITY;
    if ((rc = select_wavelength(pl.get(), iL))) return rc;
    MCM_HIP(hipStreamSynchronize(pl->stream));
    MCM_HIP(hipMemsetAsync(pl->counters.p, 0, sizeof(McCounters), pl->stream));
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t next = 0;
    McCounters c{};
    // chunk size: large enough to amortise the launch tail, small enough to
    // poll for interrupts and time budgets
    uint64_t chunk = (uint64_t)pl->blocks * pl->threads * 32;
    for (;;) {
      uint64_t n;
      if (T.timed) n = chunk;
      else if (T.requestCollected) {
        if (c.collected >= reqL) break;
        const double rate = c.launched ? (double)c.collected / (double)c.launched : 0.0;
        n = rate > 0 ? (uint64_t)std::min(4.0 * chunk, std::max(1.0, 1.1 * (reqL - c.collected) / rate)) : chunk;
      } else {
        if (c.launched >= reqL) break;
        n = std::min<uint64_t>(reqL - c.launched, 8 * chunk);
      }
      if ((rc = mcm_mc_plan_simulate(pl.get(), iL, n, in->seed, next, nullptr))) return rc;
      next += n;
      MCM_HIP(hipStreamSynchronize(pl->stream));
      MCM_HIP(hipMemcpy(&c, pl->counters.p, sizeof c, hipMemcpyDeviceToHost));
      if (c.aborted) {
        aborted = true;
        break;
      }
      if (in->isInterruptPending && in->isInterruptPending()) {
        interrupted = true;
        break;
      }
      const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
      if (T.timed && us >= budgetUs) break;
      if (T.timed && us > 0) chunk = std::max<uint64_t>(1024, (uint64_t)(chunk * std::min(4.0, std::max(0.25, 1e5 / us * (double)next / chunk / 8))));
    }
    const double minutes = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 60.0;
    cumPhotons += (double)c.launched;
    cumCollected += (double)c.collected;
    cumMinutes += minutes;
    if ((rc = mcm_mc_plan_normalize(pl.get(), iL, out))) return rc;
  }
  out->nPhotons = cumPhotons;
  out->nPhotonsCollected = cumCollected;
  out->nThreads = (double)pl->blocks * pl->threads;
  out->simulationTime = cumMinutes;
  out->examplePaths = nullptr;
  out->examplePathsElems = 0;
  if (T.nExamplePaths > 0) {
    // paths of the first packets of the first wavelength
    if ((rc = select_wavelength(pl.get(), 0))) return rc;
    McLaunch lp{};
    lp.firstPacket = 0;
    lp.nPackets = T.timed ? ~0ull : std::max<uint64_t>(1, nReq / nL);
    lp.key0 = (uint32_t)in->seed;
    lp.key1 = (uint32_t)(in->seed >> 32);
    lp.stream = 0;
    int64_t cap = 2000 * T.nExamplePaths, need = 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
      DevBuf buf, cnt;
      MCM_HIP(hipMalloc(&buf.p, (size_t)cap * 4 * sizeof(double)));
      MCM_HIP(hipMalloc(&cnt.p, sizeof(int64_t)));
      hipLaunchKernelGGL(mc_example_paths, dim3(1), dim3(64), 0, pl->stream, pl->prm, lp, (int64_t)T.nExamplePaths, buf.as<double>(),
                         cap, cnt.as<int64_t>());
      MCM_HIP(hipGetLastError());
      MCM_HIP(hipMemcpyAsync(&need, cnt.p, sizeof need, hipMemcpyDeviceToHost, pl->stream));
      MCM_HIP(hipStreamSynchronize(pl->stream));
      if (need <= cap) {
        double* host = (double*)std::malloc((size_t)std::max<int64_t>(need, 1) * 4 * sizeof(double));
        if (!host) return fail(MCM_ERR_OUT_OF_MEMORY, "Error: Out of memory");
        MCM_HIP(hipMemcpy(host, buf.p, (size_t)need * 4 * sizeof(double), hipMemcpyDeviceToHost));
        out->examplePaths = host;
        out->examplePathsElems = need;
        break;
      }
      cap = need + 16;
    }
  }
  if (aborted) set_error("all packets launch outside the simulation cuboid");
  return MCM_OK;
}

}  // extern "C"
