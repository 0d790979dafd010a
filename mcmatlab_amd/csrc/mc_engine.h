// Composition of the photon physics into the per-packet state machine that
// the GPU kernel schedules across a wavefront (and the host debug harness
// runs packet by packet).  The state after each step tells the scheduler
// what the packet needs next:
//   FETCH    -> the packet died; the lane needs a new packet number
//   MOVE     -> one voxel sub-step (propagatePhoton's movement + absorption)
//   SCATTER  -> roulette + scattering (MCmatlab.c:259-260)
//   FRESNEL  -> reflection/refraction at a refractive-index change
// Tally writes go through Sink:
//   deposit(j, w)   absorbed weight into voxel j (the sink may merge
//                   consecutive deposits into one voxel), flush() at death
//   image_add(bin, w), farField(bin, w), edge(face, idx, w), collected()
//   or, when Sink::kDeferEscapes, escape(P, escaped, edgeFlux, slot): the packet
//   is recorded and those tallies are formed later (only without deposition
//   criteria, which need the collector's verdict at death)
#pragma once
#include "mc_photon.h"

namespace mcm {

enum : int { ST_FETCH = 0, ST_MOVE = 1, ST_SCATTER = 2, ST_FRESNEL = 3, ST_DONE = 4 };

template <bool FRESNEL, bool CRITERIA>
struct PacketEngine {
  Photon P;
  PacketRng R;
  uint64_t packet;
  uint32_t slot;  // index of the packet within the current launch (escape record slot)
  int state;
  bool replay;
  float replayW;
  float pendAbsorb;
  int pendJ, pendM;

  MCM_FN void init(uint32_t k0, uint32_t k1, uint32_t stream) {
    R.key0 = k0;
    R.key1 = k1;
    R.stream = stream;
    state = ST_FETCH;
    replay = false;
    replayW = 0.f;
    packet = 0;
  }

  // launchPhoton + getNewVoxelProperties for packet `id` (MCmatlab.c:247-248).
  // False when the launch fails 10^6 times (the reference aborts the run).
  MCM_FN bool start(uint64_t id, const McParams& p, const Medium* med, const uint8_t* cdf) {
    packet = id;
    R.start(id);
    if (!launch(P, p, med, R)) {
      state = ST_FETCH;
      return false;
    }
    new_voxel_properties(P, p, med, cdf);
    state = P.stepLeft > 0.f ? ST_MOVE : ST_SCATTER;
    return true;
  }

  // Same as start(), from the launch kernel's record of packet `id`.
  MCM_FN bool start_from(uint64_t id, const LaunchRecord& L, const McParams& p, const Medium* med, const uint8_t* cdf) {
    packet = id;
    R.start(id);
    if (!unpack_launch(L, P, R)) {
      state = ST_FETCH;
      return false;
    }
    set_crossing_lengths(P, p);
    new_voxel_properties(P, p, med, cdf);
    state = P.stepLeft > 0.f ? ST_MOVE : ST_SCATTER;
    return true;
  }

  template <class Sink>
  MCM_FN void death(const McParams& p, const Medium* med, const uint8_t* cdf, Sink& sink) {
    sink.flush();
    if (CRITERIA && !replay && p.evalAtEnd && criteria_met(P, p)) {
      // end-of-life deposition (MCmatlab.c:262-264): replay the packet and
      // deposit its absorbed weights scaled by the final weight
      replay = true;
      replayW = P.weight;
      start(packet, p, med, cdf);
    } else {
      replay = false;
      state = ST_FETCH;
    }
  }

  // deposit of the sub-step, then checkEscape + getNewVoxelProperties
  // (propagatePhoton :1095-1110, MCmatlab.c:254-257).  The deposit (decided
  // on the state before the escape check, as in the reference) goes to the
  // sink only after the next voxel's medium index has been requested: a
  // deposit that flushes the sink's merged value issues an atomic, and loads
  // and atomics retire in order, so an atomic issued before the load would
  // make every voxel change wait for the atomic's round trip.
  template <class Sink>
  MCM_FN void finish_substep(float absorb, const McParams& p, const Medium* med, const uint8_t* cdf, Sink& sink) {
    bool dep = false;
    double dv = 0.0;
    if (P.insideVolume && p.calcNFR) {
      if (CRITERIA && replay) {
        dep = true;
        dv = (double)(replayW * absorb);
      } else if (!(CRITERIA && p.evalAtEnd) && criteria_met(P, p)) {
        dep = true;
        dv = (double)absorb;
      }
    }
    const int jOld = P.j;
    if (!P.sameVoxel) {
      const EscapeResult e = check_escape(P, p);
      if constexpr (Sink::kDeferEscapes) {
        if ((e.escaped && (p.useLC || p.farFieldRes)) || e.edgeFlux) sink.escape(P, e.escaped, e.edgeFlux, slot);
      } else if (!replay) {
        if (e.escaped) {
          if (p.useLC) {
            const int b = collector_bin(P, p);
            if (b >= 0 && criteria_met(P, p)) {
              sink.image_add(b, (double)P.weight);
              sink.collected();
            }
          }
          if (p.farFieldRes && criteria_met(P, p)) sink.farField(far_field_bin(P, p), (double)P.weight);
        }
        if (e.edgeFlux && criteria_met(P, p)) {
          int face;
          const int idx = edge_flux_bin(P, p, &face);
          if (face >= 0) sink.edge(face, idx, (double)P.weight);
        }
      }
      if (P.alive) {
        P.j = voxel_of(p, P.i);
        const int m = p.M[P.j];  // requested before the deposit's atomic
        if (dep) sink.deposit(jOld, dv);
        dep = false;
        const Medium md = med[m];
        P.mua = md.mua;
        P.mus = md.mus;
        P.g = md.g;
        P.cdf = cdf[m];
      }
    }
    if (dep) sink.deposit(jOld, dv);
    if (!P.alive) death(p, med, cdf, sink);
    else state = P.stepLeft > 0.f ? ST_MOVE : ST_SCATTER;
  }

  template <class Sink>
  MCM_FN void do_move(const McParams& p, const Medium* med, const uint8_t* cdf, Sink& sink) {
    if (P.substeps >= kMaxSubsteps) {
      sink.runaway();
      P.alive = false;
      death(p, med, cdf, sink);
      return;
    }
    const MoveResult mr = move<FRESNEL>(P, p, med);
    if (FRESNEL && mr.fresnel) {
      pendAbsorb = mr.absorb;
      pendJ = mr.jNew;
      pendM = mr.mNew;
      state = ST_FRESNEL;
    } else {
      finish_substep(mr.absorb, p, med, cdf, sink);
    }
  }

  template <class Sink>
  MCM_FN void do_fresnel(const McParams& p, const Medium* med, const uint8_t* cdf, Sink& sink) {
    fresnel(P, p, med, pendJ, pendM, R);
    finish_substep(pendAbsorb, p, med, cdf, sink);
  }

  template <class Sink>
  MCM_FN void do_scatter(const McParams& p, const Medium* med, const uint8_t* cdf, Sink& sink) {
    roulette(P, R);
    if (!P.alive) {
      death(p, med, cdf, sink);
      return;
    }
    scatter(P, p, p.cdfs, CRITERIA ? (int)p.M[P.j] : 0, R);
    state = P.stepLeft > 0.f ? ST_MOVE : ST_SCATTER;
  }
};

// The tallies of a recorded escape (the inline code of finish_substep with
// the deposition criteria inactive): light collector image, far field and
// edge fluxes.
template <class Sink>
MCM_FN void tally_escape(const EscapeRecord& E, const McParams& p, Sink& sink) {
  Photon P;
  P.i[0] = E.a.x;
  P.i[1] = E.a.y;
  P.i[2] = E.a.z;
  P.weight = E.a.w;
  P.u[0] = E.b.x;
  P.u[1] = E.b.y;
  P.u[2] = E.b.z;
  P.time = E.b.w;
  P.ri = E.c.x;
  P.alive = false;
  P.kec = 0;
  const int flags = (int)E.c.y;
  if (flags & 1) {
    if (p.useLC) {
      const int b = collector_bin(P, p);
      if (b >= 0) {
        sink.image_add(b, (double)P.weight);
        sink.collected();
      }
    }
    if (p.farFieldRes) sink.farField(far_field_bin(P, p), (double)P.weight);
  }
  if (flags & 2) {
    int face;
    const int idx = edge_flux_bin(P, p, &face);
    if (face >= 0) sink.edge(face, idx, (double)P.weight);
  }
}

}  // namespace mcm
