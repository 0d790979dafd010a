// Counter-based RNG for the photon kernels: Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), written
// from the paper.  Draw k of packet p of stream s under key K is word k%4 of
// Philox(ctr = {k/4, p_lo, p_hi, s}, key = K); a word w maps to the float
// ((w >> 8) + 1) * 2^-24, uniform on (0, 1] like the reference's RandomNum
// (MCmatlab.c:130, :146).  Because a packet's stream depends only on
// (key, packet id, stream), any packet can be replayed exactly, and the packets
// a launch simulates do not depend on how lanes pick them up.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Every device function of the transport is force-inlined: an outlined call
// would pass the packet state and the launch parameters by pointer, which
// moves them into per-lane scratch memory.
#define MCM_FN __host__ __device__ inline __attribute__((always_inline))

namespace mcm {

MCM_FN void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct PacketRng {
  uint32_t key0, key1;
  uint32_t plo, phi, stream;
  uint32_t block;
  uint32_t b0, b1, b2, b3;  // unused words of the current block, next first
  int have;

  MCM_FN void start(uint64_t packet) {
    plo = (uint32_t)packet;
    phi = (uint32_t)(packet >> 32);
    block = 0;
    have = 0;
  }
  // The block's words are consumed by shifting them down rather than by
  // indexing, so the state stays in registers.
  MCM_FN uint32_t next_word() {
    if (have == 0) {
      uint32_t c[4] = {block++, plo, phi, stream};
      philox4x32_10(c, key0, key1);
      b0 = c[0];
      b1 = c[1];
      b2 = c[2];
      b3 = c[3];
      have = 4;
    }
    const uint32_t w = b0;
    b0 = b1;
    b1 = b2;
    b2 = b3;
    --have;
    return w;
  }
  // uniform float on (0, 1]
  MCM_FN float next() { return (float)((next_word() >> 8) + 1u) * (1.0f / 16777216.0f); }
  // uniform double on (0, 1] with 53 random bits (3D-source voxel choice)
  MCM_FN double next53() {
    const uint64_t hi = next_word() >> 5, lo = next_word() >> 6;  // 27 + 26 bits
    return ((double)((hi << 26) | lo) + 1.0) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace mcm
