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

namespace mcm {

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
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
  uint32_t buf[4];
  int have;

  __host__ __device__ void start(uint64_t packet) {
    plo = (uint32_t)packet;
    phi = (uint32_t)(packet >> 32);
    block = 0;
    have = 0;
  }
  __host__ __device__ uint32_t next_word() {
    if (have == 0) {
      buf[0] = block++;
      buf[1] = plo;
      buf[2] = phi;
      buf[3] = stream;
      philox4x32_10(buf, key0, key1);
      have = 4;
    }
    const int k = 4 - have;
    --have;
    // select without dynamic register indexing
    uint32_t w = buf[0];
    w = k == 1 ? buf[1] : w;
    w = k == 2 ? buf[2] : w;
    w = k == 3 ? buf[3] : w;
    return w;
  }
  // uniform float on (0, 1]
  __host__ __device__ float next() { return (float)((next_word() >> 8) + 1u) * (1.0f / 16777216.0f); }
  // uniform double on (0, 1] with 53 random bits (3D-source voxel choice)
  __host__ __device__ double next53() {
    const uint64_t hi = next_word() >> 5, lo = next_word() >> 6;  // 27 + 26 bits
    return ((double)((hi << 26) | lo) + 1.0) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace mcm
