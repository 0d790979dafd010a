/*
 * dsfmt.c -- project restatement of dSFMT-19937 (see dsfmt.h).
 *
 * Generator constants (the published parameter set "dSFMT2-19937:117-19:
 * ffafffffffb3f-ffdfffc90fffd"): pick-up position 117, left shift 19, right
 * shift 12, the two 52-bit masks, and the two fix / parity-check words used
 * by the period certification.  They are the generator's definition, so they
 * are necessarily the same numbers as in the reference's vendored copy
 * (dSFMT-params19937.h:6-17).
 */
#include "dsfmt.h"

#define POS1 117
#define SL1 19
#define SR 12
#define MSK1 UINT64_C(0x000ffafffffffb3f)
#define MSK2 UINT64_C(0x000ffdfffc90fffd)
#define FIX1 UINT64_C(0x90014964b32f4329)
#define FIX2 UINT64_C(0x3b8d12ac548a7c7a)
#define PCV1 UINT64_C(0x3d84e1ac0dc82880)
#define PCV2 UINT64_C(0x0000000000000001)
#define MANTISSA_MASK UINT64_C(0x000fffffffffffff)
#define EXPONENT_ONE UINT64_C(0x3ff0000000000000)

/* One step of the recursion on 128-bit word w[k] (two uint64 halves).
 * a = old word k, b = word k+POS1 (mod N, already updated when wrapped),
 * lung = running feedback word, updated in place. */
static void recurse(uint64_t *out, const uint64_t *a, const uint64_t *b, uint64_t *lung) {
  const uint64_t a0 = a[0], a1 = a[1];
  const uint64_t l0 = lung[0], l1 = lung[1];
  /* the feedback word with its two 64-bit halves exchanged and each half's
   * 32-bit words exchanged (a 32-bit word reversal of the 128-bit value) */
  const uint64_t y0 = (a0 << SL1) ^ ((l1 >> 32) | (l1 << 32)) ^ b[0];
  const uint64_t y1 = (a1 << SL1) ^ ((l0 >> 32) | (l0 << 32)) ^ b[1];
  lung[0] = y0;
  lung[1] = y1;
  out[0] = a0 ^ (y0 >> SR) ^ (y0 & MSK1);
  out[1] = a1 ^ (y1 >> SR) ^ (y1 & MSK2);
}

static void refill(orc_dsfmt *st) {
  uint64_t lung[2] = {st->s[2 * ORC_DSFMT_N], st->s[2 * ORC_DSFMT_N + 1]};
  for (int k = 0; k < ORC_DSFMT_N; k++) {
    int kb = k + POS1;
    if (kb >= ORC_DSFMT_N) kb -= ORC_DSFMT_N;
    uint64_t w[2];
    recurse(w, &st->s[2 * k], &st->s[2 * kb], lung);
    st->s[2 * k] = w[0];
    st->s[2 * k + 1] = w[1];
  }
  st->s[2 * ORC_DSFMT_N] = lung[0];
  st->s[2 * ORC_DSFMT_N + 1] = lung[1];
}

void orc_dsfmt_init(orc_dsfmt *st, uint32_t seed) {
  /* Knuth-style 32-bit linear seeding over the whole state viewed as
   * (N+1)*4 little-endian 32-bit words. */
  enum { NW = (ORC_DSFMT_N + 1) * 4 };
  uint32_t w = seed;
  uint32_t prev = seed;
  st->s[0] = seed;
  for (int i = 1; i < NW; i++) {
    w = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
    prev = w;
    int q = i >> 1;
    if (i & 1)
      st->s[q] |= (uint64_t)w << 32;
    else
      st->s[q] = w;
  }
  /* force every generator word (not the lung) into [1,2) as an IEEE double */
  for (int k = 0; k < 2 * ORC_DSFMT_N; k++) st->s[k] = (st->s[k] & MANTISSA_MASK) | EXPONENT_ONE;
  /* period certification on the lung: parity of (lung ^ fix) & pcv must be 1 */
  uint64_t t = ((st->s[2 * ORC_DSFMT_N] ^ FIX1) & PCV1) ^ ((st->s[2 * ORC_DSFMT_N + 1] ^ FIX2) & PCV2);
  t ^= t >> 32;
  t ^= t >> 16;
  t ^= t >> 8;
  t ^= t >> 4;
  t ^= t >> 2;
  t ^= t >> 1;
  if ((t & 1) == 0) st->s[2 * ORC_DSFMT_N + 1] ^= 1; /* PCV2's lowest set bit is bit 0 */
  st->idx = ORC_DSFMT_N64;
}

double orc_dsfmt_close1_open2(orc_dsfmt *st) {
  if (st->idx >= ORC_DSFMT_N64) {
    refill(st);
    st->idx = 0;
  }
  union { uint64_t u; double d; } c;
  c.u = st->s[st->idx++];
  return c.d;
}
