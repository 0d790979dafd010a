/*
 * dsfmt.h -- project restatement of the dSFMT-19937 generator (Saito &
 * Matsumoto, "SIMD-oriented Fast Mersenne Twister", double precision variant,
 * v2.2), the RNG of the reference's CPU path.
 *
 * TEST INFRASTRUCTURE ONLY (oracle).  The reference draws with
 * dsfmt_genrand_open_close() = 2 - [1,2) (dSFMT.h:327-329), seeded per thread
 * with dsfmt_init_gen_rand(seed) (MCmatlab.c:241).  Only those two operations
 * are restated here.  Pinned by tests/test_oracle_rng.py against the
 * generator's published known-answer file (dSFMT.19937.out.txt, committed as
 * tests/golden/dsfmt19937_kat.json).
 */
#ifndef MCM_ORACLE_DSFMT_H
#define MCM_ORACLE_DSFMT_H
#include <stdint.h>

#define ORC_DSFMT_N 191              /* (19937 - 128) / 104 + 1 128-bit words */
#define ORC_DSFMT_N64 (2 * ORC_DSFMT_N)

typedef struct orc_dsfmt {
  uint64_t s[2 * (ORC_DSFMT_N + 1)]; /* N words + the "lung" word, 2 x 64 bits each */
  int idx;                           /* next 64-bit word to hand out */
} orc_dsfmt;

void orc_dsfmt_init(orc_dsfmt *st, uint32_t seed);
/* uniform double in [1,2) */
double orc_dsfmt_close1_open2(orc_dsfmt *st);
/* uniform double in (0,1], the reference's RandomNum */
static inline double orc_dsfmt_open_close(orc_dsfmt *st) { return 2.0 - orc_dsfmt_close1_open2(st); }

#endif
