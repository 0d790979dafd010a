/*
 * mc_oracle.c -- FP64 CPU restatement of MCmatlab's voxel Monte Carlo photon
 * transport.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every function names the reference function it restates.  Floating-point
 * expressions keep the reference's association order and its float-typed
 * literals (e.g. n/2.0f, 0.5f, 1.50087f) so that, with the dSFMT stream and a
 * single thread, results equal the reference build bit for bit.
 */
#define _GNU_SOURCE
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "dsfmt.h"
#include "oracle.h"

typedef double real;

#define PI_R acos(-1.0f)             /* MCmatlab.c:158 */
#define C_LIGHT ((real)29979245800)   /* cm/s, MCmatlab.c:159 */
#define THRESHOLD ((real)0.01)        /* MCmatlab.c:160 */
#define CHANCE ((real)0.1)            /* MCmatlab.c:161 */
#define KILLRANGE 5                   /* MCmatlab.c:165 */
#define CDFSIZE 201                   /* MCmatlab.c:171 */
#define INITIALPATHSSIZE 2000         /* MCmatlab.c:163 */
#define INITIALRECORDSIZE 1000        /* MCmatlab.c:164 */
#define LAUNCH_ATTEMPTS 1000000       /* MCmatlablib.c:708 */

static __thread char g_err[512];
const char *oracle_last_error(void) { return g_err; }
static int fail(int code, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

/* the reference's min/max macros (MCmatlab.c:75-76): min(a,b) = a > b ? b : a */
static inline real rmin(real a, real b) { return a > b ? b : a; }
static inline real sqr(real x) { return x * x; }

/* ------------------------------------------------------------------ RNG */

static inline uint32_t mulhilo(uint32_t a, uint32_t b, uint32_t *hi) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  return (uint32_t)p;
}

/* Philox4x32-10: 10 rounds of the Philox S-box with the published round
 * multipliers and Weyl key increments (Salmon, Moraes, Dror, Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11, section 4). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint32_t hi0, hi1;
    uint32_t lo0 = mulhilo(0xD2511F53u, c0, &hi0);
    uint32_t lo1 = mulhilo(0xCD9E8D57u, c2, &hi1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void oracle_dsfmt_stream(uint32_t seed, int n, double *out) {
  orc_dsfmt st;
  orc_dsfmt_init(&st, seed);
  for (int k = 0; k < n; k++) out[k] = orc_dsfmt_open_close(&st);
}

typedef struct rng_t {
  int kind;
  orc_dsfmt ds;
  uint32_t key[2];
  uint64_t photon;   /* Philox: packet id */
  uint32_t stream;   /* Philox: wavelength (or work-item) tag */
  uint32_t block;    /* Philox: next counter block */
  uint32_t buf[4];
  int have;
} rng_t;

/* RandomNum (MCmatlab.c:146): uniform in (0,1].  The Philox variant maps a
 * 32-bit word to ((w >> 8) + 1) * 2^-24, the same (0,1] grid the GPU uses. */
static inline real rnd(rng_t *R) {
  if (R->kind == ORC_RNG_DSFMT) return orc_dsfmt_open_close(&R->ds);
  if (R->have == 0) {
    uint32_t ctr[4] = {R->block++, (uint32_t)R->photon, (uint32_t)(R->photon >> 32), R->stream};
    oracle_philox4x32_10(ctr, R->key, R->buf);
    R->have = 4;
  }
  uint32_t w = R->buf[4 - R->have--];
  return (real)((float)((w >> 8) + 1u) * (1.0f / 16777216.0f));
}

static inline void rng_new_photon(rng_t *R, uint64_t photon) {
  R->photon = photon;
  R->block = 0;
  R->have = 0;
}

/* ------------------------------------------------------------ model state */

typedef struct geom_t {
  real d[3];
  long n[3];
  long farFieldRes;
  int boundaryType;
  real *muav, *musv, *gv, *RIv;
  unsigned char *CDFidxv;
  real *CDFs;
  unsigned char *M;
  const float *interfaceNormals;
} geom_t;

typedef struct src_t {
  int beamType;
  real emitterLength;
  real *FPIDdist1; long L_FPID1; real FPIDwidth1;
  real *FPIDdist2; long L_FPID2; real FPIDwidth2;
  real *AIDdist1; long L_AID1; real AIDwidth1;
  real *AIDdist2; long L_AID2; real AIDwidth2;
  real *S;
  real power;
  real focus[3], u[3], v[3], w[3];
} src_t;

typedef struct lc_t {
  real r[3];
  real theta, phi, f, diam, FSorNA;
  long res[2];
  real tStart, tEnd;
} lc_t;

typedef struct dc_t {
  unsigned long minS, maxS, minRefr, maxRefr, minRefl, maxRefl, minI, maxI, minIdx, maxIdx;
  bool onlyCollected, evalAtEnd;
} dc_t;

typedef struct tally_t {
  unsigned long long nPhotons, nPhotonsCollected;
  real *NFR, *image, *FF, *NI_xpos, *NI_xneg, *NI_ypos, *NI_yneg, *NI_zpos, *NI_zneg;
} tally_t;

typedef struct paths_t {
  long pathsElems, nExamplePaths, nFinished;
  bool started;
  long pathsSize;
  real *data;
} paths_t;

typedef struct photon_t {
  real i[3], u[3], D[3];
  long j;
  real mua, mus, g, RI;
  unsigned char CDFidx;
  real stepLeft, weight, time;
  bool insideVolume, alive, sameVoxel;
  long recordSize, recordElems;
  long *j_record;
  real *weight_record;
  unsigned long scatterings, refractions, reflections, interfaceTransitions;
  char kec; /* 0 killed, 1 escaped, 2 collected (killed_escaped_collected) */
  rng_t R;
  bool master;
  orc_mc_stats st;
} photon_t;

typedef struct ctx_t {
  geom_t *G;
  src_t *B;
  lc_t *LC;
  dc_t *DC;
  tally_t *O;
  paths_t *Pa;
  volatile bool *aborting;
  int threads;
} ctx_t;

static inline void atomic_add(real *p, real v) {
#ifdef _OPENMP
#pragma omp atomic
#endif
  *p += v;
}
static inline void atomic_inc(unsigned long long *p) {
#ifdef _OPENMP
#pragma omp atomic
#endif
  (*p)++;
}

/* depositionCriteriaMet, MCmatlablib.c:126-136 */
static bool criteria_met(const photon_t *P, const dc_t *DC) {
  const bool dead_killed = !P->alive && P->kec == 0;
  return P->scatterings >= DC->minS && (dead_killed ? ULONG_MAX : P->scatterings) <= DC->maxS &&
         P->refractions >= DC->minRefr && (dead_killed ? ULONG_MAX : P->refractions) <= DC->maxRefr &&
         P->reflections >= DC->minRefl && (dead_killed ? ULONG_MAX : P->reflections) <= DC->maxRefl &&
         P->interfaceTransitions >= DC->minI && (dead_killed ? ULONG_MAX : P->interfaceTransitions) <= DC->maxI &&
         (DC->onlyCollected ? P->kec == 2 : true);
}

/* infCast, MCmatlablib.c:138 */
static unsigned long inf_cast(double x) { return isinf(x) ? ULONG_MAX : (unsigned long)x; }

/* unitcrossprod, MCmatlablib.c:228-234 */
static void unit_cross(const real *a, const real *b, real *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  real norm = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int k = 0; k < 3; k++) c[k] /= norm;
}

/* xyztoXYZ, MCmatlablib.c:239-246: coordinates in the light-collector frame */
static void to_lc_frame(const real *r, real theta, real phi, real *o) {
  o[0] = sin(phi) * r[0] - cos(phi) * r[1];
  o[1] = cos(theta) * cos(phi) * r[0] + cos(theta) * sin(phi) * r[1] - sin(theta) * r[2];
  o[2] = sin(theta) * cos(phi) * r[0] + sin(theta) * sin(phi) * r[1] + cos(theta) * r[2];
}

/* axisrotate, MCmatlablib.c:251-258: rotate r by angle th about unit axis a */
static void rotate_about(const real *r, const real *a, real th, real *o) {
  const real st = sin(th), ct = cos(th);
  o[0] = (a[0] * a[0] * (1 - ct) + ct) * r[0] + (a[0] * a[1] * (1 - ct) - a[2] * st) * r[1] + (a[0] * a[2] * (1 - ct) + a[1] * st) * r[2];
  o[1] = (a[1] * a[0] * (1 - ct) + a[2] * st) * r[0] + (a[1] * a[1] * (1 - ct) + ct) * r[1] + (a[1] * a[2] * (1 - ct) - a[0] * st) * r[2];
  o[2] = (a[2] * a[0] * (1 - ct) - a[1] * st) * r[0] + (a[2] * a[1] * (1 - ct) + a[0] * st) * r[1] + (a[2] * a[2] * (1 - ct) + ct) * r[2];
}

/* binaryTreeSearch, MCmatlablib.c:263-276: index j with D[j] < x <= D[j+1]
 * in a nondecreasing table D[0..N].  The probe sequence (and hence which
 * index wins on ties) is the reference's. */
static long cdf_search(real x, long N, const real *D) {
  long span = N - 1;
  long j = span / 2;
  while (!(D[j] < x && D[j + 1] >= x)) {
    if (D[j] >= x) {
      j = j + (span - 2) / 4 - span / 2;
      span = span / 2 - 1;
    } else {
      span = span - span / 2 - 1;
      j = j + span / 2 + 1;
    }
  }
  return j;
}

/* getNewj, MCmatlablib.c:281-285 */
static inline long clamp_axis(real i, long n) { return i < 0 ? 0 : (i >= n ? n - 1 : (long)floor(i)); }
static inline long voxel_index(const geom_t *G, const real *i) {
  return clamp_axis(i[2], G->n[2]) * G->n[0] * G->n[1] + clamp_axis(i[1], G->n[1]) * G->n[0] + clamp_axis(i[0], G->n[0]);
}

/* Lambert W, lower branch W_{-1}(x) for -1/e <= x < 0 (LG01 beam radius,
 * MCmatlablib.c:586-588).  Halley's iteration (Corless et al. 1996, eq. 5.12)
 * from the asymptotic start log(-x) - log(-log(-x)); near the branch point
 * from the first terms of the branch-point series in p = -sqrt(2(1 + e x)). */
static real lambert_wm1(real x) {
  if (x >= 0) return 0;
  const real q = x + exp(-1.0);
  if (q <= 0) return -1.0;
  real w;
  if (q < 1e-3) {
    const real p = -sqrt(2.0 * exp(1.0) * q);
    w = -1.0 + p * (1.0 + p * (-1.0 / 3.0 + p * 11.0 / 72.0));
  } else {
    const real l1 = log(-x), l2 = log(-l1);
    w = l1 - l2 + l2 / l1;
  }
  for (int it = 0; it < 64; it++) {
    const real e = exp(w), p = w + 1.0;
    real t = w * e - x;
    t /= e * p - 0.5 * (p + 1.0) * t / p;
    w -= t;
    if (fabs(t) < 10 * DBL_EPSILON * fmax(fabs(w), 1.0 / (fabs(p) * e))) break;
  }
  return w;
}

double oracle_lambert_wm1(double x) { return (double)lambert_wm1((real)x); }

/* ------------------------------------------------------------ paths */

/* addToPhotonPath, MCmatlablib.c:478-502 (CPU branch) */
static void path_add(photon_t *P, paths_t *Pa, const geom_t *G, int nan_cols) {
  if (Pa->pathsElems == Pa->pathsSize - nan_cols) {
    Pa->pathsSize *= 2;
    real *nd = (real *)realloc(Pa->data, 4 * Pa->pathsSize * sizeof(real));
    if (!nd) abort();
    Pa->data = nd;
  }
  for (int k = 0; k < nan_cols; k++) {
    for (int c = 0; c < 4; c++) Pa->data[4 * Pa->pathsElems + c] = NAN;
    Pa->pathsElems++;
  }
  Pa->data[4 * Pa->pathsElems + 0] = (P->i[0] - G->n[0] / 2.0f) * G->d[0];
  Pa->data[4 * Pa->pathsElems + 1] = (P->i[1] - G->n[1] / 2.0f) * G->d[1];
  Pa->data[4 * Pa->pathsElems + 2] = (P->i[2]) * G->d[2];
  Pa->data[4 * Pa->pathsElems + 3] = P->weight;
  Pa->pathsElems++;
  Pa->started = true;
}

/* updatePaths, MCmatlablib.c:507-514 */
static void path_update(photon_t *P, ctx_t *X, bool teleported) {
  paths_t *Pa = X->Pa;
  if (!P->master || Pa->nFinished >= Pa->nExamplePaths) return;
  if (X->DC->evalAtEnd || criteria_met(P, X->DC)) {
    int nan_cols = Pa->started ? (teleported ? 1 : 0) : 2;
    path_add(P, Pa, X->G, nan_cols);
  }
}

/* deletePhotonPath, MCmatlablib.c:465-473 */
static void path_delete(paths_t *Pa) {
  for (; Pa->pathsElems > 1; Pa->pathsElems--) {
    if (isnan(Pa->data[4 * (Pa->pathsElems - 1)]) && isnan(Pa->data[4 * (Pa->pathsElems - 2)])) {
      Pa->pathsElems -= 2;
      break;
    }
  }
}

/* ------------------------------------------------------------ transport */

static inline real boundary_distance(real i, real u, real d) {
  return u ? (floor(i) + (u > 0) - i) * d / u : INFINITY;
}

/* launchPhoton, MCmatlablib.c:519-732 */
static void launch(photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  const src_t *B = X->B;
  real Xv, Yv, r, phi, tanphiX, tanphiY, costheta, sintheta;
  real target[3] = {0, 0, 0}, w0[3];
  long j;
  P->sameVoxel = false;
  P->weight = 1;
  P->recordElems = 0;
  P->kec = 0;
  long attempts = 0;
  do {
    if (B->S) {
      j = cdf_search(rnd(&P->R), G->n[0] * G->n[1] * G->n[2], B->S);
      P->i[0] = j % G->n[0] + 1 - rnd(&P->R);
      P->i[1] = j / G->n[0] % G->n[1] + 1 - rnd(&P->R);
      P->i[2] = j / G->n[0] / G->n[1] + 1 - rnd(&P->R);
      costheta = 1 - 2 * rnd(&P->R);
      sintheta = sqrt(1 - costheta * costheta);
      phi = 2 * PI_R * rnd(&P->R);
      P->u[0] = sintheta * cos(phi);
      P->u[1] = sintheta * sin(phi);
      P->u[2] = costheta;
      P->j = voxel_index(G, P->i);
      P->RI = G->RIv[G->M[P->j]];
      P->time = 0;
    } else {
      switch (B->beamType) {
        case 0: /* pencil beam */
          P->i[0] = (B->focus[0] - B->focus[2] * B->u[0] / B->u[2]) / G->d[0] + G->n[0] / 2.0f;
          P->i[1] = (B->focus[1] - B->focus[2] * B->u[1] / B->u[2]) / G->d[1] + G->n[1] / 2.0f;
          P->i[2] = 0;
          for (int k = 0; k < 3; k++) P->u[k] = B->u[k];
          P->j = voxel_index(G, P->i);
          P->RI = G->RIv[G->M[P->j]];
          P->time = -P->RI / C_LIGHT * sqrt(sqr((P->i[0] - G->n[0] / 2.0f) * G->d[0] - B->focus[0]) +
                                           sqr((P->i[1] - G->n[1] / 2.0f) * G->d[1] - B->focus[1]) +
                                           sqr((P->i[2]) * G->d[2] - B->focus[2]));
          break;
        case 1: /* isotropic point or line emitter */
          r = rnd(&P->R);
          P->i[0] = (B->focus[0] + B->u[0] * (r - 0.5) * B->emitterLength) / G->d[0] + G->n[0] / 2.0f;
          P->i[1] = (B->focus[1] + B->u[1] * (r - 0.5) * B->emitterLength) / G->d[1] + G->n[1] / 2.0f;
          P->i[2] = (B->focus[2] + B->u[2] * (r - 0.5) * B->emitterLength) / G->d[2];
          costheta = 1 - 2 * rnd(&P->R);
          sintheta = sqrt(1 - costheta * costheta);
          phi = 2 * PI_R * rnd(&P->R);
          P->u[0] = sintheta * cos(phi);
          P->u[1] = sintheta * sin(phi);
          P->u[2] = costheta;
          P->j = voxel_index(G, P->i);
          P->RI = G->RIv[G->M[P->j]];
          P->time = 0;
          break;
        case 2: { /* infinite plane wave */
          const real span = (G->boundaryType == 1 || G->boundaryType == 3) ? 1 : KILLRANGE;
          P->i[0] = span * G->n[0] * (rnd(&P->R) - 0.5f) + G->n[0] / 2.0f;
          P->i[1] = span * G->n[1] * (rnd(&P->R) - 0.5f) + G->n[1] / 2.0f;
          P->i[2] = 0;
          for (int k = 0; k < 3; k++) P->u[k] = B->u[k];
          P->j = voxel_index(G, P->i);
          P->RI = G->RIv[G->M[P->j]];
          P->time = P->RI / C_LIGHT * ((P->i[0] - G->n[0] / 2.0f) * G->d[0] * B->u[0] + (P->i[1] - G->n[1] / 2.0f) * G->d[1] * B->u[1] +
                                      (P->i[2]) * G->d[2] * B->u[2]);
          break;
        }
        case 3: /* Laguerre-Gaussian LG01 */
          phi = rnd(&P->R) * 2 * PI_R;
          rotate_about(B->v, B->u, phi, w0);
          r = B->FPIDwidth1 * sqrt(((real)lambert_wm1(-rnd(&P->R) * exp(-1.0f)) + 1) / (-2)) / 1.50087f;
          for (int k = 0; k < 3; k++) target[k] = B->focus[k] + r * w0[k];
          phi = B->AIDwidth1 * sqrt(((real)lambert_wm1(-rnd(&P->R) * exp(-1.0f)) + 1) / (-2)) / 1.50087f;
          rotate_about(B->u, w0, phi, P->u);
          goto ray_to_surface;
        case 4: /* radially factorizable beam */
          rotate_about(B->v, B->u, rnd(&P->R) * 2 * PI_R, w0);
          if (*B->FPIDdist1 == -1) {
            r = B->FPIDwidth1 * sqrt(rnd(&P->R));
          } else if (*B->FPIDdist1 == -2) {
            r = B->FPIDwidth1 * sqrt(-0.5f * log(rnd(&P->R)));
          } else {
            const real a = rnd(&P->R); /* table draw before the in-bin draw */
            const long jb = cdf_search(a, B->L_FPID1 - 1, B->FPIDdist1);
            r = B->FPIDwidth1 * (jb + rnd(&P->R)) / (B->L_FPID1 - 1);
          }
          for (int k = 0; k < 3; k++) target[k] = B->focus[k] + r * w0[k];
          rotate_about(B->v, B->u, rnd(&P->R) * 2 * PI_R, w0);
          if (*B->AIDdist1 == -1) {
            phi = atan(tan(B->AIDwidth1) * sqrt(rnd(&P->R)));
          } else if (*B->AIDdist1 == -2) {
            phi = atan(tan(B->AIDwidth1) * sqrt(-0.5f * log(rnd(&P->R))));
          } else if (*B->AIDdist1 == -3) {
            phi = asin(sqrt(rnd(&P->R)));
          } else {
            const real a = rnd(&P->R);
            const long jb = cdf_search(a, B->L_AID1 - 1, B->AIDdist1);
            phi = atan(tan(B->AIDwidth1) * (jb + rnd(&P->R)) / (B->L_AID1 - 1));
          }
          rotate_about(B->u, w0, phi, P->u);
          goto ray_to_surface;
        case 5: /* X/Y factorizable beam */
          if (*B->FPIDdist1 == -1) {
            Xv = B->FPIDwidth1 * (rnd(&P->R) * 2 - 1);
          } else if (*B->FPIDdist1 == -2) {
            const real a = rnd(&P->R), b = rnd(&P->R); /* Box-Muller: radius draw, then angle draw */
            Xv = B->FPIDwidth1 * sqrt(-0.5f * log(a)) * cos(2 * PI_R * b);
          } else {
            const real a = rnd(&P->R);
            const long jb = cdf_search(a, B->L_FPID1 - 1, B->FPIDdist1);
            Xv = B->FPIDwidth1 * ((jb + rnd(&P->R)) / (B->L_FPID1 - 1) * 2 - 1);
          }
          if (*B->FPIDdist2 == -1) {
            Yv = B->FPIDwidth2 * (rnd(&P->R) * 2 - 1);
          } else if (*B->FPIDdist2 == -2) {
            const real a = rnd(&P->R), b = rnd(&P->R);
            Yv = B->FPIDwidth2 * sqrt(-0.5f * log(a)) * cos(2 * PI_R * b);
          } else {
            const real a = rnd(&P->R);
            const long jb = cdf_search(a, B->L_FPID2 - 1, B->FPIDdist2);
            Yv = B->FPIDwidth2 * ((jb + rnd(&P->R)) / (B->L_FPID2 - 1) * 2 - 1);
          }
          for (int k = 0; k < 3; k++) target[k] = B->focus[k] + Xv * B->v[k] + Yv * B->w[k];
          if (*B->AIDdist1 == -3) {
            rotate_about(B->v, B->u, rnd(&P->R) * 2 * PI_R, w0);
            rotate_about(B->u, w0, asin(sqrt(rnd(&P->R))), P->u);
          } else {
            if (*B->AIDdist1 == -1) {
              tanphiX = tan(B->AIDwidth1) * (rnd(&P->R) * 2 - 1);
            } else if (*B->AIDdist1 == -2) {
              const real a = rnd(&P->R), b = rnd(&P->R);
              tanphiX = tan(B->AIDwidth1) * sqrt(-0.5f * log(a)) * cos(2 * PI_R * b);
            } else {
              const real a = rnd(&P->R);
              const long jb = cdf_search(a, B->L_AID1, B->AIDdist1);
              tanphiX = tan(B->AIDwidth1) * ((jb + rnd(&P->R)) / B->L_AID1 * 2 - 1);
            }
            if (*B->AIDdist2 == -1) {
              tanphiY = tan(B->AIDwidth2) * (rnd(&P->R) * 2 - 1);
            } else if (*B->AIDdist2 == -2) {
              const real a = rnd(&P->R), b = rnd(&P->R);
              tanphiY = tan(B->AIDwidth2) * sqrt(-0.5f * log(a)) * cos(2 * PI_R * b);
            } else {
              const real a = rnd(&P->R);
              const long jb = cdf_search(a, B->L_AID2, B->AIDdist2);
              tanphiY = tan(B->AIDwidth2) * ((jb + rnd(&P->R)) / B->L_AID2 * 2 - 1);
            }
            rotate_about(B->v, B->u, atan2(tanphiX, tanphiY), w0);
            rotate_about(B->u, w0, atan(sqrt(tanphiX * tanphiX + tanphiY * tanphiY)), P->u);
          }
        ray_to_surface:
          /* start where the ray through `target` crosses z = 0, at a time
           * such that it passes the focal plane at t = 0 */
          P->i[0] = (target[0] - target[2] * P->u[0] / P->u[2]) / G->d[0] + G->n[0] / 2.0f;
          P->i[1] = (target[1] - target[2] * P->u[1] / P->u[2]) / G->d[1] + G->n[1] / 2.0f;
          P->i[2] = 0;
          P->j = voxel_index(G, P->i);
          P->RI = G->RIv[G->M[P->j]];
          P->time = -P->RI / C_LIGHT * sqrt(sqr((P->i[0] - G->n[0] / 2.0f) * G->d[0] - target[0]) +
                                           sqr((P->i[1] - G->n[1] / 2.0f) * G->d[1] - target[1]) +
                                           sqr((P->i[2]) * G->d[2] - target[2]));
          break;
      }
    }
    switch (G->boundaryType) {
      case 0:
        P->alive = (fabs(P->i[0] / G->n[0] - 1.0f / 2) < KILLRANGE / 2.0f && fabs(P->i[1] / G->n[1] - 1.0f / 2) < KILLRANGE / 2.0f &&
                    fabs(P->i[2] / G->n[2] - 1.0f / 2) < KILLRANGE / 2.0f);
        break;
      case 2:
        P->alive = (fabs(P->i[0] / G->n[0] - 1.0f / 2) < KILLRANGE / 2.0f && fabs(P->i[1] / G->n[1] - 1.0f / 2) < KILLRANGE / 2.0f &&
                    P->i[2] / G->n[2] - 1.0f / 2 < KILLRANGE / 2.0f && P->i[2] >= 0);
        break;
      default: /* 1 and 3 */
        P->alive = P->i[0] < G->n[0] && P->i[0] >= 0 && P->i[1] < G->n[1] && P->i[1] >= 0 && P->i[2] < G->n[2] && P->i[2] >= 0;
        break;
    }
  } while (!P->alive && ++attempts < LAUNCH_ATTEMPTS);
  if (attempts >= LAUNCH_ATTEMPTS) *X->aborting = true;

  for (int k = 0; k < 3; k++) P->D[k] = boundary_distance(P->i[k], P->u[k], G->d[k]);
  P->insideVolume = P->i[0] < G->n[0] && P->i[0] >= 0 && P->i[1] < G->n[1] && P->i[1] >= 0 && P->i[2] < G->n[2] && P->i[2] >= 0;
  P->stepLeft = -log(rnd(&P->R));
  P->scatterings = P->refractions = P->reflections = P->interfaceTransitions = 0;
  if (P->master) {
    X->Pa->started = false;
    path_update(P, X, false);
  }
}

/* formImage, MCmatlablib.c:737-788 */
static void form_image(photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  const lc_t *LC = X->LC;
  real U[3];
  to_lc_frame(P->u, LC->theta, LC->phi, U);
  if (!(U[2] < 0)) return;
  real resc[3] = {(P->i[0] - G->n[0] / 2.0f) * G->d[0] - LC->r[0], (P->i[1] - G->n[1] / 2.0f) * G->d[1] - LC->r[1],
                  (P->i[2]) * G->d[2] - LC->r[2]};
  real Resc[3];
  to_lc_frame(resc, LC->theta, LC->phi, Resc);
  if (!(Resc[2] > 0)) return;
  real RLCP[2];
  RLCP[0] = Resc[0] - Resc[2] * U[0] / U[2];
  RLCP[1] = Resc[1] - Resc[2] * U[1] / U[2];
  real distLCP = sqrt(RLCP[0] * RLCP[0] + RLCP[1] * RLCP[1]);
  if (!(distLCP < LC->diam / 2)) return;
  if (isfinite(LC->f)) { /* objective lens */
    real RImP[2];
    RImP[0] = RLCP[0] + LC->f * U[0] / U[2];
    RImP[1] = RLCP[1] + LC->f * U[1] / U[2];
    real distImP = sqrt(RImP[0] * RImP[0] + RImP[1] * RImP[1]);
    if (distImP < LC->FSorNA / 2) {
      long Xi = (long)(LC->res[0] * (RImP[0] / LC->FSorNA + 1.0f / 2));
      long Yi = (long)(LC->res[0] * (RImP[1] / LC->FSorNA + 1.0f / 2));
      long ti = 0;
      if (LC->res[1] > 1) {
        long t = (long)(1 + (LC->res[1] - 2) * (P->time - (Resc[2] - LC->f) / U[2] * P->RI / C_LIGHT - LC->tStart) / (LC->tEnd - LC->tStart));
        ti = t < 0L ? 0L : t;
        if (ti > LC->res[1] - 1) ti = LC->res[1] - 1;
      }
      P->kec = 2;
      if (criteria_met(P, X->DC)) {
        atomic_add(&X->O->image[Xi + Yi * LC->res[0] + ti * LC->res[0] * LC->res[0]], P->weight);
        atomic_inc(&X->O->nPhotonsCollected);
      }
    }
  } else { /* fiber tip */
    real thetaLCFF = atan(-sqrt(U[0] * U[0] + U[1] * U[1]) / U[2]);
    if (thetaLCFF < asin(fmin(1.0f, LC->FSorNA))) {
      long ti = 0;
      if (LC->res[1] > 1) {
        long t = (long)(1 + (LC->res[1] - 2) * (P->time - Resc[2] / U[2] * P->RI / C_LIGHT - LC->tStart) / (LC->tEnd - LC->tStart));
        ti = t < 0L ? 0L : t;
        if (ti > LC->res[1] - 1) ti = LC->res[1] - 1;
      }
      P->kec = 2;
      if (criteria_met(P, X->DC)) {
        atomic_add(&X->O->image[ti], P->weight);
        atomic_inc(&X->O->nPhotonsCollected);
      }
    }
  }
}

/* formFarField, MCmatlablib.c:793-797 */
static void form_far_field(const photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  real theta = (1 - DBL_EPSILON) * acos(P->u[2]);
  real phi_shifted = (1 - DBL_EPSILON) * (PI_R + atan2(P->u[1], P->u[0]));
  atomic_add(&X->O->FF[(long)floor(theta / PI_R * G->farFieldRes) + G->farFieldRes * (long)floor(phi_shifted / (2 * PI_R) * G->farFieldRes)],
             P->weight);
}

/* formEdgeFluxes, MCmatlablib.c:802-816 */
static void form_edge_fluxes(const photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  tally_t *O = X->O;
  if (G->boundaryType == 1) {
    if (P->i[2] < 0) atomic_add(&O->NI_zneg[(long)P->i[0] + G->n[0] * (long)P->i[1]], P->weight);
    else if (P->i[2] >= G->n[2]) atomic_add(&O->NI_zpos[(long)P->i[0] + G->n[0] * (long)P->i[1]], P->weight);
    else if (P->i[1] < 0) atomic_add(&O->NI_yneg[(long)P->i[0] + G->n[0] * (long)P->i[2]], P->weight);
    else if (P->i[1] >= G->n[1]) atomic_add(&O->NI_ypos[(long)P->i[0] + G->n[0] * (long)P->i[2]], P->weight);
    else if (P->i[0] < 0) atomic_add(&O->NI_xneg[(long)P->i[1] + G->n[1] * (long)P->i[2]], P->weight);
    else if (P->i[0] >= G->n[0]) atomic_add(&O->NI_xpos[(long)P->i[1] + G->n[1] * (long)P->i[2]], P->weight);
  } else if (G->boundaryType == 2) {
    if (P->i[2] < 0)
      atomic_add(&O->NI_zneg[(long)(P->i[0] + G->n[0] * (KILLRANGE - 1) / 2.0f) +
                             (KILLRANGE * G->n[0]) * ((long)(P->i[1] + G->n[1] * (KILLRANGE - 1) / 2.0f))],
                 P->weight);
  } else {
    if (P->i[2] < 0) atomic_add(&O->NI_zneg[(long)P->i[0] + G->n[0] * (long)P->i[1]], P->weight);
    else if (P->i[2] >= G->n[2]) atomic_add(&O->NI_zpos[(long)P->i[0] + G->n[0] * (long)P->i[1]], P->weight);
  }
}

/* checkEscape, MCmatlablib.c:821-889 */
static void check_escape(photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  bool escaped = false;
  switch (G->boundaryType) {
    case 0:
      P->alive = (fabs(P->i[0] / G->n[0] - 1.0f / 2) < KILLRANGE / 2.0f && fabs(P->i[1] / G->n[1] - 1.0f / 2) < KILLRANGE / 2.0f &&
                  fabs(P->i[2] / G->n[2] - 1.0f / 2) < KILLRANGE / 2.0f);
      break;
    case 1:
      P->alive = P->i[0] < G->n[0] && P->i[0] >= 0 && P->i[1] < G->n[1] && P->i[1] >= 0 && P->i[2] < G->n[2] && P->i[2] >= 0;
      escaped = !P->alive && P->RI == 1;
      break;
    case 2:
      P->alive = (fabs(P->i[0] / G->n[0] - 1.0f / 2) < KILLRANGE / 2.0f && fabs(P->i[1] / G->n[1] - 1.0f / 2) < KILLRANGE / 2.0f &&
                  P->i[2] / G->n[2] - 1.0f / 2 < KILLRANGE / 2.0f && P->i[2] >= 0);
      escaped = (P->i[2] < 0) && P->RI == 1;
      break;
    case 3: {
      bool tooOld = P->time > 1000 / C_LIGHT * sqrt(sqr(G->d[0] * G->n[0]) + sqr(G->d[1] * G->n[1]));
      P->alive = P->i[2] < G->n[2] && P->i[2] >= 0 && !tooOld;
      escaped = !P->alive && P->RI == 1 && !tooOld;
      bool teleported = false;
      if (P->i[0] >= G->n[0]) { P->i[0] = 0; teleported = true; }
      if (P->i[0] < 0) { P->i[0] = G->n[0] * (1 - DBL_EPSILON); teleported = true; }
      if (P->i[1] >= G->n[1]) { P->i[1] = 0; teleported = true; }
      if (P->i[1] < 0) { P->i[1] = G->n[1] * (1 - DBL_EPSILON); teleported = true; }
      if (P->master && teleported && (X->DC->evalAtEnd || criteria_met(P, X->DC))) path_update(P, X, teleported);
      break;
    }
  }
  P->insideVolume = P->i[0] < G->n[0] && P->i[0] >= 0 && P->i[1] < G->n[1] && P->i[1] >= 0 && P->i[2] < G->n[2] && P->i[2] >= 0;
  if (escaped) {
    P->kec = 1;
    if (X->O->image) form_image(P, X);
    if (X->O->FF && criteria_met(P, X->DC)) form_far_field(P, X);
  }
  if (!P->alive && G->boundaryType && criteria_met(P, X->DC)) form_edge_fluxes(P, X);
}

/* getNewVoxelProperties, MCmatlablib.c:894-904 (the refractive index is
 * deliberately not refreshed here) */
static inline void new_voxel_properties(photon_t *P, const geom_t *G) {
  P->j = voxel_index(G, P->i);
  const unsigned char m = G->M[P->j];
  P->mua = G->muav[m];
  P->mus = G->musv[m];
  P->g = G->gv[m];
  P->CDFidx = G->CDFidxv[m];
}

/* getNormal, MCmatlablib.c:909-915 */
static void voxel_normal(const geom_t *G, long j, real *nx, real *ny, real *nz) {
  float theta = G->interfaceNormals[2 * j];
  float phi = G->interfaceNormals[2 * j + 1];
  *nx = sin(theta) * cos(phi);
  *ny = sin(theta) * sin(phi);
  *nz = cos(theta);
}

/* getInterpolatedNormal, MCmatlablib.c:920-1014.  The reference floors the
 * already-integer voxel index minus 0.5, so the interpolation cell is always
 * voxels (ic-1, ic) along each axis with weight 1/2 each; restated as is.
 * Corners outside the grid or with a different refractive index than voxel
 * jt are skipped. */
static void interpolated_normal(const geom_t *G, const photon_t *P, long jt, real *nx, real *ny, real *nz) {
  long ic[3], lo[3];
  real w[3];
  for (int a = 0; a < 3; a++) {
    ic[a] = (long)((P->i[a] < 0) ? 0 : ((P->i[a] >= G->n[a]) ? G->n[a] - 1 : P->i[a]));
    lo[a] = (long)floor(ic[a] - 0.5);
    w[a] = (real)(ic[a] - 0.5 - lo[a]);
  }
  *nx = *ny = *nz = 0;
  const real ri_t = G->RIv[G->M[jt]];
  /* corner order of the reference: x outer, then y, then z innermost */
  for (int cx = 0; cx < 2; cx++) {
    for (int cy = 0; cy < 2; cy++) {
      for (int cz = 0; cz < 2; cz++) {
        const bool okx = cx ? lo[0] < G->n[0] - 1 : lo[0] >= 0;
        const bool oky = cy ? lo[1] < G->n[1] - 1 : lo[1] >= 0;
        const bool okz = cz ? lo[2] < G->n[2] - 1 : lo[2] >= 0;
        if (!(okx && oky && okz)) continue;
        long jc = (lo[0] + cx) + (lo[1] + cy) * G->n[0] + (lo[2] + cz) * G->n[0] * G->n[1];
        if (G->RIv[G->M[jc]] != ri_t) continue;
        real vx, vy, vz;
        voxel_normal(G, jc, &vx, &vy, &vz);
        const real wx = cx ? w[0] : (1 - w[0]);
        const real wy = cy ? w[1] : (1 - w[1]);
        const real wz = cz ? w[2] : (1 - w[2]);
        *nx += wx * wy * wz * vx;
        *ny += wx * wy * wz * vy;
        *nz += wx * wy * wz * vz;
      }
    }
  }
  real norm = sqrt(sqr(*nx) + sqr(*ny) + sqr(*nz));
  *nx /= norm;
  *ny /= norm;
  *nz /= norm;
}

static inline bool in_dc_range(unsigned char m, const dc_t *DC) { return m >= DC->minIdx && m <= DC->maxIdx; }

static void record_deposit(photon_t *P, real absorb) {
  if (P->recordElems == P->recordSize) {
    P->recordSize *= 2;
    long *nj = (long *)realloc(P->j_record, P->recordSize * sizeof(long));
    real *nw = (real *)realloc(P->weight_record, P->recordSize * sizeof(real));
    if (!nj || !nw) abort();
    P->j_record = nj;
    P->weight_record = nw;
  }
  P->j_record[P->recordElems] = P->j;
  P->weight_record[P->recordElems] = absorb;
  P->recordElems++;
}

/* propagatePhoton, MCmatlablib.c:1019-1119 */
static void propagate(photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  const dc_t *DC = X->DC;
  P->sameVoxel = true;
  P->st.substeps += 1;
  real s = rmin(P->stepLeft / P->mus, rmin(P->D[0], rmin(P->D[1], P->D[2])));
  P->stepLeft = s == P->stepLeft / P->mus ? 0 : P->stepLeft - s * P->mus;
  P->time += s * P->RI / C_LIGHT;
  for (int a = 0; a < 3; a++) {
    long i_old = (long)floor(P->i[a]);
    if (s == P->D[a]) {
      P->i[a] = (P->u[a] > 0) ? i_old + 1 : i_old - DBL_EPSILON * (labs(i_old) + 1);
      P->D[a] = G->d[a] / fabs(P->u[a]);
      P->sameVoxel = false;
    } else {
      P->i[a] += s * P->u[a] / G->d[a];
      if (floor(P->i[a]) != i_old) P->i[a] = (P->u[a] > 0) ? i_old + 1 - DBL_EPSILON * (labs(i_old) + 1) : i_old;
      P->D[a] -= s;
    }
  }
  if (!P->sameVoxel) {
    P->st.voxelChanges += 1;
    long j_new = voxel_index(G, P->i);
    const unsigned char m_new = G->M[j_new];
    real mu = P->RI / G->RIv[m_new];
    if (mu != 1) {
      P->st.riCrossings += 1;
      bool reflected = false;
      real nx, ny, nz;
      interpolated_normal(G, P, j_new, &nx, &ny, &nz);
      real cos_in = nx * P->u[0] + ny * P->u[1] + nz * P->u[2];
      if (cos_in > 0) {
        real cos_out_sqr = 1 - sqr(mu) * (1 - sqr(cos_in));
        if (cos_out_sqr > 0) {
          real cos_out = sqrt(cos_out_sqr);
          real R = sqr((mu * cos_in - cos_out) / (mu * cos_in + cos_out)) / 2 + sqr((mu * cos_out - cos_in) / (mu * cos_out + cos_in)) / 2;
          reflected = rnd(&P->R) <= R;
        } else {
          reflected = true;
        }
        if (reflected) {
          P->u[0] -= 2 * nx * cos_in;
          P->u[1] -= 2 * ny * cos_in;
          P->u[2] -= 2 * nz * cos_in;
          for (int a = 0; a < 3; a++) P->D[a] = boundary_distance(P->i[a], P->u[a], G->d[a]);
          if (in_dc_range(m_new, DC)) P->reflections++;
        } else {
          real ncoeff = sqrt(cos_out_sqr) - mu * cos_in;
          P->u[0] = ncoeff * nx + mu * P->u[0];
          P->u[1] = ncoeff * ny + mu * P->u[1];
          P->u[2] = ncoeff * nz + mu * P->u[2];
          for (int a = 0; a < 3; a++) P->D[a] = boundary_distance(P->i[a], P->u[a], G->d[a]);
          P->RI = G->RIv[m_new];
          if (in_dc_range(m_new, DC)) {
            P->refractions++;
            P->interfaceTransitions++;
          }
        }
      }
    } else if (m_new != G->M[P->j] && G->RIv[G->M[P->j]] == P->RI) {
      if (in_dc_range(m_new, DC)) P->interfaceTransitions++;
    }
  }
  real absorb = -P->weight * expm1(-P->mua * s);
  P->weight -= absorb;
  if (P->insideVolume) {
    if (!DC->evalAtEnd) {
      if (X->O->NFR && criteria_met(P, DC)) atomic_add(&X->O->NFR[P->j], absorb);
    } else {
      record_deposit(P, absorb);
    }
  }
  if (P->master) path_update(P, X, false);
}

/* checkRoulette, MCmatlablib.c:1124-1133 */
static void roulette(photon_t *P) {
  if (P->weight < THRESHOLD) {
    if (rnd(&P->R) <= CHANCE) P->weight /= CHANCE;
    else P->alive = false;
  }
}

/* scatterPhoton, MCmatlablib.c:1138-1184 */
static void scatter(photon_t *P, ctx_t *X) {
  const geom_t *G = X->G;
  P->st.scatterings += 1;
  real costheta;
  if (isnan(P->g)) {
    long jTheta = cdf_search(rnd(&P->R), CDFSIZE, G->CDFs + P->CDFidx * CDFSIZE);
    costheta = cos((jTheta + rnd(&P->R)) * PI_R / (CDFSIZE - 1));
  } else {
    costheta = fabs(P->g) == 1.0 ? P->g
               : fabs(P->g) <= sqrt(DBL_EPSILON) ? 2 * rnd(&P->R) - 1
                                                 : (1 + P->g * P->g - sqr((1 - P->g * P->g) / (1 - P->g + 2 * P->g * rnd(&P->R)))) / (2 * P->g);
  }
  real sintheta = sqrt(1 - costheta * costheta);
  real phi = 2 * PI_R * rnd(&P->R);
  real cosphi = cos(phi);
  real sinphi = sin(phi);
  if (fabs(P->u[2]) < 1) {
    real ux = sintheta * (P->u[0] * P->u[2] * cosphi - P->u[1] * sinphi) / sqrt(P->u[0] * P->u[0] + P->u[1] * P->u[1]) + P->u[0] * costheta;
    real uy = sintheta * (P->u[1] * P->u[2] * cosphi + P->u[0] * sinphi) / sqrt(P->u[0] * P->u[0] + P->u[1] * P->u[1]) + P->u[1] * costheta;
    P->u[2] = -sintheta * (cosphi)*sqrt(P->u[0] * P->u[0] + P->u[1] * P->u[1]) + P->u[2] * costheta;
    P->u[1] = uy;
    P->u[0] = ux;
  } else {
    P->u[0] = sintheta * cosphi;
    P->u[1] = sintheta * sinphi;
    P->u[2] = costheta * (P->u[2] >= 0 ? 1 : -1);
  }
  for (int a = 0; a < 3; a++) P->D[a] = boundary_distance(P->i[a], P->u[a], G->d[a]);
  P->stepLeft = -log(rnd(&P->R));
  if (in_dc_range(G->M[P->j], X->DC)) P->scatterings++;
  if (P->master) path_update(P, X, false);
}

/* ------------------------------------------------------------ photon loop */

static long long now_us(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (long long)t.tv_sec * 1000000 + t.tv_nsec / 1000;
}

typedef struct loop_args {
  ctx_t *X;
  int iL, nL;
  unsigned long long nRequested;
  bool requestCollected;
  double timeBudgetUs;     /* per wavelength; INFINITY when photon-limited */
  long long tStart;
  int rngKind;
  uint64_t seed;
  unsigned long long *nextPhoton; /* Philox: global packet counter */
  const mcm_mc_input *in;
  orc_mc_stats *stats;
} loop_args;

/* threadInitAndLoop, MCmatlab.c:178-315 (CPU branch) */
static void photon_loop(loop_args *A, int thread) {
  ctx_t *X = A->X;
  photon_t Pv;
  memset(&Pv, 0, sizeof Pv);
  photon_t *P = &Pv;
  P->master = (thread == 0);
  P->recordSize = X->DC->evalAtEnd ? INITIALRECORDSIZE : 0;
  P->j_record = X->DC->evalAtEnd ? (long *)malloc(P->recordSize * sizeof(long)) : NULL;
  P->weight_record = X->DC->evalAtEnd ? (real *)malloc(P->recordSize * sizeof(real)) : NULL;
  if (P->recordSize && (!P->j_record || !P->weight_record)) abort();
  P->R.kind = A->rngKind;
  if (A->rngKind == ORC_RNG_DSFMT) {
    orc_dsfmt_init(&P->R.ds, (uint32_t)((unsigned long)A->seed + thread));
  } else {
    P->R.key[0] = (uint32_t)A->seed;
    P->R.key[1] = (uint32_t)(A->seed >> 32);
    P->R.stream = (uint32_t)A->iL;
  }
  tally_t *O = X->O;
  int pctThisWavelength = 0;
  for (;;) {
    unsigned long long count = A->requestCollected ? O->nPhotonsCollected : O->nPhotons;
    if (!(pctThisWavelength < 100 && count + (unsigned)thread < A->nRequested && !*X->aborting)) break;
    if (A->rngKind == ORC_RNG_PHILOX) {
      unsigned long long id;
#ifdef _OPENMP
#pragma omp atomic capture
#endif
      id = (*A->nextPhoton)++;
      if (!A->requestCollected && id >= A->nRequested) break;
      rng_new_photon(&P->R, id);
    }
    launch(P, X);
    if (P->alive) new_voxel_properties(P, X->G);
    if (P->alive) {
      atomic_inc(&O->nPhotons);
      P->st.launched += 1;
    }
    while (P->alive) {
      while (P->alive && P->stepLeft > 0) {
        propagate(P, X);
        if (!P->sameVoxel) {
          check_escape(P, X);
          if (P->alive) new_voxel_properties(P, X->G);
        }
      }
      if (P->alive) roulette(P);
      if (P->alive) scatter(P, X);
    }
    if (X->DC->evalAtEnd && criteria_met(P, X->DC)) {
      for (long k = 0; k < P->recordElems; k++) atomic_add(&O->NFR[P->j_record[k]], P->weight * P->weight_record[k]);
    }
    if (P->master && X->Pa->started) {
      if (X->DC->evalAtEnd && !criteria_met(P, X->DC)) path_delete(X->Pa);
      else X->Pa->nFinished++;
    }
    /* progress (MCmatlab.c:281-310): a time budget or the photon count ends the wavelength */
    int pctTime = 0;
    if (isfinite(A->timeBudgetUs)) pctTime = (int)(100.0 * (now_us() - A->tStart) / A->timeBudgetUs);
    unsigned long long cnt = A->requestCollected ? O->nPhotonsCollected : O->nPhotons;
    int pctPhotons = (int)(100.0 * cnt / A->nRequested);
    if (P->master && A->in->isInterruptPending && A->in->isInterruptPending()) *X->aborting = true;
    pctThisWavelength = pctTime > pctPhotons ? pctTime : pctPhotons;
  }
  free(P->j_record);
  free(P->weight_record);
  if (A->stats) {
#ifdef _OPENMP
#pragma omp critical(orc_stats)
#endif
    {
      A->stats->launched += P->st.launched;
      A->stats->substeps += P->st.substeps;
      A->stats->voxelChanges += P->st.voxelChanges;
      A->stats->riCrossings += P->st.riCrossings;
      A->stats->scatterings += P->st.scatterings;
    }
  }
}

/* Build the FP64 cumulative table of a beam profile (MCmatlab.c:539-575).
 * radial = radius-weighted trapezoids (sourceType 4); a single element d
 * encodes a named profile as -1 - d. */
static void build_beam_cdf(real *dst, const double *src, long L, bool radial) {
  if (L == 1) {
    dst[0] = -1 - (real)src[0];
    return;
  }
  dst[0] = 0;
  for (long k = 1; k < L; k++) dst[k] = dst[k - 1] + (radial ? k - 1 : 1) * (real)src[k - 1] + (radial ? k : 1) * (real)src[k];
  for (long k = 1; k < L; k++) dst[k] /= dst[L - 1];
}

/* normalizeDepositionAndResetO, MCmatlablib.c:1186-1250 */
static void normalize(const src_t *B, const geom_t *G, const lc_t *LC, tally_t *O, mcm_mc_output *out, long iL, double Pfraction) {
  const double V = G->d[0] * G->d[1] * G->d[2];
  const long L = G->n[0] * G->n[1] * G->n[2];
  const long L_LC = LC->res[0] * LC->res[0];
  const long L_FF = G->farFieldRes * G->farFieldRes;
  double normfactor = (double)O->nPhotons / Pfraction;
  if (B->S) normfactor /= B->power;
  else if (B->beamType == 2 && (G->boundaryType == 0 || G->boundaryType == 2)) normfactor /= KILLRANGE * KILLRANGE;
  O->nPhotons = 0;
  O->nPhotonsCollected = 0;
  if (O->NFR)
    for (long j = 0; j < L; j++) {
      out->NFR[j + (unsigned long long)iL * L] = (float)(O->NFR[j] / (V * normfactor * G->muav[G->M[j]]));
      O->NFR[j] = 0;
    }
  if (O->FF)
    for (long j = 0; j < L_FF; j++) {
      out->farField[j + iL * L_FF] = (float)(O->FF[j] / normfactor);
      O->FF[j] = 0;
    }
  if (O->image) {
    if (L_LC > 1)
      for (long j = 0; j < L_LC * LC->res[1]; j++) {
        out->image[j + iL * L_LC * LC->res[1]] = (float)(O->image[j] / (LC->FSorNA * LC->FSorNA / L_LC * normfactor));
        O->image[j] = 0;
      }
    else
      for (long j = 0; j < LC->res[1]; j++) {
        out->image[j + iL * L_LC * LC->res[1]] = (float)(O->image[j] / normfactor);
        O->image[j] = 0;
      }
  }
  const long nxy = G->n[0] * G->n[1], nxz = G->n[0] * G->n[2], nyz = G->n[1] * G->n[2];
  if (G->boundaryType == 1) {
    for (long j = 0; j < nyz; j++) {
      out->NI_xpos[j + iL * nyz] = (float)(O->NI_xpos[j] / (G->d[1] * G->d[2] * normfactor));
      out->NI_xneg[j + iL * nyz] = (float)(O->NI_xneg[j] / (G->d[1] * G->d[2] * normfactor));
      O->NI_xpos[j] = O->NI_xneg[j] = 0;
    }
    for (long j = 0; j < nxz; j++) {
      out->NI_ypos[j + iL * nxz] = (float)(O->NI_ypos[j] / (G->d[0] * G->d[2] * normfactor));
      out->NI_yneg[j + iL * nxz] = (float)(O->NI_yneg[j] / (G->d[0] * G->d[2] * normfactor));
      O->NI_ypos[j] = O->NI_yneg[j] = 0;
    }
    for (long j = 0; j < nxy; j++) {
      out->NI_zpos[j + iL * nxy] = (float)(O->NI_zpos[j] / (G->d[0] * G->d[1] * normfactor));
      out->NI_zneg[j + iL * nxy] = (float)(O->NI_zneg[j] / (G->d[0] * G->d[1] * normfactor));
      O->NI_zpos[j] = O->NI_zneg[j] = 0;
    }
  } else if (G->boundaryType == 2) {
    for (long j = 0; j < KILLRANGE * G->n[0] * KILLRANGE * G->n[1]; j++) {
      out->NI_zneg[j + iL * nxy * KILLRANGE * KILLRANGE] = (float)(O->NI_zneg[j] / (G->d[0] * G->d[1] * normfactor));
      O->NI_zneg[j] = 0;
    }
  } else if (G->boundaryType == 3) {
    for (long j = 0; j < nxy; j++) {
      out->NI_zpos[j + iL * nxy] = (float)(O->NI_zpos[j] / (G->d[0] * G->d[1] * normfactor));
      out->NI_zneg[j + iL * nxy] = (float)(O->NI_zneg[j] / (G->d[0] * G->d[1] * normfactor));
      O->NI_zpos[j] = O->NI_zneg[j] = 0;
    }
  }
}

/* mexFunction body, MCmatlab.c:317-791 (CPU branch), with fields taken from
 * mcm_mc_input instead of mx objects. */
int oracle_mc_run(const mcm_mc_input *in, mcm_mc_output *out, int nthreads, int rngKind, uint64_t seed, orc_mc_stats *stats) {
  if (!in || !out) return fail(MCM_ERR_INVALID_ARGUMENT, "null argument");
  if (in->nx <= 0 || in->ny <= 0 || in->nz <= 0 || in->nM <= 0 || in->nL <= 0 || !in->M)
    return fail(MCM_ERR_INVALID_ARGUMENT, "bad grid or media sizes");
  if (nthreads < 1) nthreads = 1;
  const float *S_PDF = in->sourceDistribution;
  const bool timed = isnan(in->nPhotonsRequested);
  const bool requestCollected = in->requestCollectedPhotons != 0;
  volatile bool aborting = false;
  const double timeRequested = timed ? in->simulationTimeRequested : INFINITY;
  const unsigned long long nRequested = timed ? ULLONG_MAX : (unsigned long long)in->nPhotonsRequested;
  const int nL = (int)in->nL, nM = (int)in->nM;
  double nPhotonsCum = 0, nCollectedCum = 0, simTimeCum = 0;
  const int sourceType = S_PDF ? -1 : (int)in->sourceType;
  const long L_FPID1 = sourceType >= 4 ? (sourceType == 4 ? in->FPID.n_radialDistr : in->FPID.n_XDistr) : 0;
  const long L_AID1 = sourceType >= 4 ? (sourceType == 4 ? in->AID.n_radialDistr : in->AID.n_XDistr) : 0;
  const long L_FPID2 = sourceType == 5 ? in->FPID.n_YDistr : 0;
  const long L_AID2 = sourceType == 5 ? in->AID.n_YDistr : 0;

  geom_t G;
  memset(&G, 0, sizeof G);
  const long L = in->nx * in->ny * in->nz;
  G.d[0] = in->dx; G.d[1] = in->dy; G.d[2] = in->dz;
  G.n[0] = in->nx; G.n[1] = in->ny; G.n[2] = in->nz;
  G.farFieldRes = in->farFieldRes;
  G.boundaryType = in->boundaryType;
  G.M = (unsigned char *)malloc(L);
  G.interfaceNormals = in->interfaceNormals;
  for (long k = 0; k < L; k++) G.M[k] = (unsigned char)(in->M[k] - 1);
  const long nCDF = in->nCDFs;
  real *small = (real *)malloc((nM * 4 + nCDF + L_FPID1 + L_AID1 + L_FPID2 + L_AID2) * sizeof(real) + nM + 8);
  if (!G.M || !small) return fail(MCM_ERR_OUT_OF_MEMORY, "Error: Out of memory");
  G.muav = small;
  G.musv = G.muav + nM;
  G.gv = G.musv + nM;
  G.RIv = G.gv + nM;
  G.CDFs = G.RIv + nM;
  for (long k = 0; k < nCDF; k++) G.CDFs[k] = (real)in->CDFs[k];

  dc_t DC = {inf_cast(in->minScatterings), inf_cast(in->maxScatterings), inf_cast(in->minRefractions), inf_cast(in->maxRefractions),
             inf_cast(in->minReflections), inf_cast(in->maxReflections), inf_cast(in->minInterfaceTransitions),
             inf_cast(in->maxInterfaceTransitions), inf_cast(in->minSubmediaIdx) - 1, inf_cast(in->maxSubmediaIdx) - 1,
             in->onlyCollected != 0, in->evaluateOnlyAtEndOfLife != 0};
  if (DC.minS == 0 && DC.maxS == ULONG_MAX && DC.minRefr == 0 && DC.maxRefr == ULONG_MAX && DC.minRefl == 0 && DC.maxRefl == ULONG_MAX &&
      DC.minI == 0 && DC.maxI == ULONG_MAX && !DC.onlyCollected)
    DC.evalAtEnd = false;

  const bool useLC = in->useLightCollector != 0;
  const real f = (real)in->LC_f;
  const long nTimeBins = S_PDF ? 0 : (long)in->LC_nTimeBins;
  lc_t LC = {{(real)in->LC_x - (isfinite(f) ? f * sin((real)in->LC_theta) * cos((real)in->LC_phi) : 0),
              (real)in->LC_y - (isfinite(f) ? f * sin((real)in->LC_theta) * sin((real)in->LC_phi) : 0),
              (real)in->LC_z - (isfinite(f) ? f * cos((real)in->LC_theta) : 0)},
             (real)in->LC_theta, (real)in->LC_phi, f, (real)in->LC_diam,
             (real)(isfinite(f) ? in->LC_fieldSize : in->LC_NA),
             {(long)in->LC_res, nTimeBins ? nTimeBins + 2 : 1},
             (real)(S_PDF ? 0 : in->LC_tStart), (real)(S_PDF ? 0 : in->LC_tEnd)};

  paths_t Pa;
  memset(&Pa, 0, sizeof Pa);
  Pa.nExamplePaths = in->nExamplePaths;
  Pa.pathsSize = INITIALPATHSSIZE * Pa.nExamplePaths;
  Pa.data = Pa.nExamplePaths ? (real *)malloc(4 * Pa.pathsSize * sizeof(real)) : NULL;

  const bool calcNFR = in->calcNFR != 0;
  const long nxy = G.n[0] * G.n[1], nxz = G.n[0] * G.n[2], nyz = G.n[1] * G.n[2];
  tally_t O;
  memset(&O, 0, sizeof O);
  O.NFR = calcNFR ? (real *)calloc(L, sizeof(real)) : NULL;
  O.image = useLC ? (real *)calloc(LC.res[0] * LC.res[0] * LC.res[1], sizeof(real)) : NULL;
  O.FF = G.farFieldRes ? (real *)calloc(G.farFieldRes * G.farFieldRes, sizeof(real)) : NULL;
  if (G.boundaryType == 1) {
    O.NI_xpos = (real *)calloc(nyz, sizeof(real));
    O.NI_xneg = (real *)calloc(nyz, sizeof(real));
    O.NI_ypos = (real *)calloc(nxz, sizeof(real));
    O.NI_yneg = (real *)calloc(nxz, sizeof(real));
  }
  if (G.boundaryType == 1 || G.boundaryType == 3) O.NI_zpos = (real *)calloc(nxy, sizeof(real));
  if (G.boundaryType != 0) O.NI_zneg = (real *)calloc(nxy * (G.boundaryType == 2 ? KILLRANGE * KILLRANGE : 1), sizeof(real));

  real *Scdf = NULL;
  if (S_PDF) {
    Scdf = (real *)malloc((L + 1) * sizeof(real));
    if (!Scdf) return fail(MCM_ERR_OUT_OF_MEMORY, "Error: Out of memory");
    Scdf[0] = 0;
  }
  real *FPIDdist1 = G.CDFs + nCDF;
  real *AIDdist1 = FPIDdist1 + L_FPID1;
  real *FPIDdist2 = AIDdist1 + L_AID1;
  real *AIDdist2 = FPIDdist2 + L_FPID2;
  if (sourceType >= 4) {
    build_beam_cdf(FPIDdist1, sourceType == 4 ? in->FPID.radialDistr : in->FPID.XDistr, L_FPID1, sourceType == 4);
    build_beam_cdf(AIDdist1, sourceType == 4 ? in->AID.radialDistr : in->AID.XDistr, L_AID1, sourceType == 4);
    if (sourceType == 5) {
      build_beam_cdf(FPIDdist2, in->FPID.YDistr, L_FPID2, false);
      build_beam_cdf(AIDdist2, in->AID.YDistr, L_AID2, false);
    }
  }
  G.CDFidxv = (unsigned char *)(AIDdist2 + L_AID2);

  const real tb = (real)(Scdf ? 0 : in->LS_theta), pb = (real)(Scdf ? 0 : in->LS_phi), psib = (real)(Scdf ? 0 : in->LS_psi);
  real u[3] = {sin(tb) * cos(pb), sin(tb) * sin(pb), cos(tb)};
  real w[3] = {1, 0, 0};
  real v[3] = {0, 0, 1};
  if (u[2] != 1) unit_cross(u, v, w);
  rotate_about(w, u, psib, v);
  unit_cross(u, v, w);

  src_t B;
  memset(&B, 0, sizeof B);
  if (!Scdf) {
    B.beamType = sourceType;
    B.emitterLength = (real)in->emitterLength;
    B.FPIDdist1 = FPIDdist1; B.L_FPID1 = L_FPID1;
    B.FPIDwidth1 = (real)(sourceType == 5 ? in->FPID.XWidth : in->FPID.radialWidth);
    B.FPIDdist2 = FPIDdist2; B.L_FPID2 = L_FPID2;
    B.FPIDwidth2 = (real)(sourceType == 5 ? in->FPID.YWidth : 0);
    B.AIDdist1 = AIDdist1; B.L_AID1 = L_AID1;
    B.AIDwidth1 = (real)(sourceType == 5 ? in->AID.XWidth : in->AID.radialWidth);
    B.AIDdist2 = AIDdist2; B.L_AID2 = L_AID2;
    B.AIDwidth2 = (real)(sourceType == 5 ? in->AID.YWidth : 0);
    B.focus[0] = (real)in->xFocus; B.focus[1] = (real)in->yFocus; B.focus[2] = (real)in->zFocus;
  }
  B.S = Scdf;
  for (int k = 0; k < 3; k++) { B.u[k] = u[k]; B.v[k] = v[k]; B.w[k] = w[k]; }

  ctx_t X = {&G, &B, &LC, &DC, &O, &Pa, &aborting, nthreads};
  for (int iL = 0; iL < nL && !aborting; iL++) {
    for (int m = 0; m < nM; m++) {
      G.muav[m] = (real)in->mua[m + iL * nM];
      G.musv[m] = (real)in->mus[m + iL * nM];
      G.gv[m] = (real)in->g[m + iL * nM];
      G.RIv[m] = (real)in->n[m + iL * nM];
      G.CDFidxv[m] = (unsigned char)in->CDFidx[m + iL * nM];
    }
    if (Scdf) {
      for (long k = 1; k < L + 1; k++) Scdf[k] = Scdf[k - 1] + (real)S_PDF[iL * L + k - 1];
      B.power = (real)(Scdf[L] * G.d[0] * G.d[1] * G.d[2]);
      for (long k = 1; k < L + 1; k++) Scdf[k] /= Scdf[L];
    } else {
      B.power = (real)in->spectrum[iL];
    }
    const unsigned long long nReqL = timed ? ULLONG_MAX
                                           : (unsigned long long)(iL == nL - 1 ? nRequested - (requestCollected ? nCollectedCum : nPhotonsCum)
                                                                               : nRequested / nL);
    const double timeL = timeRequested / nL;
    unsigned long long nextPhoton = 0;
    loop_args A = {&X, iL, nL, nReqL, requestCollected, timed ? timeL * 60000000.0 : INFINITY, now_us(), rngKind, seed, &nextPhoton, in, stats};
    const long long t0 = now_us();
    if (nthreads == 1) {
      photon_loop(&A, 0);
    } else {
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
      photon_loop(&A, omp_get_thread_num());
#else
      photon_loop(&A, 0);
#endif
    }
    const double nPh = (double)O.nPhotons;
    nPhotonsCum += nPh;
    nCollectedCum += (double)O.nPhotonsCollected;
    simTimeCum += (now_us() - t0) / 60000000.0;
    normalize(&B, &G, &LC, &O, out, iL, B.power);
  }
  out->nPhotons = nPhotonsCum;
  out->nPhotonsCollected = nCollectedCum;
  out->nThreads = nthreads;
  out->simulationTime = simTimeCum;
  out->examplePaths = NULL;
  out->examplePathsElems = 0;
  if (Pa.nExamplePaths) {
    out->examplePaths = Pa.data;
    out->examplePathsElems = Pa.pathsElems;
  }
  free(Scdf);
  free(small);
  free(G.M);
  free(O.NFR); free(O.image); free(O.FF);
  free(O.NI_xpos); free(O.NI_xneg); free(O.NI_ypos); free(O.NI_yneg); free(O.NI_zpos); free(O.NI_zneg);
  return MCM_OK;
}
