/*
 * heat_oracle.c -- CPU restatement of MCmatlab's heat propagator
 * (+MCmatlab/src/finiteElementHeatPropagator.c).  TEST INFRASTRUCTURE ONLY.
 *
 * One step of the reference's Douglas-Gunn ADI scheme (:767-782), with
 * L_d T_i = sum over the two neighbours n along axis d of c[M_i,M_n,d](T_n - T_i)
 * (c = P.dTdtperdeltaT, index M_i + nM*M_n + nM*nM*d) and Q = dTdt_abs:
 *   1a  T* = T + dt (Q [lightsOn] + L_x T / 2 + L_y T + L_z T)        (:288-312)
 *   1b  (I - dt/2 L_x) T' = T*      Thomas solve along x             (:352-389)
 *   2a  T' -= dt/2 L_y T                                             (:429-448)
 *   2b  (I - dt/2 L_y) T'' = T'     Thomas solve along y             (:486-521)
 *   3a  T'' -= dt/2 L_z T                                            (:546-565)
 *   3b  (I - dt/2 L_z) T''' = T''   Thomas solve along z, then       (:606-646)
 *       Omega += dt exp(log A - E / (R ((T''' + T)/2 + 273.15)))  where E != 0
 *       and the per-medium running maximum of T'''.
 * With heatBoundaryType 1 the outermost voxel layer is held fixed.  Sensors are
 * trilinear interpolations of the new T after each step (:652-663).
 *
 * Variants: FP32 with the reference's float expressions (bit-identical to the
 * reference build single-threaded), the same in FP64, and FP64 forward Euler
 * (T_new = T + dt (Q + L_x T + L_y T + L_z T), stable for dt <= dtmax).
 */
#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define R_GAS_F 8.3144598f      /* finiteElementHeatPropagator.c:86 */
#define CELSIUS_ZERO_F 273.15f   /* :87 */

typedef struct heat_ctx {
  long nx, ny, nz, nM;
  const unsigned char *M;
  bool sinked, lightsOn, calcDamage;
  int nthreads;
} heat_ctx;

#ifdef _OPENMP
#include <omp.h>
static inline int omp_tid(void) { return omp_get_thread_num(); }
#else
static inline int omp_tid(void) { return 0; }
#endif

static inline bool frozen(const heat_ctx *H, long ix, long iy, long iz) {
  return H->sinked && (ix == 0 || ix == H->nx - 1 || iy == 0 || iy == H->ny - 1 || iz == 0 || iz == H->nz - 1);
}

/* The three generic bodies below are instantiated for float and double. */
#define RT float
#define SFX f32
#define EXPF expf
#include "heat_oracle_impl.h"
#undef RT
#undef SFX
#undef EXPF
#define RT double
#define SFX f64
#define EXPF exp
#include "heat_oracle_impl.h"
#undef RT
#undef SFX
#undef EXPF

int oracle_heat_run(const mcm_heat_input *in, mcm_heat_output *out, int variant, int nthreads) {
  if (!in || !out || !in->T || !in->M || !in->dTdtperdeltaT || !in->dTdt_abs) return MCM_ERR_INVALID_ARGUMENT;
  heat_ctx H = {in->nx, in->ny, in->nz, in->nM, in->M, in->heatBoundaryType != 0, in->lightsOn != 0, in->Omega != NULL,
                nthreads > 1 ? nthreads : 1};
#ifndef _OPENMP
  H.nthreads = 1;
#endif
  if (variant == ORC_HEAT_FP32_DGADI) return heat_run_f32(&H, in, out, 0);
  if (variant == ORC_HEAT_FP64_DGADI) return heat_run_f64(&H, in, out, 0);
  if (variant == ORC_HEAT_FP64_EXPLICIT) return heat_run_f64(&H, in, out, 1);
  return MCM_ERR_INVALID_ARGUMENT;
}
