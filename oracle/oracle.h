/*
 * oracle.h -- CPU restatement of the reference's two hot paths.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle; the product library never
 * links or calls it.
 *
 *   oracle_mc_run    follows +MCmatlab/src/MCmatlab.c:178-315 (photon loop),
 *                    :317-791 (MEX body) and MCmatlablib.c (transport), in
 *                    FP64 with the reference's operation order.  With
 *                    rng = ORC_RNG_DSFMT and nthreads = 1 it reproduces the
 *                    reference compiled from source bit for bit for the same
 *                    seed (pinned in tests/test_oracle_mc.py).
 *   oracle_heat_run  follows +MCmatlab/src/finiteElementHeatPropagator.c:
 *                    FP32 Douglas-Gunn ADI with the reference's operation
 *                    order (pinned bit for bit against the reference build),
 *                    the same scheme in FP64, and an FP64 explicit-Euler
 *                    variant (pinned by closed forms in tests/test_oracle_heat.py).
 *   oracle_downsample_params follows +MCmatlab/src/getDownsampledParamVals.c
 *                    (farthest-point media binning), single-threaded build
 *                    (pinned bit for bit against the reference build and by
 *                    brute-force cases in tests/test_oracle_binning.py).
 */
#ifndef MCM_ORACLE_H
#define MCM_ORACLE_H
#include <stdint.h>
#include "mcm_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_RNG_DSFMT 0   /* one dSFMT stream per thread, seeded seed+thread (reference CPU path) */
#define ORC_RNG_PHILOX 1  /* the GPU's counter-based stream: draw k of photon p = Philox4x32-10 */

/* Per-run event counters used for the roofline's algorithmic-byte model. */
typedef struct orc_mc_stats {
  double launched;       /* photons launched alive */
  double substeps;       /* propagatePhoton calls */
  double voxelChanges;   /* substeps that left the voxel */
  double riCrossings;    /* voxel changes into a different refractive index (Fresnel evaluations) */
  double scatterings;    /* scatterPhoton calls */
} orc_mc_stats;

int oracle_mc_run(const mcm_mc_input *in, mcm_mc_output *out, int nthreads, int rng,
                  uint64_t seed, orc_mc_stats *stats);

#define ORC_HEAT_FP32_DGADI 0
#define ORC_HEAT_FP64_DGADI 1
#define ORC_HEAT_FP64_EXPLICIT 2
int oracle_heat_run(const mcm_heat_input *in, mcm_heat_output *out, int variant, int nthreads);

/* [out, binidx] = getDownsampledParamVals(pv, N2): pv is 3 x N1 (mua, mus, g)
 * column-major single, out 3 x N2, binidx N1 (1-based).  0, or -1 / -2 on bad
 * arguments / allocation failure. */
int oracle_downsample_params(const float *pv, int64_t N1, int64_t N2, float *out, uint8_t *binidx);

/* Philox4x32-10 (Salmon et al., SC'11), exposed for the RNG known-answer test. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* dSFMT (0,1] stream, exposed for the known-answer test. */
void oracle_dsfmt_stream(uint32_t seed, int n, double *out);
const char *oracle_last_error(void);
/* The Lambert W lower branch the LG01 source samples with, exposed for its
 * known-answer test against the reference's lambert.c. */
double oracle_lambert_wm1(double x);

#ifdef __cplusplus
}
#endif
#endif
