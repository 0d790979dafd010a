/*
 * mcm_abi.h -- C ABI of libmcm, the MI355X (gfx950) implementation of
 * MCmatlab's two native hot paths:
 *
 *   - voxel Monte Carlo photon transport, replacing the MEX entry
 *       model = MCmatlab(model, simType)
 *     (+MCmatlab/src/MCmatlab.c:317 `mexFunction`, field reads :322-615,
 *      field writes :461-506 and :755-776);
 *   - heat diffusion + Arrhenius damage, replacing the MEX entry
 *       [T, Omega, sensorTemps, maxMediaTemps] = finiteElementHeatPropagator(T, Omega, P)
 *     (+MCmatlab/src/finiteElementHeatPropagator.c:688 `mexFunction`,
 *      field reads :694-736).
 *
 * Conventions
 *   - Plain C: POD structs, raw pointers, explicit sizes.  No torch or HIP
 *     types in signatures; HIP streams are passed as `void *` (NULL = the
 *     library's own stream).
 *   - Every entry point returns an int status (MCM_OK == 0).  Nothing longjmps
 *     and nothing spins on a device error (unlike the reference's gpuAssert,
 *     MCmatlablib.c:165-171).  mcm_last_error() gives the message for the
 *     calling thread.
 *   - Arrays are column major exactly as MATLAB hands them to the MEX files
 *     (x fastest, then y, then z, then wavelength).
 *   - Input structs mirror the model fields the MEX files read, one member per
 *     field, in the reference's read order; the comment on each member names
 *     the MATLAB field and the reference line that reads it.
 *   - Output buffers are owned by the caller and sized as mcm_mc_sizes()
 *     reports, except examplePaths, which the library allocates (its length is
 *     only known after the run) and the caller releases with mcm_free().
 */
#ifndef MCM_ABI_H
#define MCM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCM_ABI_VERSION 1

/* ---- status codes ---- */
#define MCM_OK 0
#define MCM_ERR_INVALID_ARGUMENT (-1)   /* malformed input (sizes, indices, NULLs) */
#define MCM_ERR_OUT_OF_MEMORY (-2)      /* host or device allocation failed ("MCmatlab:OutOfMemory") */
#define MCM_ERR_DEVICE (-3)             /* a HIP runtime call failed */
#define MCM_ERR_NO_DEVICE (-4)          /* no usable gfx950 device */
#define MCM_ERR_UNSUPPORTED (-5)        /* valid input this build does not implement */

/* Heat propagator numerics (mcm_heat_input.precision / .scheme) */
#define MCM_HEAT_FP64 0           /* FP64 state on the device (default) */
#define MCM_HEAT_FP32 1           /* FP32 state, operation order of the reference's FP32 code */
#define MCM_HEAT_DGADI 0          /* Douglas-Gunn ADI, finiteElementHeatPropagator.c:767-782 (default) */
#define MCM_HEAT_EXPLICIT 1       /* forward Euler 7-point stencil; stable only for dt <= dtmax */

/* One MCmatlab.sourceIntensityDistribution object (LS.FPID or LS.AID). */
typedef struct mcm_beam_distr {
  const double *radialDistr; int64_t n_radialDistr; /* radialDistr, MCmatlab.c:367-368, 539, 548 */
  double radialWidth;                               /* radialWidth, MCmatlab.c:596, 602 */
  const double *XDistr; int64_t n_XDistr;           /* XDistr, MCmatlab.c:367-368 */
  double XWidth;                                    /* XWidth, MCmatlab.c:596, 602 */
  const double *YDistr; int64_t n_YDistr;           /* YDistr, MCmatlab.c:369-370, 558, 567 */
  double YWidth;                                    /* YWidth, MCmatlab.c:599, 605 */
} mcm_beam_distr;

/* Everything MCmatlab.c's mexFunction reads from model.MC (or model.FMC),
 * model.MC.LS/LC/depositionCriteria and model.G. */
typedef struct mcm_mc_input {
  int32_t simType;                    /* prhs[1]: 1 = MC, 2 = FMC (MCmatlab.c:322) */
  int32_t silentMode;                 /* silentMode (:325) */
  int32_t calcNFR;                    /* calcNFR (:326) */
  const float *sourceDistribution;    /* sourceDistribution single [nx,ny,nz,nL], NULL when NaN (:329-330) */
  double nPhotonsRequested;           /* nPhotonsRequested; NaN = timed run (:333, :337) */
  int32_t requestCollectedPhotons;    /* requestCollectedPhotons (:334) */
  double simulationTimeRequested;     /* simulationTimeRequested [min] (:336) */
  int64_t nM, nL;                     /* size(mediaProperties.mua) (:356-357) */
  const double *mua, *mus, *g, *n;    /* mediaProperties.{mua,mus,g,n} [nM,nL] (:351-354) */
  const double *CDFidx;               /* mediaProperties.CDFidx [nM,nL] (:355) */
  int64_t nx, ny, nz;                 /* size(M) (:362, :380-382) */
  double dx, dy, dz;                  /* G.dx, G.dy, G.dz (:377-379) */
  int64_t farFieldRes;                /* farFieldRes (:383) */
  int32_t boundaryType;               /* boundaryType 0..3 (:384) */
  const float *interfaceNormals;      /* interfaceNormals single [2,nx*ny*nz] (theta,phi) or NULL (:386) */
  const uint8_t *M;                   /* M uint8 [nx,ny,nz], 1-based (:387-388) */
  const double *CDFs; int64_t nCDFs;  /* CDFs double, numel (:390-399) */
  /* depositionCriteria (:402-416); Inf allowed for the max* and idx fields */
  double minScatterings, maxScatterings;
  double minRefractions, maxRefractions;
  double minReflections, maxReflections;
  double minInterfaceTransitions, maxInterfaceTransitions;
  double minSubmediaIdx, maxSubmediaIdx;   /* 1-based */
  int32_t onlyCollected, evaluateOnlyAtEndOfLife;
  /* light collector (:427-447) */
  int32_t useLightCollector;
  double LC_x, LC_y, LC_z, LC_theta, LC_phi, LC_f, LC_diam;
  double LC_fieldSize;                /* used when f is finite */
  double LC_NA;                       /* used when f is infinite */
  double LC_res, LC_nTimeBins, LC_tStart, LC_tEnd;
  int64_t nExamplePaths;              /* nExamplePaths (:453) */
  const double *spectrum;             /* spectrum [nL] (:631) */
  int32_t useAllCPUs;                 /* useAllCPUs (:727); ignored by the GPU build */
  int32_t GPUdevice;                  /* GPUdevice (:638) */
  /* light source LS (:363-366, :579-614) */
  double sourceType, emitterLength;
  double LS_theta, LS_phi, LS_psi;
  double xFocus, yFocus, zFocus;
  mcm_beam_distr FPID, AID;
  /* ---- extensions with no reference counterpart ---- */
  uint64_t seed;                      /* RNG key; the reference seeds from the wall clock (MCmatlab.c:241, 725) */
  int (*isInterruptPending)(void);    /* polled between launches; NULL = never (reference: utIsInterruptPending) */
} mcm_mc_input;

/* What MCmatlab.c's mexFunction writes into the duplicated model.  Pointers
 * for outputs that the input does not request may be NULL. */
typedef struct mcm_mc_output {
  float *NFR;            /* NFR single [nx,ny,nz,nL] when calcNFR (:467, normalisation MCmatlablib.c:1204-1207) */
  float *image;          /* LC.image single [res,res,nTimeBins+2 or 1,nL] when useLightCollector (:469-470) */
  float *farField;       /* farField single [ffr,ffr,nL] when farFieldRes (:472-474) */
  float *NI_xpos, *NI_xneg, *NI_ypos, *NI_yneg, *NI_zpos, *NI_zneg;  /* (:476-493) */
  double *examplePaths;  /* library-allocated [4,examplePathsElems]; release with mcm_free (:772-776) */
  int64_t examplePathsElems;
  double nPhotons, nPhotonsCollected, nThreads, simulationTime;  /* (:755-763); simulationTime in minutes */
} mcm_mc_output;

/* Element counts of every output array for a given input (0 = not produced). */
typedef struct mcm_mc_sizes {
  int64_t NFR, image, farField;
  int64_t NI_xpos, NI_xneg, NI_ypos, NI_yneg, NI_zpos, NI_zneg;
  int64_t image_dims[4], NI_dims[6][3];
} mcm_mc_sizes;

/* Everything finiteElementHeatPropagator.c's mexFunction reads. */
typedef struct mcm_heat_input {
  int64_t nx, ny, nz;                 /* size(T) (:694-698) */
  int64_t nM;                         /* size(P.dTdtperdeltaT,1) (:699) */
  const float *T;                     /* T single [nx,ny,nz] (:702) */
  const float *Omega;                 /* Omega single [nx,ny,nz], or NULL when Omega is NaN (:722-723) */
  int64_t steps;                      /* P.steps (:701) */
  double dt;                          /* P.dt (:707) */
  const uint8_t *M;                   /* P.M uint8 [nx,ny,nz], 0-based (:708) */
  const float *dTdtperdeltaT;         /* P.dTdtperdeltaT single [nM,nM,3] (:709) */
  const float *dTdt_abs;              /* P.dTdt_abs single [nx,ny,nz] (:710) */
  const double *A, *E;                /* P.A, P.E double [nM] (:711-718) */
  int32_t lightsOn;                   /* P.lightsOn (:719) */
  int32_t heatBoundaryType;           /* P.heatBoundaryType: 0 insulating, 1 fixed-temperature boundary voxels (:720) */
  const double *tempSensorCornerIdxs; /* P.tempSensorCornerIdxs double [nS], 0-based (:728) */
  int64_t nSensors;                   /* numel(P.tempSensorCornerIdxs) (:729) */
  const double *tempSensorInterpWeights; /* P.tempSensorInterpWeights double [nS,3] (:730) */
  int32_t useAllCPUs;                 /* P.useAllCPUs (:755); ignored by the GPU build */
  /* ---- extensions with no reference counterpart ---- */
  int32_t precision;                  /* MCM_HEAT_FP64 (default) or MCM_HEAT_FP32 */
  int32_t scheme;                     /* MCM_HEAT_DGADI (default) or MCM_HEAT_EXPLICIT */
  const double *T_f64;                /* optional FP64 state replacing T (chained FP64 calls) */
  const double *Omega_f64;            /* optional FP64 state replacing Omega */
  int32_t GPUdevice;
  int (*isInterruptPending)(void);    /* polled between steps; NULL = never */
} mcm_heat_input;

typedef struct mcm_heat_output {
  float *T;               /* plhs[0] single [nx,ny,nz] */
  float *Omega;           /* plhs[1] single [nx,ny,nz], or 1 element set to NaN without damage */
  double *sensorTemps;    /* plhs[2] double [nS, steps+1] */
  float *maxMediaTemps;   /* plhs[3] single [1,nM] */
  double *T_f64;          /* optional FP64 copy of the final state */
  double *Omega_f64;      /* optional FP64 copy of the final damage integral */
  int64_t stepsDone;      /* < steps only when interrupted */
} mcm_heat_output;

/* ---- library ---- */
const char *mcm_last_error(void);
int mcm_version(void);
/* Number of visible gfx950 devices (0 when none); never fails. */
int mcm_device_count(void);
void mcm_free(void *p);

/* ---- Monte Carlo ---- */
int mcm_mc_output_sizes(const mcm_mc_input *in, mcm_mc_sizes *sizes);
/* Whole MEX-equivalent call: all wavelengths, host buffers in and out. */
int mcm_mc_run(const mcm_mc_input *in, mcm_mc_output *out);

/* Device-resident form of the same call, for callers that keep the model on
 * the GPU across many launches (benchmarks, sweeps, gradient loops).
 * plan_create uploads the grid and tables once; plan_simulate launches
 * nPhotons packets of wavelength iL into the plan's FP64 tallies (async on
 * `stream`); plan_normalize waits, normalises like the reference and writes
 * the wavelength's slice of `out`, then clears the tallies. */
typedef struct mcm_mc_plan mcm_mc_plan;
int mcm_mc_plan_create(const mcm_mc_input *in, mcm_mc_plan **plan);
int mcm_mc_plan_simulate(mcm_mc_plan *plan, int64_t iL, uint64_t nPhotons, uint64_t seed,
                         uint64_t firstPhoton, void *stream);
int mcm_mc_plan_normalize(mcm_mc_plan *plan, int64_t iL, mcm_mc_output *out);
/* Raw FP64 tallies on the device (absorbed weight per voxel, before
 * normalisation), valid until the next plan call. */
int mcm_mc_plan_tallies(mcm_mc_plan *plan, const double **d_deposit, uint64_t *nPhotons);
/* Event counters accumulated since the last plan_normalize (synchronises):
 * [0] packets launched alive, [1] collected, [2] voxel sub-steps,
 * [3] scatterings, [4] interface (Fresnel) events, [5] packets retired at the
 * sub-step cap, [6] aborted (1 = a packet failed 10^6 launch attempts),
 * [7] lanes per launch. */
int mcm_mc_plan_counters(mcm_mc_plan *plan, double counters[8]);
int mcm_mc_plan_destroy(mcm_mc_plan *plan);
/* Launch tuning (no reference counterpart).  The fluence accumulates into 1,
 * 2, 4 or 8 FP64 copies, one per group of the GPU's XCDs, summed before any
 * read: more copies relieve atomic contention on hot voxels (narrow beams),
 * fewer keep the working set small (broad beams).  Times one launch of
 * nPackets packets (0: 2^21) of wavelength iL per copy count (within a 2 GiB
 * budget), keeps the fewest copies within 10 % of the fastest trial, then
 * clears the plan's tallies and counters.
 * ms[4] (optional) receives the trial times for 1, 2, 4, 8 copies (0 =
 * skipped).  mcm_mc_run calls it for runs of 2e7 packets or more. */
int mcm_mc_plan_autotune(mcm_mc_plan *plan, int64_t iL, uint64_t nPackets, double ms[4]);
/* Set the number of XCD-private FP64 fluence copies (1, 2, 4 or 8) by hand;
 * tallies so far are folded into the first copy and kept.  No reference
 * counterpart. */
int mcm_mc_plan_set_fluence_copies(mcm_mc_plan *plan, int32_t copies);

/* Batched launch (no reference counterpart: the reference runs one
 * wavelength after another, MCmatlab.c:619-747).  Item i asks for nPhotons
 * packets of wavelength iL drawn from Philox stream (seed, stream); an item
 * reproduces mcm_mc_plan_simulate(plan, iL, nPhotons, seed, 0, .) when
 * stream == iL.  All items run in one sequence of persistent launches whose
 * lanes take packets of any item, so small items do not leave the GPU idle. */
typedef struct mcm_mc_item {
  int64_t iL;          /* wavelength (column of mediaProperties) */
  uint64_t nPhotons;   /* packets of this item */
  uint64_t seed;       /* Philox key */
  uint32_t stream;     /* Philox stream id */
  uint32_t reserved;   /* 0 */
} mcm_mc_item;
/* Runs nItems (1..256) items; outs[i] receives item i's outputs normalised
 * like one wavelength of mcm_mc_run (NFR [nx,ny,nz], image, farField and NI
 * faces sized as mcm_mc_output_sizes reports divided by nL), nPhotons,
 * nPhotonsCollected, nThreads and the batch's simulationTime.  Deposition
 * criteria, 3D sources and wavelength-dependent refractive indices are
 * MCM_ERR_UNSUPPORTED; nItems * nM must not exceed 2048. */
int mcm_mc_batch_run(mcm_mc_plan *plan, const mcm_mc_item *items, int64_t nItems, mcm_mc_output *outs);

/* Work items on different grids (no reference counterpart): item i runs
 * nPhotons packets of plans[plan] at wavelength iL with Philox key seed and
 * stream id `stream`, and reproduces mcm_mc_plan_simulate of that plan when
 * stream == iL.  The items run in one sequence of persistent launches per
 * interface class (plans with refractive-index changes, plans without);
 * each lane reads the grid, media and tally layout of its packet's plan.
 * outs[i] is normalised like mcm_mc_batch_run's.  The plans share one
 * device; deposition criteria and 3D sources are MCM_ERR_UNSUPPORTED, the
 * items of one plan need equal refractive indices, nItems is 1..256 and
 * nItems * (largest nM) must not exceed 2048.  Uses plans[0]'s stream,
 * counters and packet records. */
typedef struct mcm_mc_multi_item {
  int64_t plan;        /* index into plans */
  int64_t iL;          /* wavelength of that plan */
  uint64_t nPhotons;
  uint64_t seed;
  uint32_t stream;
  uint32_t reserved;   /* 0 */
} mcm_mc_multi_item;
int mcm_mc_multi_run(mcm_mc_plan *const *plans, int64_t nPlans, const mcm_mc_multi_item *items, int64_t nItems,
                     mcm_mc_output *outs);

/* Differentiable absorbed power (no reference counterpart).  J = power
 * absorbed inside a box of voxels per watt of source power (spectrum[iL]
 * scales it like the NFR normalisation, MCmatlablib.c:1195-1200), its
 * gradient with respect to a per-voxel absorption coefficient map, and its
 * derivative with respect to a common scale s of every medium's mus
 * (mus -> s mus; for one shared mus, dJ/dmus = dJ_dlogmus / mus).  The map
 * overrides the media table's mua voxel by voxel (NULL: the table's). */
typedef struct mcm_mc_grad_input {
  const float *muaMap;   /* [nx,ny,nz] 1/cm, or NULL */
  int64_t box[6];        /* voxels x0 <= x < x1, y0 <= y < y1, z0 <= z < z1 (0-based) */
  int64_t iL;            /* wavelength */
  uint64_t nPhotons;
  uint64_t seed;         /* Philox key; the stream is iL, so equal seeds give common random numbers */
} mcm_mc_grad_input;
typedef struct mcm_mc_grad_output {
  double J, J_se;                  /* estimate and its standard error */
  double *dJ_dmua;                 /* caller-owned [nx,ny,nz] or NULL: dJ/dmua per voxel [W cm / W];
                                      NULL skips the replay pass (J and dJ_dlogmus need one pass) */
  double dJ_dlogmus, dJ_dlogmus_se;
} mcm_mc_grad_output;
int mcm_mc_grad_run(mcm_mc_plan *plan, const mcm_mc_grad_input *in, mcm_mc_grad_output *out);

/* ---- heat ---- */
int mcm_heat_run(const mcm_heat_input *in, mcm_heat_output *out);

typedef struct mcm_heat_plan mcm_heat_plan;
int mcm_heat_plan_create(const mcm_heat_input *in, mcm_heat_plan **plan);
/* Advance `steps` steps of size dt with the given illumination; async on the
 * plan's stream.  Sensors are recorded into the plan's device buffer. */
int mcm_heat_plan_advance(mcm_heat_plan *plan, int64_t steps, double dt, int32_t lightsOn);
/* Runs of equal (dt, lightsOn) steps are replayed as captured HIP graphs of
 * `steps` (even) time steps each, one launch per graph instead of 2-4 per
 * step; 0 launches every step.  Default 32.  No reference counterpart (the
 * reference launches its kernels step by step,
 * finiteElementHeatPropagator.c:764-798). */
int mcm_heat_plan_set_graph_steps(mcm_heat_plan *plan, int64_t steps);
/* Explicit scheme only: z planes per workgroup of the stencil (8, 12, 16, 24
 * or 32).  A plan picks it at creation by timing the candidates once per
 * grid shape in the process (plans created for fewer than 64 steps take
 * 16 unless the shape was timed before; environment MCM_HEAT_ZRUN forces
 * one); the decomposition does not change any result.  No reference
 * counterpart (launch shape of this implementation). */
int mcm_heat_plan_set_explicit_zrun(mcm_heat_plan *plan, int32_t zPlanes);
int32_t mcm_heat_plan_explicit_zrun(const mcm_heat_plan *plan);
int mcm_heat_plan_download(mcm_heat_plan *plan, mcm_heat_output *out);

/* Several plans (different grids and models, one device) advanced together:
 * runs of stepsPerGraph (even) steps of every plan are captured as one graph
 * whose branches (one per plan, on the plan's stream) run concurrently, and
 * replayed; remaining steps are launched per plan.  dt[i] and lightsOn[i]
 * are plan i's.  Results equal advancing each plan alone.  A plan stays
 * usable on its own between group calls; destroy the group before its plans.
 * No reference counterpart (batched sweeps of independent models). */
typedef struct mcm_heat_group mcm_heat_group;
int mcm_heat_group_create(mcm_heat_plan *const *plans, int64_t nPlans, mcm_heat_group **group);
int mcm_heat_group_advance(mcm_heat_group *group, int64_t steps, const double *dt, const int32_t *lightsOn, int64_t stepsPerGraph);
int mcm_heat_group_destroy(mcm_heat_group *group);
int mcm_heat_plan_destroy(mcm_heat_plan *plan);

/* Discrete adjoint of the FP64 DG-ADI propagation (no reference
 * counterpart).  Runs the schedule (nSegments segments of `steps` steps of
 * size dt with illumination lightsOn; nSegments == 0: in->steps, in->dt,
 * in->lightsOn) from in->T / in->Omega and returns
 *   J      = sum over the target voxels of the final damage integral Omega
 *            (targetMask [nx,ny,nz] non-zero, or NULL = voxels whose medium
 *            has E != 0),
 *   dJ_dQ  = dJ / d dTdt_abs per voxel (lights-on steps only) [1 per K/s],
 *   dJ_dT0 = dJ / d T(initial) per voxel [1/K],
 *   dJ_dc  = dJ / d dTdtperdeltaT [nM,nM,3] (the conduction coefficients
 *            c[a,b,d] = TC_eff(a,b) / VHC(a) / d^2 of
 *            simulateHeatDistribution.m:469-494, which the thermal
 *            conductivities enter through; chain rule to TC on the caller's
 *            side) [K s],
 * the exact derivatives of the discrete scheme (up to rounding).  States are
 * checkpointed every checkpointEvery steps (0: 100) and recomputed one
 * segment at a time on the way back.  Needs precision MCM_HEAT_FP64, scheme
 * MCM_HEAT_DGADI and a damage input (Omega). */
typedef struct mcm_heat_segment {
  int64_t steps;
  double dt;
  int32_t lightsOn;
  int32_t reserved;
} mcm_heat_segment;
typedef struct mcm_heat_adjoint_output {
  double J;
  double *dJ_dQ;      /* caller-owned [nx,ny,nz] or NULL */
  double *dJ_dT0;     /* caller-owned [nx,ny,nz] or NULL */
  double *Omega_f64;  /* caller-owned final damage [nx,ny,nz] or NULL */
  double *dJ_dc;      /* caller-owned [nM,nM,3] (index a + b nM + d nM^2) or NULL */
} mcm_heat_adjoint_output;
int mcm_heat_adjoint_run(const mcm_heat_input *in, const mcm_heat_segment *segments, int64_t nSegments,
                         const uint8_t *targetMask, int64_t checkpointEvery, mcm_heat_adjoint_output *out);

/* ---- Media binning ------------------------------------------------------
 * [downsampledParamVals, binidx] = getDownsampledParamVals(paramVals, N2)
 * (replaces the MEX +MCmatlab/src/getDownsampledParamVals.c:54-202, called
 * from getOpticalMediaProperties.m:157 and getThermalMediaProperties.m:80).
 * paramVals: host [3 x n1] single, column-major (one (mua, mus, g) or
 * (VHC, TC, 1) set per voxel of a medium); downsampled: host [3 x n2];
 * binidx: host [n1], 1-based label of the nearest downsampled set (uint8,
 * wrapping like the reference's for n2 > 255).  Farthest-point sampling in
 * the range-normalised space of the rows that vary, with the reference's
 * arithmetic and its single-threaded selection order, on GPU `device`.
 * n1, n2 >= 1; non-finite values are MCM_ERR_INVALID_ARGUMENT. */
int mcm_downsample_params(const float *paramVals, int64_t n1, int64_t n2, float *downsampled, uint8_t *binidx,
                          int32_t device);
/* The same without any device allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory, such as serving loops and
 * harnesses that forbid hipMalloc.  Every intermediate buffer lives in a
 * file-scope device array of the library (about 30 MB, held by each GPU it
 * runs on until the process ends), so n1 <= 2^20 parameter sets and
 * n2 <= 4096 bins (larger requests are MCM_ERR_INVALID_ARGUMENT).  Runs on
 * `stream` (a hipStream_t, or NULL for the default stream) and returns when
 * the outputs are in the host arrays.  Calls on one GPU share the array:
 * calls that overlap on the same GPU race and must be serialised by the
 * caller. */
int mcm_downsample_params_noalloc(const float *paramVals, int64_t n1, int64_t n2, float *downsampled, uint8_t *binidx,
                                  int32_t device, void *stream);

#ifdef __cplusplus
}
#endif
#endif
