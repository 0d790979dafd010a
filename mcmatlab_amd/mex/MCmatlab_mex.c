/*
 * MCmatlab_mex.c -- MEX gateway for the GPU Monte Carlo (libmcm).
 *
 * Drop-in for the reference MEX file MCmatlab.<mexext>
 * (+MCmatlab/src/MCmatlab.c:317-791): model = MCmatlab(model, simType), called
 * by @model/runMonteCarlo.m:99-104 and :168-175.  It reads the same
 * properties of model.MC (or model.FMC), model.MC.LS/LC/depositionCriteria
 * and model.G, runs mcm_mc_run, and writes NFR, farField, NI_*, LC.image,
 * examplePaths, nPhotons, nPhotonsCollected, nThreads and simulationTime into
 * a duplicate of the model, with the reference's array classes and sizes.
 *
 * Build (MATLAB):   mex -R2018a MCmatlab_mex.c -I<repo>/include -L<repo>/mcmatlab_amd -lmcm -output MCmatlab
 * Errors raise MATLAB errors with identifiers MCmatlab:<reason>.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mcm_abi.h"
#include "mex.h"

static double num(const mxArray *obj, const char *name) { return *mxGetPr(mxGetPropertyShared(obj, 0, name)); }
static int flag(const mxArray *obj, const char *name) { return mxIsLogicalScalarTrue(mxGetPropertyShared(obj, 0, name)) ? 1 : 0; }

static void fill_distr(const mxArray *d, mcm_beam_distr *b) {
  mxArray *a;
  a = mxGetPropertyShared(d, 0, "radialDistr");
  b->radialDistr = mxGetPr(a);
  b->n_radialDistr = (int64_t)mxGetN(a);
  b->radialWidth = num(d, "radialWidth");
  a = mxGetPropertyShared(d, 0, "XDistr");
  b->XDistr = mxGetPr(a);
  b->n_XDistr = (int64_t)mxGetN(a);
  b->XWidth = num(d, "XWidth");
  a = mxGetPropertyShared(d, 0, "YDistr");
  b->YDistr = mxGetPr(a);
  b->n_YDistr = (int64_t)mxGetN(a);
  b->YWidth = num(d, "YWidth");
}

static int interrupt_poll(void) { return utIsInterruptPending() ? 1 : 0; }

static mxArray *single_array(int nd, int64_t d0, int64_t d1, int64_t d2, int64_t d3) {
  mwSize dims[4] = {(mwSize)d0, (mwSize)d1, (mwSize)d2, (mwSize)d3};
  return mxCreateNumericArray((mwSize)nd, dims, mxSINGLE_CLASS, mxREAL);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  (void)nlhs;
  if (nrhs != 2) mexErrMsgIdAndTxt("MCmatlab:args", "usage: model = MCmatlab(model, simType)");
  const int simType = *mxGetPr(prhs[1]) == 2 ? 2 : 1;
  const mxArray *MC = mxGetPropertyShared(prhs[0], 0, simType == 2 ? "FMC" : "MC");
  const mxArray *G = mxGetPropertyShared(prhs[0], 0, "G");
  mcm_mc_input in;
  memset(&in, 0, sizeof in);
  in.simType = simType;
  in.silentMode = flag(MC, "silentMode");
  in.calcNFR = flag(MC, "calcNFR");
  const mxArray *LS = mxGetPropertyShared(MC, 0, "LS");
  const float *S_PDF = (const float *)mxGetData(mxGetPropertyShared(MC, 0, "sourceDistribution"));
  in.sourceDistribution = mxIsNaN(S_PDF[0]) ? NULL : S_PDF;
  in.nPhotonsRequested = num(MC, "nPhotonsRequested");
  in.requestCollectedPhotons = flag(MC, "requestCollectedPhotons");
  in.simulationTimeRequested = num(MC, "simulationTimeRequested");
  const mxArray *mp = mxGetPropertyShared(MC, 0, "mediaProperties");
  const mxArray *mua = mxGetField(mp, 0, "mua");
  in.nM = (int64_t)mxGetM(mua);
  in.nL = (int64_t)mxGetN(mua);
  in.mua = mxGetPr(mua);
  in.mus = mxGetPr(mxGetField(mp, 0, "mus"));
  in.g = mxGetPr(mxGetField(mp, 0, "g"));
  in.n = mxGetPr(mxGetField(mp, 0, "n"));
  in.CDFidx = mxGetPr(mxGetField(mp, 0, "CDFidx"));
  const mxArray *Marr = mxGetPropertyShared(MC, 0, "M");
  const mwSize *dims = mxGetDimensions(Marr);
  in.nx = (int64_t)dims[0];
  in.ny = (int64_t)dims[1];
  in.nz = mxGetNumberOfDimensions(Marr) > 2 ? (int64_t)dims[2] : 1;
  in.M = (const uint8_t *)mxGetData(Marr);
  in.dx = num(G, "dx");
  in.dy = num(G, "dy");
  in.dz = num(G, "dz");
  in.farFieldRes = (int64_t)num(MC, "farFieldRes");
  in.boundaryType = (int32_t)num(MC, "boundaryType");
  const float *IN = (const float *)mxGetData(mxGetPropertyShared(MC, 0, "interfaceNormals"));
  in.interfaceNormals = (mxGetNumberOfElements(mxGetPropertyShared(MC, 0, "interfaceNormals")) >= 2) ? IN : NULL;
  const mxArray *cdfs = mxGetPropertyShared(MC, 0, "CDFs");
  in.nCDFs = (int64_t)mxGetNumberOfElements(cdfs);
  in.CDFs = in.nCDFs ? mxGetPr(cdfs) : NULL;
  const mxArray *DC = mxGetPropertyShared(MC, 0, "depositionCriteria");
  in.minScatterings = num(DC, "minScatterings");
  in.maxScatterings = num(DC, "maxScatterings");
  in.minRefractions = num(DC, "minRefractions");
  in.maxRefractions = num(DC, "maxRefractions");
  in.minReflections = num(DC, "minReflections");
  in.maxReflections = num(DC, "maxReflections");
  in.minInterfaceTransitions = num(DC, "minInterfaceTransitions");
  in.maxInterfaceTransitions = num(DC, "maxInterfaceTransitions");
  in.minSubmediaIdx = num(DC, "minSubmediaIdx");
  in.maxSubmediaIdx = num(DC, "maxSubmediaIdx");
  in.onlyCollected = flag(DC, "onlyCollected");
  in.evaluateOnlyAtEndOfLife = flag(DC, "evaluateOnlyAtEndOfLife");
  const mxArray *LC = mxGetPropertyShared(MC, 0, "LC");
  in.useLightCollector = flag(MC, "useLightCollector");
  in.LC_x = num(LC, "x");
  in.LC_y = num(LC, "y");
  in.LC_z = num(LC, "z");
  in.LC_theta = num(LC, "theta");
  in.LC_phi = num(LC, "phi");
  in.LC_f = num(LC, "f");
  in.LC_diam = num(LC, "diam");
  in.LC_fieldSize = num(LC, "fieldSize");
  in.LC_NA = num(LC, "NA");
  in.LC_res = num(LC, "res");
  in.LC_nTimeBins = num(LC, "nTimeBins");
  in.LC_tStart = num(LC, "tStart");
  in.LC_tEnd = num(LC, "tEnd");
  in.nExamplePaths = (int64_t)num(MC, "nExamplePaths");
  in.spectrum = mxGetPr(mxGetPropertyShared(MC, 0, "spectrum"));
  in.useAllCPUs = flag(MC, "useAllCPUs");
  {
    const mxArray *gd = mxGetPropertyShared(MC, 0, "GPUdevice");
    in.GPUdevice = mxIsDouble(gd) ? (int32_t)*mxGetPr(gd) : *(const int32_t *)mxGetData(gd);
  }
  if (!in.sourceDistribution) {
    in.sourceType = num(LS, "sourceType");
    in.emitterLength = num(LS, "emitterLength");
    in.LS_theta = num(LS, "theta");
    in.LS_phi = num(LS, "phi");
    in.LS_psi = num(LS, "psi");
    in.xFocus = num(LS, "xFocus");
    in.yFocus = num(LS, "yFocus");
    in.zFocus = num(LS, "zFocus");
    if (in.sourceType >= 4) {
      fill_distr(mxGetPropertyShared(LS, 0, "FPID"), &in.FPID);
      fill_distr(mxGetPropertyShared(LS, 0, "AID"), &in.AID);
    }
  }
  /* The reference seeds from the microsecond clock (MCmatlab.c:241, 725); so
   * does this gateway, unless the environment variable MCM_MC_SEED fixes the
   * packet streams (setenv('MCM_MC_SEED','42') in MATLAB) for reproducible runs. */
  {
    const char *env = getenv("MCM_MC_SEED");
    if (env && *env) {
      in.seed = (uint64_t)strtoull(env, NULL, 10);
    } else {
      struct timespec ts;
      clock_gettime(CLOCK_REALTIME, &ts);
      in.seed = (uint64_t)ts.tv_sec * 1000000u + (uint64_t)(ts.tv_nsec / 1000);
    }
  }
  in.isInterruptPending = interrupt_poll;

  mcm_mc_sizes sz;
  int rc = mcm_mc_output_sizes(&in, &sz);
  if (rc != MCM_OK) mexErrMsgIdAndTxt("MCmatlab:invalidInput", "%s", mcm_last_error());

  plhs[0] = mxDuplicateArray(prhs[0]);
  mxArray *MCout = mxGetPropertyShared(plhs[0], 0, simType == 2 ? "FMC" : "MC");
  mxArray *LCout = mxGetPropertyShared(MCout, 0, "LC");
  mcm_mc_output out;
  memset(&out, 0, sizeof out);
  if (sz.NFR) {
    mxArray *a = single_array(4, in.nx, in.ny, in.nz, in.nL);
    mxSetPropertyShared(MCout, 0, "NFR", a);
    out.NFR = (float *)mxGetData(a);
  }
  if (sz.image) {
    mxArray *a = single_array(4, sz.image_dims[0], sz.image_dims[1], sz.image_dims[2], sz.image_dims[3]);
    mxSetPropertyShared(LCout, 0, "image", a);
    out.image = (float *)mxGetData(a);
  }
  if (sz.farField) {
    mxArray *a = single_array(3, in.farFieldRes, in.farFieldRes, in.nL, 1);
    mxSetPropertyShared(MCout, 0, "farField", a);
    out.farField = (float *)mxGetData(a);
  }
  {
    static const char *names[6] = {"NI_xpos", "NI_xneg", "NI_ypos", "NI_yneg", "NI_zpos", "NI_zneg"};
    float **slot[6] = {&out.NI_xpos, &out.NI_xneg, &out.NI_ypos, &out.NI_yneg, &out.NI_zpos, &out.NI_zneg};
    const int64_t n[6] = {sz.NI_xpos, sz.NI_xneg, sz.NI_ypos, sz.NI_yneg, sz.NI_zpos, sz.NI_zneg};
    for (int f = 0; f < 6; ++f) {
      if (!n[f]) continue;
      mxArray *a = single_array(3, sz.NI_dims[f][0], sz.NI_dims[f][1], sz.NI_dims[f][2], 1);
      mxSetPropertyShared(MCout, 0, names[f], a);
      *slot[f] = (float *)mxGetData(a);
    }
  }
  if (!in.silentMode)
    mexPrintf(simType == 2 ? "-----------Fluorescence Monte Carlo Simulation-----------\n"
                           : "-----------------Monte Carlo Simulation------------------\n");
  rc = mcm_mc_run(&in, &out);
  if (rc != MCM_OK) mexErrMsgIdAndTxt(rc == MCM_ERR_OUT_OF_MEMORY ? "MCmatlab:OutOfMemory" : "MCmatlab:GPU", "%s", mcm_last_error());
  if (!in.silentMode) {
    if (out.nPhotons == 0) mexPrintf("\nERROR: All photons launch outside simulation cuboid. Check your model definition.\n");
    else if (out.simulationTime > 0)
      mexPrintf("Simulated %0.2e photons over %0.2e minutes at a rate of %0.2e photons per minute\n", out.nPhotons, out.simulationTime,
                out.nPhotons / out.simulationTime);
  }
  mxArray *s = mxCreateDoubleMatrix(1, 1, mxREAL);
  *mxGetPr(s) = out.nPhotons;
  mxSetProperty(MCout, 0, "nPhotons", s);
  *mxGetPr(s) = out.nPhotonsCollected;
  mxSetProperty(MCout, 0, "nPhotonsCollected", s);
  *mxGetPr(s) = out.nThreads;
  mxSetProperty(MCout, 0, "nThreads", s);
  *mxGetPr(s) = out.simulationTime;
  mxSetProperty(MCout, 0, "simulationTime", s);
  mxDestroyArray(s);
  if (sz.image == 1) {
    /* a single-pixel, single-bin collector returns a scalar image (MCmatlab.c:765-770) */
    mxArray *o = mxCreateNumericMatrix(1, 1, mxSINGLE_CLASS, mxREAL);
    *(float *)mxGetData(o) = out.image[0];
    mxSetProperty(LCout, 0, "image", o);
    mxDestroyArray(o);
  }
  if (out.examplePaths) {
    mxArray *p = mxCreateNumericMatrix(4, (mwSize)out.examplePathsElems, mxDOUBLE_CLASS, mxREAL);
    memcpy(mxGetPr(p), out.examplePaths, (size_t)(4 * out.examplePathsElems) * sizeof(double));
    mxSetPropertyShared(MCout, 0, "examplePaths", p);
    mcm_free(out.examplePaths);
  }
}
