/*
 * finiteElementHeatPropagator_mex.c -- MEX gateway for the GPU heat
 * propagator (libmcm).
 *
 * Drop-in for the reference MEX file finiteElementHeatPropagator.<mexext>
 * (+MCmatlab/src/finiteElementHeatPropagator.c:688-821):
 *   [T, Omega, sensorTemps, maxMediaTemps] = finiteElementHeatPropagator(T, Omega, P)
 * called by @model/simulateHeatDistribution.m:259-264, :320-325, :381-386,
 * with P holding M (0-based uint8), A, E, dTdtperdeltaT, dTdt_abs, useAllCPUs,
 * heatBoundaryType, tempSensorCornerIdxs, tempSensorInterpWeights, lightsOn,
 * steps and dt.  Outputs have the reference's classes and sizes.  The device
 * state is FP64; an optional field P.precision = 'single' selects the
 * reference's FP32 arithmetic ('double', the default, FP64).
 *
 * Build (MATLAB):   mex -R2018a finiteElementHeatPropagator_mex.c -I<repo>/include -L<repo>/mcmatlab_amd -lmcm
 *                       -output finiteElementHeatPropagator
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "mcm_abi.h"
#include "mex.h"

static int interrupt_poll(void) { return utIsInterruptPending() ? 1 : 0; }

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  (void)nlhs;
  if (nrhs != 3) mexErrMsgIdAndTxt("MCmatlab:args", "usage: [T,Omega,sensorTemps,maxMediaTemps] = finiteElementHeatPropagator(T,Omega,P)");
  const mxArray *P = prhs[2];
  const mwSize *dims = mxGetDimensions(prhs[0]);
  mcm_heat_input in;
  memset(&in, 0, sizeof in);
  in.nx = (int64_t)dims[0];
  in.ny = (int64_t)dims[1];
  in.nz = mxGetNumberOfDimensions(prhs[0]) > 2 ? (int64_t)dims[2] : 1;
  in.nM = (int64_t)mxGetM(mxGetField(P, 0, "dTdtperdeltaT"));
  in.steps = (int64_t)*mxGetPr(mxGetField(P, 0, "steps"));
  in.T = (const float *)mxGetData(prhs[0]);
  in.dt = *mxGetPr(mxGetField(P, 0, "dt"));
  in.M = (const uint8_t *)mxGetData(mxGetField(P, 0, "M"));
  in.dTdtperdeltaT = (const float *)mxGetData(mxGetField(P, 0, "dTdtperdeltaT"));
  in.dTdt_abs = (const float *)mxGetData(mxGetField(P, 0, "dTdt_abs"));
  in.A = mxGetPr(mxGetField(P, 0, "A"));
  in.E = mxGetPr(mxGetField(P, 0, "E"));
  in.lightsOn = mxIsLogicalScalarTrue(mxGetField(P, 0, "lightsOn"));
  in.heatBoundaryType = (int32_t)*mxGetPr(mxGetField(P, 0, "heatBoundaryType"));
  const float *Om = (const float *)mxGetData(prhs[1]);
  in.Omega = mxIsNaN(Om[0]) ? NULL : Om;
  in.tempSensorCornerIdxs = mxGetPr(mxGetField(P, 0, "tempSensorCornerIdxs"));
  in.nSensors = (int64_t)mxGetM(mxGetField(P, 0, "tempSensorCornerIdxs"));
  in.tempSensorInterpWeights = mxGetPr(mxGetField(P, 0, "tempSensorInterpWeights"));
  in.useAllCPUs = mxIsLogicalScalarTrue(mxGetField(P, 0, "useAllCPUs"));
  in.precision = MCM_HEAT_FP64;
  if (mxGetFieldNumber(P, "precision") >= 0) {  /* optional: 'single' or 'double' */
    const mxArray *pr = mxGetField(P, 0, "precision");
    char buf[8];
    if (!mxIsChar(pr) || mxGetString(pr, buf, sizeof buf) != 0 || (strcmp(buf, "single") != 0 && strcmp(buf, "double") != 0))
      mexErrMsgIdAndTxt("MCmatlab:invalidInput", "P.precision must be 'single' or 'double'");
    if (strcmp(buf, "single") == 0) in.precision = MCM_HEAT_FP32;
  }
  in.scheme = MCM_HEAT_DGADI;
  in.isInterruptPending = interrupt_poll;

  plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), dims, mxSINGLE_CLASS, mxREAL);
  plhs[1] = in.Omega ? mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), dims, mxSINGLE_CLASS, mxREAL)
                     : mxCreateNumericMatrix(1, 1, mxSINGLE_CLASS, mxREAL);
  plhs[2] = in.nSensors ? mxCreateDoubleMatrix((mwSize)in.nSensors, (mwSize)(in.steps + 1), mxREAL) : mxCreateDoubleMatrix(0, 0, mxREAL);
  plhs[3] = mxCreateNumericMatrix(1, (mwSize)in.nM, mxSINGLE_CLASS, mxREAL);
  mcm_heat_output out;
  memset(&out, 0, sizeof out);
  out.T = (float *)mxGetData(plhs[0]);
  out.Omega = (float *)mxGetData(plhs[1]);
  out.sensorTemps = in.nSensors ? mxGetPr(plhs[2]) : NULL;
  out.maxMediaTemps = (float *)mxGetData(plhs[3]);
  const int rc = mcm_heat_run(&in, &out);
  if (rc != MCM_OK) mexErrMsgIdAndTxt(rc == MCM_ERR_OUT_OF_MEMORY ? "MCmatlab:OutOfMemory" : "MCmatlab:GPU", "%s", mcm_last_error());
}
