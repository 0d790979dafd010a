/*
 * getDownsampledParamVals_mex.c -- MEX gateway for the GPU media binning
 * (libmcm).
 *
 * Drop-in for the reference MEX file getDownsampledParamVals.<mexext>
 * (+MCmatlab/src/getDownsampledParamVals.c:54-202):
 *   [downsampledParamVals, binidx] = getDownsampledParamVals(paramVals, N2)
 * called by @model/private/getOpticalMediaProperties.m:157 and
 * getThermalMediaProperties.m:80 with paramVals a 3 x N1 single matrix and N2
 * the medium's nBins.  Outputs: 3 x N2 single and 1 x N1 uint8, bit for bit
 * the reference's.
 *
 * Build (MATLAB):   mex -R2018a getDownsampledParamVals_mex.c -I<repo>/include -L<repo>/mcmatlab_amd -lmcm
 *                       -output getDownsampledParamVals
 */
#include <stdint.h>

#include "mcm_abi.h"
#include "mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  (void)nlhs;
  if (nrhs != 2) mexErrMsgIdAndTxt("MCmatlab:args", "usage: [downsampledParamVals,binidx] = getDownsampledParamVals(paramVals,N2)");
  if (mxGetClassID(prhs[0]) != mxSINGLE_CLASS || mxGetM(prhs[0]) != 3)
    mexErrMsgIdAndTxt("MCmatlab:invalidInput", "paramVals must be a 3 x N single matrix");
  const int64_t n1 = (int64_t)mxGetN(prhs[0]);
  const int64_t n2 = (int64_t)*mxGetPr(prhs[1]);
  if (n1 < 1 || n2 < 1) mexErrMsgIdAndTxt("MCmatlab:invalidInput", "paramVals and N2 must be non-empty");
  plhs[0] = mxCreateNumericMatrix(3, (mwSize)n2, mxSINGLE_CLASS, mxREAL);
  plhs[1] = mxCreateNumericMatrix(1, (mwSize)n1, mxUINT8_CLASS, mxREAL);
  const int rc = mcm_downsample_params((const float *)mxGetData(prhs[0]), n1, n2, (float *)mxGetData(plhs[0]),
                                       (uint8_t *)mxGetData(plhs[1]), 0);
  if (rc != MCM_OK) mexErrMsgIdAndTxt(rc == MCM_ERR_OUT_OF_MEMORY ? "MCmatlab:OutOfMemory" : "MCmatlab:GPU", "%s", mcm_last_error());
}
