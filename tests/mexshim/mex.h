/*
 * mex.h -- a minimal, project-written stand-in for the subset of MATLAB's
 * MEX/matrix C API that MCmatlab's two MEX entry points use.
 *
 * TEST INFRASTRUCTURE ONLY.  It exists so that this project's own MEX layer
 * (the gateways under mcmatlab_amd/mex) can be exercised without MATLAB.
 * It is never linked into the shipped library.
 *
 * Data model: every mxArray is a scalar numeric/logical array, or a scalar
 * struct / value object whose fields are named mxArrays.  Arrays are column
 * major like MATLAB.  Property access on objects and field access on structs
 * share one implementation (MATLAB value objects behave like structs for
 * what the MEX files do with them).
 *
 * Functions named mxs_* are shim-only helpers used by the Python driver
 * (tests/mexdriver.py) to build inputs and read outputs.
 */
#ifndef MCM_MEXSHIM_MEX_H
#define MCM_MEXSHIM_MEX_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;

typedef enum {
  mxUNKNOWN_CLASS = 0,
  mxCELL_CLASS,
  mxSTRUCT_CLASS,
  mxLOGICAL_CLASS,
  mxCHAR_CLASS,
  mxVOID_CLASS,
  mxDOUBLE_CLASS,
  mxSINGLE_CLASS,
  mxINT8_CLASS,
  mxUINT8_CLASS,
  mxINT16_CLASS,
  mxUINT16_CLASS,
  mxINT32_CLASS,
  mxUINT32_CLASS,
  mxINT64_CLASS,
  mxUINT64_CLASS,
  mxFUNCTION_CLASS,
  mxOPAQUE_CLASS,
  mxOBJECT_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

/* ---- the MATLAB API subset ---- */
mxArray *mxGetProperty(const mxArray *pa, mwIndex i, const char *name);
mxArray *mxGetPropertyShared(const mxArray *pa, mwIndex i, const char *name);
void mxSetProperty(mxArray *pa, mwIndex i, const char *name, const mxArray *value);
void mxSetPropertyShared(mxArray *pa, mwIndex i, const char *name, const mxArray *value);
mxArray *mxGetField(const mxArray *pa, mwIndex i, const char *name);
int mxGetFieldNumber(const mxArray *pa, const char *name);
bool mxIsChar(const mxArray *pa);
int mxGetString(const mxArray *pa, char *buf, mwSize buflen);
double *mxGetPr(const mxArray *pa);
void *mxGetData(const mxArray *pa);
const mwSize *mxGetDimensions(const mxArray *pa);
mwSize mxGetNumberOfDimensions(const mxArray *pa);
size_t mxGetM(const mxArray *pa);
size_t mxGetN(const mxArray *pa);
size_t mxGetNumberOfElements(const mxArray *pa);
mxClassID mxGetClassID(const mxArray *pa);
bool mxIsSingle(const mxArray *pa);
bool mxIsDouble(const mxArray *pa);
bool mxIsUint8(const mxArray *pa);
bool mxIsLogical(const mxArray *pa);
bool mxIsStruct(const mxArray *pa);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID classid, mxComplexity flag);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray *mxCreateDoubleScalar(double value);
mxArray *mxCreateLogicalScalar(bool value);
mxArray *mxDuplicateArray(const mxArray *in);
void mxDestroyArray(mxArray *pa);
bool mxIsLogicalScalarTrue(const mxArray *pa);
bool mxIsNaN(double x);
bool mxIsInf(double x);
bool mxIsFinite(double x);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexPrintf(const char *fmt, ...);
int mexEvalString(const char *cmd);
bool utIsInterruptPending(void);

/* ---- shim-only helpers for the Python driver ---- */
mxArray *mxs_create_numeric(int classid, int ndim, const int64_t *dims, const void *data);
mxArray *mxs_create_object(void);
mxArray *mxs_create_struct(void);
int mxs_set(mxArray *obj, const char *name, mxArray *value); /* takes ownership */
mxArray *mxs_get(const mxArray *obj, const char *name);
int mxs_class(const mxArray *pa);
int mxs_ndim(const mxArray *pa);
int64_t mxs_dim(const mxArray *pa, int k);
void *mxs_data(const mxArray *pa);
int64_t mxs_numel(const mxArray *pa);
void mxs_destroy(mxArray *pa);
/* Call a mexFunction; returns 0 on success, 1 if it raised via
 * mexErrMsgIdAndTxt (message in mxs_last_error()). */
typedef void (*mxs_mexfunc_t)(int, mxArray **, int, const mxArray **);
int mxs_call(mxs_mexfunc_t fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs);
const char *mxs_last_error(void);
const char *mxs_last_error_id(void);
/* Make utIsInterruptPending() return true from its k-th call on (k<0: never). */
void mxs_set_interrupt_after(int64_t k);
int64_t mxs_interrupt_polls(void);

#ifdef __cplusplus
}
#endif
#endif
