/*
 * mexshim.c -- implementation of the project-written MATLAB mx-API stand-in
 * declared in mex.h.  TEST INFRASTRUCTURE ONLY (see mex.h).
 */
#define _GNU_SOURCE
#include "mex.h"
#include <setjmp.h>
#include <stdarg.h>
#include <time.h>

struct mxArray_tag {
  mxClassID cls;
  mwSize ndim;
  mwSize dims[8];
  size_t nel;
  void *data;
  /* struct / object fields */
  int nfields;
  int capfields;
  char **names;
  mxArray **values;
};

static size_t elem_size(mxClassID c) {
  switch (c) {
    case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
    case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
    case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
    case mxINT8_CLASS: case mxUINT8_CLASS: case mxLOGICAL_CLASS: return 1;
    default: return 0;
  }
}

static char g_err[1024];
static char g_err_id[256];
static jmp_buf *g_jmp = NULL;
static int64_t g_interrupt_after = -1;
static int64_t g_interrupt_polls = 0;

static void fail(const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  snprintf(g_err_id, sizeof g_err_id, "mexshim:internal");
  if (g_jmp) longjmp(*g_jmp, 1);
  fprintf(stderr, "mexshim: %s\n", msg);
  abort();
}

static mxArray *alloc_array(void) {
  mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
  if (!a) fail("out of memory");
  return a;
}

static mxArray *find_field(const mxArray *pa, const char *name) {
  if (!pa) return NULL;
  for (int k = 0; k < pa->nfields; k++)
    if (strcmp(pa->names[k], name) == 0) return pa->values[k];
  return NULL;
}

static void set_field(mxArray *pa, const char *name, mxArray *value) {
  for (int k = 0; k < pa->nfields; k++) {
    if (strcmp(pa->names[k], name) == 0) {
      if (pa->values[k] && pa->values[k] != value) mxs_destroy(pa->values[k]);
      pa->values[k] = value;
      return;
    }
  }
  if (pa->nfields == pa->capfields) {
    int cap = pa->capfields ? 2 * pa->capfields : 16;
    pa->names = (char **)realloc(pa->names, cap * sizeof(char *));
    pa->values = (mxArray **)realloc(pa->values, cap * sizeof(mxArray *));
    if (!pa->names || !pa->values) fail("out of memory");
    pa->capfields = cap;
  }
  pa->names[pa->nfields] = strdup(name);
  pa->values[pa->nfields] = value;
  pa->nfields++;
}

mxArray *mxGetPropertyShared(const mxArray *pa, mwIndex i, const char *name) {
  (void)i;
  mxArray *v = find_field(pa, name);
  if (!v) {
    char msg[300];
    snprintf(msg, sizeof msg, "no property/field '%s'", name);
    fail(msg);
  }
  return v;
}
/* MATLAB returns a copy here; the MEX files only read the result, so the
 * shared array is returned (the reference never frees these copies). */
mxArray *mxGetProperty(const mxArray *pa, mwIndex i, const char *name) { return mxGetPropertyShared(pa, i, name); }
mxArray *mxGetField(const mxArray *pa, mwIndex i, const char *name) { return mxGetPropertyShared(pa, i, name); }
int mxGetFieldNumber(const mxArray *pa, const char *name) {
  for (int k = 0; pa && k < pa->nfields; k++)
    if (strcmp(pa->names[k], name) == 0) return k;
  return -1;
}
bool mxIsChar(const mxArray *pa) { return pa && pa->cls == mxCHAR_CLASS; }
/* mxChar (16-bit) to a NUL-terminated char string; 1 if it does not fit */
int mxGetString(const mxArray *pa, char *buf, mwSize buflen) {
  if (!pa || pa->cls != mxCHAR_CLASS || buflen == 0) return 1;
  const uint16_t *c = (const uint16_t *)pa->data;
  size_t k = 0;
  for (; k < pa->nel && k + 1 < buflen; k++) buf[k] = (char)c[k];
  buf[k] = 0;
  return k < pa->nel ? 1 : 0;
}
void mxSetPropertyShared(mxArray *pa, mwIndex i, const char *name, const mxArray *value) {
  (void)i;
  set_field(pa, name, (mxArray *)value);
}
void mxSetProperty(mxArray *pa, mwIndex i, const char *name, const mxArray *value) {
  (void)i;
  set_field(pa, name, mxDuplicateArray(value));
}

double *mxGetPr(const mxArray *pa) { return (double *)pa->data; }
void *mxGetData(const mxArray *pa) { return pa->data; }
const mwSize *mxGetDimensions(const mxArray *pa) { return pa->dims; }
mwSize mxGetNumberOfDimensions(const mxArray *pa) { return pa->ndim; }
size_t mxGetM(const mxArray *pa) { return pa->dims[0]; }
size_t mxGetN(const mxArray *pa) {
  size_t n = 1;
  for (mwSize k = 1; k < pa->ndim; k++) n *= pa->dims[k];
  return n;
}
size_t mxGetNumberOfElements(const mxArray *pa) { return pa->nel; }
mxClassID mxGetClassID(const mxArray *pa) { return pa->cls; }
bool mxIsSingle(const mxArray *pa) { return pa->cls == mxSINGLE_CLASS; }
bool mxIsDouble(const mxArray *pa) { return pa->cls == mxDOUBLE_CLASS; }
bool mxIsUint8(const mxArray *pa) { return pa->cls == mxUINT8_CLASS; }
bool mxIsLogical(const mxArray *pa) { return pa->cls == mxLOGICAL_CLASS; }
bool mxIsStruct(const mxArray *pa) { return pa->cls == mxSTRUCT_CLASS; }

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID classid, mxComplexity flag) {
  (void)flag;
  if (ndim > 8) fail("too many dimensions");
  mxArray *a = alloc_array();
  a->cls = classid;
  /* MATLAB drops trailing singleton dimensions beyond the second */
  while (ndim > 2 && dims[ndim - 1] == 1) ndim--;
  a->ndim = ndim < 2 ? 2 : ndim;
  a->nel = 1;
  for (mwSize k = 0; k < a->ndim; k++) {
    a->dims[k] = k < ndim ? dims[k] : 1;
    a->nel *= a->dims[k];
  }
  size_t es = elem_size(classid);
  a->data = calloc(a->nel ? a->nel : 1, es ? es : 1);
  if (!a->data) fail("out of memory");
  return a;
}
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag) {
  mwSize d[2] = {m, n};
  return mxCreateNumericArray(2, d, classid, flag);
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag) {
  return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, flag);
}
mxArray *mxCreateDoubleScalar(double value) {
  mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
  *(double *)a->data = value;
  return a;
}
mxArray *mxCreateLogicalScalar(bool value) {
  mxArray *a = mxCreateNumericMatrix(1, 1, mxLOGICAL_CLASS, mxREAL);
  *(uint8_t *)a->data = value ? 1 : 0;
  return a;
}

mxArray *mxDuplicateArray(const mxArray *in) {
  if (!in) return NULL;
  mxArray *a = alloc_array();
  a->cls = in->cls;
  a->ndim = in->ndim;
  memcpy(a->dims, in->dims, sizeof a->dims);
  a->nel = in->nel;
  if (in->data) {
    size_t bytes = (in->nel ? in->nel : 1) * elem_size(in->cls);
    a->data = malloc(bytes ? bytes : 1);
    if (!a->data) fail("out of memory");
    memcpy(a->data, in->data, bytes);
  }
  for (int k = 0; k < in->nfields; k++) set_field(a, in->names[k], mxDuplicateArray(in->values[k]));
  return a;
}

void mxs_destroy(mxArray *pa) {
  if (!pa) return;
  for (int k = 0; k < pa->nfields; k++) {
    free(pa->names[k]);
    mxs_destroy(pa->values[k]);
  }
  free(pa->names);
  free(pa->values);
  free(pa->data);
  free(pa);
}
void mxDestroyArray(mxArray *pa) { mxs_destroy(pa); }

bool mxIsLogicalScalarTrue(const mxArray *pa) {
  return pa && pa->cls == mxLOGICAL_CLASS && pa->nel == 1 && ((uint8_t *)pa->data)[0] != 0;
}
bool mxIsNaN(double x) { return isnan(x); }
bool mxIsInf(double x) { return isinf(x); }
bool mxIsFinite(double x) { return isfinite(x); }

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  snprintf(g_err_id, sizeof g_err_id, "%s", id ? id : "");
  if (g_jmp) longjmp(*g_jmp, 1);
  fprintf(stderr, "mexErrMsgIdAndTxt outside mxs_call: %s: %s\n", g_err_id, g_err);
  abort();
}
int mexPrintf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  int r = vprintf(fmt, ap);
  va_end(ap);
  return r;
}
int mexEvalString(const char *cmd) {
  (void)cmd;
  fflush(stdout);
  return 0;
}
bool utIsInterruptPending(void) {
  g_interrupt_polls++;
  return g_interrupt_after >= 0 && g_interrupt_polls >= g_interrupt_after;
}

/* ---- driver helpers ---- */
mxArray *mxs_create_numeric(int classid, int ndim, const int64_t *dims, const void *data) {
  mwSize d[8];
  if (ndim > 8 || ndim < 1) fail("bad ndim");
  for (int k = 0; k < ndim; k++) d[k] = (mwSize)dims[k];
  mxArray *a = mxCreateNumericArray((mwSize)ndim, d, (mxClassID)classid, mxREAL);
  if (data) memcpy(a->data, data, a->nel * elem_size(a->cls));
  return a;
}
mxArray *mxs_create_object(void) {
  mxArray *a = alloc_array();
  a->cls = mxOBJECT_CLASS;
  a->ndim = 2;
  a->dims[0] = a->dims[1] = 1;
  a->nel = 1;
  return a;
}
mxArray *mxs_create_struct(void) {
  mxArray *a = mxs_create_object();
  a->cls = mxSTRUCT_CLASS;
  return a;
}
int mxs_set(mxArray *obj, const char *name, mxArray *value) {
  set_field(obj, name, value);
  return 0;
}
mxArray *mxs_get(const mxArray *obj, const char *name) { return find_field(obj, name); }
int mxs_class(const mxArray *pa) { return (int)pa->cls; }
int mxs_ndim(const mxArray *pa) { return (int)pa->ndim; }
int64_t mxs_dim(const mxArray *pa, int k) { return k < (int)pa->ndim ? (int64_t)pa->dims[k] : 1; }
void *mxs_data(const mxArray *pa) { return pa->data; }
int64_t mxs_numel(const mxArray *pa) { return (int64_t)pa->nel; }

int mxs_call(mxs_mexfunc_t fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
  jmp_buf jb;
  jmp_buf *prev = g_jmp;
  g_err[0] = 0;
  g_err_id[0] = 0;
  g_interrupt_polls = 0;
  g_jmp = &jb;
  if (setjmp(jb)) {
    g_jmp = prev;
    return 1;
  }
  fn(nlhs, plhs, nrhs, prhs);
  g_jmp = prev;
  fflush(stdout);
  return 0;
}
const char *mxs_last_error(void) { return g_err; }
const char *mxs_last_error_id(void) { return g_err_id; }
void mxs_set_interrupt_after(int64_t k) { g_interrupt_after = k; }
int64_t mxs_interrupt_polls(void) { return g_interrupt_polls; }
