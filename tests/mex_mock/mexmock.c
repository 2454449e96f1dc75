/*
 * Test harness: implementation of the MEX API stand-in declared in mex.h,
 * plus mexmock_call (runs one mexFunction, catching mexErrMsgIdAndTxt) and
 * mexmock_run_at_exit.  Test infrastructure only.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
  mxClassID cls;
  int complex_, sparse;
  mwSize ndim;
  mwSize dims[8];
  void* data;          /* double / mxComplexDouble / uint64 / char */
  mxArray** cells;     /* cell elements */
  int nfields;         /* 1x1 struct */
  char** fnames;
  mxArray** fvals;
  mwIndex *ir, *jc;    /* sparse (CSC) */
  mwSize nzmax;
};

/* ---- temporaries (freed after every call, as MATLAB frees mxMalloc) ---- */
typedef struct Tmp {
  void* p;
  struct Tmp* next;
} Tmp;
static Tmp* g_tmp = NULL;

void* mxMalloc(size_t n) {
  Tmp* t = (Tmp*)malloc(sizeof(Tmp));
  t->p = malloc(n ? n : 1);
  t->next = g_tmp;
  g_tmp = t;
  return t->p;
}

void mxFree(void* p) {
  Tmp** q = &g_tmp;
  while (*q) {
    if ((*q)->p == p) {
      Tmp* t = *q;
      *q = t->next;
      free(t->p);
      free(t);
      return;
    }
    q = &(*q)->next;
  }
}

static void free_temporaries(void) {
  while (g_tmp) {
    Tmp* t = g_tmp;
    g_tmp = t->next;
    free(t->p);
    free(t);
  }
}

/* ---- arrays ------------------------------------------------------------ */
static size_t numel_dims(mwSize nd, const mwSize* d) {
  size_t n = 1;
  mwSize i;
  for (i = 0; i < nd; ++i) n *= d[i];
  return n;
}

static size_t elem_size(const mxArray* a) {
  switch (a->cls) {
    case mxDOUBLE_CLASS: return a->complex_ ? sizeof(mxComplexDouble) : sizeof(double);
    case mxUINT64_CLASS: return sizeof(uint64_t);
    case mxCHAR_CLASS: return 1;
    default: return 0;
  }
}

static mxArray* new_array(mxClassID cls, mwSize ndim, const mwSize* dims, int cplx) {
  mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
  mwSize i;
  a->cls = cls;
  a->complex_ = cplx;
  a->ndim = ndim < 2 ? 2 : ndim;
  a->dims[0] = a->dims[1] = 1;
  for (i = 0; i < ndim && i < 8; ++i) a->dims[i] = dims[i];
  if (cls == mxCELL_CLASS)
    a->cells = (mxArray**)calloc(numel_dims(a->ndim, a->dims) + 1, sizeof(mxArray*));
  else if (cls != mxSTRUCT_CLASS)
    a->data = calloc(numel_dims(a->ndim, a->dims) + 1, elem_size(a) ? elem_size(a) : 1);
  return a;
}

size_t mxGetNumberOfElements(const mxArray* a) {
  return a->sparse ? a->dims[0] * a->dims[1] : numel_dims(a->ndim, a->dims);
}
size_t mxGetM(const mxArray* a) { return a->dims[0]; }
size_t mxGetN(const mxArray* a) { return numel_dims(a->ndim, a->dims) / (a->dims[0] ? a->dims[0] : 1); }
void mxSetM(mxArray* a, size_t m) { a->dims[0] = m; }
void mxSetN(mxArray* a, size_t n) {
  a->ndim = 2;
  a->dims[1] = n;
}
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->ndim; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims; }
int mxIsCell(const mxArray* a) { return a && a->cls == mxCELL_CLASS; }
int mxIsStruct(const mxArray* a) { return a && a->cls == mxSTRUCT_CLASS; }
int mxIsComplex(const mxArray* a) { return a && a->complex_; }
int mxIsSparse(const mxArray* a) { return a && a->sparse; }
int mxIsUint64(const mxArray* a) { return a && a->cls == mxUINT64_CLASS; }

double mxGetScalar(const mxArray* a) {
  if (!a || (!a->sparse && mxGetNumberOfElements(a) == 0)) return 0.0;
  if (a->cls == mxDOUBLE_CLASS) return a->complex_ ? ((mxComplexDouble*)a->data)[0].real : ((double*)a->data)[0];
  if (a->cls == mxUINT64_CLASS) return (double)((uint64_t*)a->data)[0];
  return 0.0;
}
double* mxGetDoubles(const mxArray* a) { return a->cls == mxDOUBLE_CLASS && !a->complex_ ? (double*)a->data : NULL; }
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
  return a->cls == mxDOUBLE_CLASS && a->complex_ ? (mxComplexDouble*)a->data : NULL;
}
void* mxGetData(const mxArray* a) { return a->data; }
mwIndex* mxGetIr(const mxArray* a) { return a->ir; }
mwIndex* mxGetJc(const mxArray* a) { return a->jc; }

mxArray* mxGetCell(const mxArray* a, mwIndex i) { return a->cells[i]; }
void mxSetCell(mxArray* a, mwIndex i, mxArray* v) {
  if (a->cells[i]) mxDestroyArray(a->cells[i]);
  a->cells[i] = v;
}

mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name) {
  int f;
  (void)i;
  for (f = 0; f < a->nfields; ++f)
    if (!strcmp(a->fnames[f], name)) return a->fvals[f];
  return NULL;
}
void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* v) {
  int f;
  (void)i;
  for (f = 0; f < a->nfields; ++f)
    if (!strcmp(a->fnames[f], name)) {
      if (a->fvals[f]) mxDestroyArray(a->fvals[f]);
      a->fvals[f] = v;
      return;
    }
}

int mxGetString(const mxArray* a, char* buf, mwSize len) {
  size_t n;
  if (!a || a->cls != mxCHAR_CLASS || len == 0) return 1;
  n = mxGetNumberOfElements(a);
  if (n + 1 > len) {
    memcpy(buf, a->data, len - 1);
    buf[len - 1] = 0;
    return 1;
  }
  memcpy(buf, a->data, n);
  buf[n] = 0;
  return 0;
}

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c) {
  return new_array(cls, ndim, dims, c == mxCOMPLEX);
}
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c) {
  const mwSize d[2] = {m, n};
  return new_array(cls, 2, d, c == mxCOMPLEX);
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, c);
}
mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  ((double*)a->data)[0] = v;
  return a;
}
mxArray* mxCreateCellMatrix(mwSize m, mwSize n) {
  const mwSize d[2] = {m, n};
  return new_array(mxCELL_CLASS, 2, d, 0);
}
static char* dup_str(const char* s) {
  char* d = (char*)malloc(strlen(s) + 1);
  memcpy(d, s, strlen(s) + 1);
  return d;
}

mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names) {
  const mwSize d[2] = {m, n};
  mxArray* a = new_array(mxSTRUCT_CLASS, 2, d, 0);
  int f;
  a->nfields = nfields;
  a->fnames = (char**)calloc(nfields + 1, sizeof(char*));
  a->fvals = (mxArray**)calloc(nfields + 1, sizeof(mxArray*));
  for (f = 0; f < nfields; ++f) a->fnames[f] = dup_str(names[f]);
  return a;
}
mxArray* mxCreateString(const char* s) {
  const mwSize d[2] = {1, strlen(s)};
  mxArray* a = new_array(mxCHAR_CLASS, 2, d, 0);
  memcpy(a->data, s, strlen(s));
  return a;
}
mxArray* mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity c) {
  mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
  a->cls = mxDOUBLE_CLASS;
  a->complex_ = c == mxCOMPLEX;
  a->sparse = 1;
  a->ndim = 2;
  a->dims[0] = m;
  a->dims[1] = n;
  a->nzmax = nzmax ? nzmax : 1;
  a->data = calloc(a->nzmax, elem_size(a));
  a->ir = (mwIndex*)calloc(a->nzmax, sizeof(mwIndex));
  a->jc = (mwIndex*)calloc(n + 1, sizeof(mwIndex));
  return a;
}

mxArray* mxDuplicateArray(const mxArray* a) {
  mxArray* b;
  size_t i, n;
  if (!a) return NULL;
  if (a->sparse) {
    b = mxCreateSparse(a->dims[0], a->dims[1], a->nzmax, a->complex_ ? mxCOMPLEX : mxREAL);
    memcpy(b->data, a->data, a->nzmax * elem_size(a));
    memcpy(b->ir, a->ir, a->nzmax * sizeof(mwIndex));
    memcpy(b->jc, a->jc, (a->dims[1] + 1) * sizeof(mwIndex));
    return b;
  }
  if (a->cls == mxSTRUCT_CLASS) {
    int f;
    b = mxCreateStructMatrix(a->dims[0], a->dims[1], a->nfields, (const char**)a->fnames);
    for (f = 0; f < a->nfields; ++f) b->fvals[f] = mxDuplicateArray(a->fvals[f]);
    return b;
  }
  b = new_array(a->cls, a->ndim, a->dims, a->complex_);
  n = numel_dims(a->ndim, a->dims);
  if (a->cls == mxCELL_CLASS)
    for (i = 0; i < n; ++i) b->cells[i] = mxDuplicateArray(a->cells[i]);
  else
    memcpy(b->data, a->data, n * elem_size(a));
  return b;
}

void mxDestroyArray(mxArray* a) {
  size_t i;
  if (!a) return;
  if (a->cls == mxCELL_CLASS && a->cells)
    for (i = 0; i < numel_dims(a->ndim, a->dims); ++i) mxDestroyArray(a->cells[i]);
  if (a->cls == mxSTRUCT_CLASS) {
    int f;
    for (f = 0; f < a->nfields; ++f) {
      mxDestroyArray(a->fvals[f]);
      free(a->fnames[f]);
    }
    free(a->fnames);
    free(a->fvals);
  }
  free(a->cells);
  free(a->data);
  free(a->ir);
  free(a->jc);
  free(a);
}

/* ---- errors, calls, exit hooks ---------------------------------------- */
static jmp_buf g_jmp;
static int g_in_call = 0;
static char g_msg[1024];
static void (*g_at_exit[64])(void);
static int g_n_at_exit = 0;

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  va_list ap;
  int k = snprintf(g_msg, sizeof(g_msg), "%s: ", id);
  va_start(ap, fmt);
  vsnprintf(g_msg + k, sizeof(g_msg) - k, fmt, ap);
  va_end(ap);
  if (g_in_call) longjmp(g_jmp, 1);
  fprintf(stderr, "mexErrMsgIdAndTxt outside a call: %s\n", g_msg);
  abort();
}

int mexAtExit(void (*fn)(void)) {
  int i;
  for (i = 0; i < g_n_at_exit; ++i)
    if (g_at_exit[i] == fn) return 0;
  if (g_n_at_exit < 64) g_at_exit[g_n_at_exit++] = fn;
  return 0;
}

void mexLock(void) {}

/* mexCallMATLAB: answered by a function the test harness registers (it
   plays the MATLAB functions a shim calls back, e.g. wfilters) */
typedef int (*mexmock_matlab_fn)(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs, const char* name);
static mexmock_matlab_fn g_matlab = NULL;

void mexmock_set_matlab(mexmock_matlab_fn f) { g_matlab = f; }

int mexCallMATLAB(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name) {
  if (!g_matlab || g_matlab(nlhs, plhs, nrhs, prhs, name) != 0)
    mexErrMsgIdAndTxt("mexmock:call", "mexCallMATLAB(%s) is not available", name);
  return 0;
}

typedef void (*mex_fn)(int, mxArray**, int, const mxArray**);

/* Runs fn like MATLAB: returns 0, or 1 with the error text in err. */
int mexmock_call(mex_fn fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* err, int errlen) {
  volatile int failed = 0;
  g_msg[0] = 0;
  g_in_call = 1;
  if (setjmp(g_jmp) == 0)
    fn(nlhs, plhs, nrhs, prhs);
  else
    failed = 1;
  g_in_call = 0;
  free_temporaries();
  if (failed && err && errlen > 0) snprintf(err, errlen, "%s", g_msg);
  return failed;
}

void mexmock_run_at_exit(void) {
  while (g_n_at_exit > 0) g_at_exit[--g_n_at_exit]();
}
