/*
 * Shared helpers of the MATLAB MEX shims over libfhs (compiled only where
 * MATLAB's mex.h exists:  mex -R2018a -I../../include apply_direct_operator.c
 * -L../fhs -lfhs).
 *
 * One fhs_ctx per MATLAB process (MEX files share it through libfhs's
 * process-wide state when loaded from the same library); freed at exit.
 * Device handles (G, NUFFT plan, data state) travel through MATLAB as
 * structs with a uint64 field 'fhs_handle' and a 'fhs_kind' string; a
 * foreign or stale handle is rejected with an error.
 * mexErrMsgIdAndTxt longjmps: every temporary comes from mxMalloc (freed by
 * MATLAB on error) and no C++ object lives across it.
 */
#ifndef FHS_MEX_COMMON_H
#define FHS_MEX_COMMON_H

#include <stdint.h>
#include <string.h>

#include "fhs.h"
#include "mex.h"

static fhs_ctx* g_fhs_ctx = NULL;

static void fhs_mex_at_exit(void) {
  if (g_fhs_ctx) fhs_ctx_destroy(g_fhs_ctx);
  g_fhs_ctx = NULL;
}

static inline fhs_ctx* fhs_mex_ctx(void) {
  if (!g_fhs_ctx) {
    if (fhs_ctx_create(0, &g_fhs_ctx) != FHS_OK) mexErrMsgIdAndTxt("fhs:ctx", "cannot create the GPU context");
    mexAtExit(fhs_mex_at_exit);
    mexLock();
  }
  return g_fhs_ctx;
}

static inline void fhs_mex_check(int st, const char* what) {
  if (st != FHS_OK) mexErrMsgIdAndTxt("fhs:call", "%s: %s", what, fhs_last_error(g_fhs_ctx));
}

/* handle stored in a struct: s.fhs_handle (uint64), s.fhs_kind (char) */
static inline void* fhs_mex_handle(const mxArray* s, const char* kind) {
  const mxArray* h;
  const mxArray* k;
  char buf[32];
  if (!mxIsStruct(s)) mexErrMsgIdAndTxt("fhs:handle", "expected an fhs %s handle struct", kind);
  h = mxGetField(s, 0, "fhs_handle");
  k = mxGetField(s, 0, "fhs_kind");
  if (!h || !k || !mxIsUint64(h) || mxGetString(k, buf, sizeof(buf)) || strcmp(buf, kind))
    mexErrMsgIdAndTxt("fhs:handle", "not an fhs %s handle", kind);
  return (void*)(uintptr_t)(*(uint64_t*)mxGetData(h));
}

/* complex double column (MATLAB) -> device complex64 */
static inline void* fhs_mex_to_device_complex(const mxArray* a, int64_t* n_out) {
  const int64_t n = (int64_t)mxGetNumberOfElements(a);
  float* h = (float*)mxMalloc(2 * n * sizeof(float) + 8);
  void* d = NULL;
  int64_t i;
  if (mxIsComplex(a)) {
    const mxComplexDouble* z = mxGetComplexDoubles(a);
    for (i = 0; i < n; ++i) {
      h[2 * i] = (float)z[i].real;
      h[2 * i + 1] = (float)z[i].imag;
    }
  } else {
    const double* r = mxGetDoubles(a);
    for (i = 0; i < n; ++i) {
      h[2 * i] = (float)r[i];
      h[2 * i + 1] = 0.f;
    }
  }
  fhs_mex_check(fhs_device_alloc(g_fhs_ctx, 2 * n * sizeof(float), &d), "device_alloc");
  fhs_mex_check(fhs_memcpy_h2d(g_fhs_ctx, d, h, 2 * n * sizeof(float)), "memcpy_h2d");
  mxFree(h);
  *n_out = n;
  return d;
}

/* real double array -> device float32 (NULL for []) */
static inline float* fhs_mex_to_device_real(const mxArray* a, int64_t* n_out) {
  const int64_t n = (int64_t)mxGetNumberOfElements(a);
  float* h;
  void* d = NULL;
  const double* r;
  int64_t i;
  *n_out = n;
  if (n == 0) return NULL;
  h = (float*)mxMalloc(n * sizeof(float));
  r = mxGetDoubles(a);
  for (i = 0; i < n; ++i) h[i] = (float)r[i];
  fhs_mex_check(fhs_device_alloc(g_fhs_ctx, n * sizeof(float), &d), "device_alloc");
  fhs_mex_check(fhs_memcpy_h2d(g_fhs_ctx, d, h, n * sizeof(float)), "memcpy_h2d");
  mxFree(h);
  return (float*)d;
}

/* device complex64[n] -> new complex double column */
static inline mxArray* fhs_mex_from_device_complex(const void* d, int64_t n) {
  mxArray* out = mxCreateDoubleMatrix((mwSize)n, 1, mxCOMPLEX);
  mxComplexDouble* z = mxGetComplexDoubles(out);
  float* h = (float*)mxMalloc(2 * n * sizeof(float) + 8);
  int64_t i;
  fhs_mex_check(fhs_memcpy_d2h(g_fhs_ctx, h, d, 2 * n * sizeof(float)), "memcpy_d2h");
  for (i = 0; i < n; ++i) {
    z[i].real = h[2 * i];
    z[i].imag = h[2 * i + 1];
  }
  mxFree(h);
  return out;
}

#endif
