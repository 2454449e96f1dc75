/*
 * [v0, g0] = fhs_update_dual_lowrankness(v0, xhat, apodization_window,
 *                                        weights0, beta0)
 *                                 (src/fhs/fhs_update_dual_lowrankness.m:1-47)
 * Drop-in MEX over fhs_update_dual_lowrankness for one facet: xhat is the
 * facet M x N x c (c <= 16), v0 its dual [M*N, c], apodization_window
 * M x N (or [] = 1), weights0 c values in descending-singular-value order
 * (or a scalar), beta0 a scalar.  v0 <- Z - U max(S - beta0 weights0, 0) V'
 * with Z = v0 + w .* xhat; g0 = w .* reshape(v0, M, N, c).  The facet's
 * Gram matrix and eigen-decomposition are FP64 on the GPU.
 */
#include "fhs_mex_common.h"

static void download_real(fhs_ctx* ctx, const float* d, int64_t n, double* out) {
  float* h = (float*)mxMalloc(n * sizeof(float) + 4);
  int64_t i;
  fhs_mex_check(fhs_memcpy_d2h(ctx, h, d, n * sizeof(float)), "memcpy_d2h");
  for (i = 0; i < n; ++i) out[i] = h[i];
  mxFree(h);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  const mwSize* d;
  int M, N, c, q, zero = 0;
  int64_t n, nv, nw, P;
  float *dv, *dx, *dg = NULL, *hw = NULL;
  double* w0;
  fhs_lowrank* R = NULL;
  (void)nlhs;
  if (nrhs != 5 || mxIsComplex(prhs[1]) || !mxGetDoubles(prhs[1]) || mxIsComplex(prhs[0]) || !mxGetDoubles(prhs[0]))
    mexErrMsgIdAndTxt("fhs:args", "[v0, g0] = fhs_update_dual_lowrankness(v0, xhat, apodization_window, weights0, "
                                  "beta0), real double arrays");
  d = mxGetDimensions(prhs[1]);
  M = (int)d[0];
  N = (int)d[1];
  c = mxGetNumberOfDimensions(prhs[1]) > 2 ? (int)d[2] : 1;
  P = (int64_t)M * N;
  if (c < 1 || c > 16) mexErrMsgIdAndTxt("fhs:args", "the facet must have 1..16 channels");
  if ((int64_t)mxGetNumberOfElements(prhs[0]) != P * c) mexErrMsgIdAndTxt("fhs:args", "v0 must be [M*N, c]");
  nw = (int64_t)mxGetNumberOfElements(prhs[2]);
  if (nw != 0 && (nw != P || !mxGetDoubles(prhs[2])))
    mexErrMsgIdAndTxt("fhs:args", "apodization_window must be a real M x N array or []");
  if (mxGetNumberOfElements(prhs[3]) != 1 && (int)mxGetNumberOfElements(prhs[3]) != c)
    mexErrMsgIdAndTxt("fhs:args", "weights0 must have one value per channel (or be a scalar)");
  w0 = (double*)mxMalloc(c * sizeof(double));
  for (q = 0; q < c; ++q) w0[q] = mxGetNumberOfElements(prhs[3]) == 1 ? mxGetScalar(prhs[3]) : mxGetDoubles(prhs[3])[q];
  if (nw) {
    const double* r = mxGetDoubles(prhs[2]);
    int64_t i;
    hw = (float*)mxMalloc(P * sizeof(float));
    for (i = 0; i < P; ++i) hw[i] = (float)r[i];
  }
  /* one facet covering the M x N image: the library's v0 layout is the
     column-major [M*N, c] of the reference */
  dv = fhs_mex_to_device_real(prhs[0], &nv);
  dx = fhs_mex_to_device_real(prhs[1], &n);
  fhs_mex_check(fhs_device_alloc(ctx, n * sizeof(float), (void**)&dg), "device_alloc");
  fhs_mex_check(fhs_lowrank_create_host(ctx, M, N, c, 1, 1, &zero, &zero, &M, &N, hw, &R), "lowrank_create");
  if (fhs_update_dual_lowrankness(ctx, R, dv, dx, w0, mxGetScalar(prhs[4]), 1.0f, dg) != FHS_OK) {
    fhs_lowrank_destroy(R);
    fhs_mex_check(FHS_ERR_ARG, "update_dual_lowrankness");
  }
  fhs_lowrank_destroy(R);
  plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), mxGetDimensions(prhs[0]), mxDOUBLE_CLASS, mxREAL);
  download_real(ctx, dv, nv, mxGetDoubles(plhs[0]));
  plhs[1] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[1]), mxGetDimensions(prhs[1]), mxDOUBLE_CLASS, mxREAL);
  download_real(ctx, dg, n, mxGetDoubles(plhs[1]));
  fhs_device_free(ctx, dv);
  fhs_device_free(ctx, dx);
  fhs_device_free(ctx, dg);
}
