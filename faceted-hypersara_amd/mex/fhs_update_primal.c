/*
 * [xsol, xhat, rel_x, norm_x] = fhs_update_primal(xsol, g)
 *                                          (src/fhs/fhs_update_primal.m:1-40)
 * Drop-in MEX over fhs_update_primal: xsol = max(real(xsol - g), 0),
 * xhat = 2 xsol - xsol_prev, rel_x = sum (xsol_prev - xsol)^2,
 * norm_x = sum xsol^2 (FP64 sums); any array shape (a facet M x N x L).
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  int64_t n, ng, i;
  float *dx, *dg, *dh = NULL;
  double rel = 0.0, nrm = 0.0;
  (void)nlhs;
  if (nrhs != 2 || mxIsComplex(prhs[0]) || !mxGetDoubles(prhs[0]))
    mexErrMsgIdAndTxt("fhs:args", "[xsol, xhat, rel_x, norm_x] = fhs_update_primal(xsol, g), xsol real double");
  if (mxGetNumberOfElements(prhs[1]) != mxGetNumberOfElements(prhs[0]))
    mexErrMsgIdAndTxt("fhs:args", "g must have the size of xsol");
  if (mxIsComplex(prhs[1]))   /* max(real(xsol - g), 0): only the real part of g enters */
  {
    mxArray* gr = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[1]), mxGetDimensions(prhs[1]), mxDOUBLE_CLASS,
                                       mxREAL);
    const mxComplexDouble* z = mxGetComplexDoubles(prhs[1]);
    double* r = mxGetDoubles(gr);
    for (i = 0; i < (int64_t)mxGetNumberOfElements(prhs[1]); ++i) r[i] = z[i].real;
    dg = fhs_mex_to_device_real(gr, &ng);
  } else {
    dg = fhs_mex_to_device_real(prhs[1], &ng);
  }
  dx = fhs_mex_to_device_real(prhs[0], &n);
  if (n > 0) fhs_mex_check(fhs_device_alloc(ctx, n * sizeof(float), (void**)&dh), "device_alloc");
  if (n > 0) fhs_mex_check(fhs_update_primal(ctx, dx, dg, NULL, NULL, dh, n, &rel, &nrm), "update_primal");
  plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), mxGetDimensions(prhs[0]), mxDOUBLE_CLASS, mxREAL);
  plhs[1] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[0]), mxGetDimensions(prhs[0]), mxDOUBLE_CLASS, mxREAL);
  if (n > 0) {
    float* h = (float*)mxMalloc(n * sizeof(float));
    double* o = mxGetDoubles(plhs[0]);
    fhs_mex_check(fhs_memcpy_d2h(ctx, h, dx, n * sizeof(float)), "memcpy_d2h");
    for (i = 0; i < n; ++i) o[i] = h[i];
    o = mxGetDoubles(plhs[1]);
    fhs_mex_check(fhs_memcpy_d2h(ctx, h, dh, n * sizeof(float)), "memcpy_d2h");
    for (i = 0; i < n; ++i) o[i] = h[i];
    mxFree(h);
  }
  plhs[2] = mxCreateDoubleScalar(rel);
  plhs[3] = mxCreateDoubleScalar(nrm);
  if (dx) fhs_device_free(ctx, dx);
  if (dg) fhs_device_free(ctx, dg);
  if (dh) fhs_device_free(ctx, dh);
}
