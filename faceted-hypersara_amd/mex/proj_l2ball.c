/*
 * p = proj_l2ball(x, eps, y)                      (src/proj_l2ball.m)
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  int64_t n, ny;
  void *dx, *dy, *dp = NULL;
  mxArray* out;
  (void)nlhs;
  if (nrhs != 3) mexErrMsgIdAndTxt("fhs:args", "p = proj_l2ball(x, eps, y)");
  dx = fhs_mex_to_device_complex(prhs[0], &n);
  dy = fhs_mex_to_device_complex(prhs[2], &ny);
  if (n != ny) mexErrMsgIdAndTxt("fhs:args", "x and y must have the same number of elements");
  fhs_mex_check(fhs_device_alloc(ctx, 2 * n * sizeof(float), &dp), "device_alloc");
  fhs_mex_check(fhs_proj_l2ball(ctx, dx, mxGetScalar(prhs[1]), dy, n, dp), "proj_l2ball");
  out = fhs_mex_from_device_complex(dp, n);
  mxSetM(out, mxGetM(prhs[0]));
  mxSetN(out, mxGetN(prhs[0]));
  plhs[0] = out;
  fhs_device_free(ctx, dx);
  fhs_device_free(ctx, dy);
  fhs_device_free(ctx, dp);
}
