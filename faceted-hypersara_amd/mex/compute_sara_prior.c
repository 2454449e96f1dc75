/*
 * f = compute_sara_prior(x, Psit, n_wavelet_coefficients)
 *                                             (src/compute_sara_prior.m:33-41)
 * Drop-in MEX over fhs_compute_sara_prior: Psit is the fhs_sara_dictionary
 * handle (kind 'sara'); the third argument is kept for the signature (the
 * dictionary knows its coefficient count).  f = sum over coefficients of
 * the l2 norm across channels (FP64).
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  const mwSize* d;
  mwSize nd;
  int L;
  int64_t n;
  float* dx;
  double f = 0.0;
  fhs_sara* S;
  (void)nlhs;
  if (nrhs != 3) mexErrMsgIdAndTxt("fhs:args", "f = compute_sara_prior(x, Psit, n_wavelet_coefficients)");
  S = (fhs_sara*)fhs_mex_handle(prhs[1], "sara");
  nd = mxGetNumberOfDimensions(prhs[0]);
  d = mxGetDimensions(prhs[0]);
  L = nd > 2 ? (int)d[2] : 1;
  dx = fhs_mex_to_device_real(prhs[0], &n);
  fhs_mex_check(fhs_compute_sara_prior(ctx, S, dx, L, &f), "compute_sara_prior");
  if (dx) fhs_device_free(ctx, dx);
  plhs[0] = mxCreateDoubleScalar(f);
}
