/*
 * z = sdwt2_sara_faceting(x_curr, Iq, offset, status_q, nlevel, wavelet, Ncoefs_q)
 * (src/fhs/fhs_update_dual_sparsity.m:73,80, fhs_update_weights.m:88,
 * fhs_initialize_dual_and_weights.m:113, fhs_compute_facet_prior.m:82).
 * Drop-in over fhs_sara_facet_analysis: offset is the handle and status_q
 * the facet index from sdwt2_setup; x_curr the facet's extended image with
 * its zero padding (dims_overlap_ref + offsetL + offsetR); z the facet's
 * coefficients, a column.
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  fhs_sara_facets* F;
  int q, ext[4];
  int64_t n = 0, nx;
  float *dx, *dz = NULL, *h;
  int64_t i;
  (void)nlhs;
  if (nrhs != 7 || mxIsComplex(prhs[0]) || !mxGetDoubles(prhs[0]))
    mexErrMsgIdAndTxt("fhs:args", "z = sdwt2_sara_faceting(x_curr, Iq, offset, status_q, nlevel, wavelet, Ncoefs_q)");
  F = (fhs_sara_facets*)fhs_mex_handle(prhs[2], "sara_facets");
  q = (int)mxGetScalar(prhs[3]);
  if (fhs_sara_facets_info(F, q, ext, &n, NULL) != FHS_OK) mexErrMsgIdAndTxt("fhs:args", "bad facet index");
  if ((int)mxGetM(prhs[0]) != ext[2] || (int64_t)mxGetNumberOfElements(prhs[0]) != (int64_t)ext[2] * ext[3])
    mexErrMsgIdAndTxt("fhs:args", "x_curr must be %d x %d (dims_overlap_ref + offsetL + offsetR)", ext[2], ext[3]);
  dx = fhs_mex_to_device_real(prhs[0], &nx);
  fhs_mex_check(fhs_device_alloc(ctx, (n > 0 ? n : 1) * sizeof(float), (void**)&dz), "device_alloc");
  fhs_mex_check(fhs_sara_facet_analysis(ctx, F, q, dx, 1, dz), "sara_facet_analysis");
  plhs[0] = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
  h = (float*)mxMalloc((n > 0 ? n : 1) * sizeof(float));
  fhs_mex_check(fhs_memcpy_d2h(ctx, h, dz, n * sizeof(float)), "memcpy_d2h");
  for (i = 0; i < n; ++i) mxGetDoubles(plhs[0])[i] = h[i];
  mxFree(h);
  fhs_device_free(ctx, dx);
  fhs_device_free(ctx, dz);
}
