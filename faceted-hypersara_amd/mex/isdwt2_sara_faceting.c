/*
 * g = isdwt2_sara_faceting(z, Iq, dims_q, I_overlap_q, dims_overlap_q,
 *                          Ncoefs_q, nlevel, wavelet, temLIdxs_q, temRIdxs_q)
 * (src/fhs/fhs_update_dual_sparsity.m:92).  Drop-in over
 * fhs_sara_facet_synthesis: temLIdxs_q is the handle with the facet index
 * from sdwt2_setup; g the facet's contribution over its in-image extended
 * region (dims_overlap_q: the zero-padding rows / columns at the image
 * edge dropped), which the solver's border reduction adds up over facets.
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  fhs_sara_facets* F;
  const mxArray* fq;
  int q, ext[4], py, px, r, c;
  int64_t n = 0, nz;
  float *dz, *dg = NULL, *h;
  double* o;
  (void)nlhs;
  if (nrhs != 10 || mxIsComplex(prhs[0]) || !mxGetDoubles(prhs[0]))
    mexErrMsgIdAndTxt("fhs:args", "g = isdwt2_sara_faceting(z, Iq, dims_q, I_overlap_q, dims_overlap_q, Ncoefs_q, "
                                  "nlevel, wavelet, temLIdxs_q, temRIdxs_q)");
  F = (fhs_sara_facets*)fhs_mex_handle(prhs[8], "sara_facets");
  fq = mxGetField(prhs[8], 0, "fhs_facet");
  if (!fq) mexErrMsgIdAndTxt("fhs:handle", "temLIdxs_q carries no facet index");
  q = (int)mxGetScalar(fq);
  if (fhs_sara_facets_info(F, q, ext, &n, NULL) != FHS_OK) mexErrMsgIdAndTxt("fhs:args", "bad facet index");
  if ((int64_t)mxGetNumberOfElements(prhs[0]) != n)
    mexErrMsgIdAndTxt("fhs:args", "z must hold the facet's %lld coefficients", (long long)n);
  dz = fhs_mex_to_device_real(prhs[0], &nz);
  fhs_mex_check(fhs_device_alloc(ctx, (size_t)ext[2] * ext[3] * sizeof(float), (void**)&dg), "device_alloc");
  fhs_mex_check(fhs_sara_facet_synthesis(ctx, F, q, dz, 1, 1.0f, dg), "sara_facet_synthesis");
  h = (float*)mxMalloc((size_t)ext[2] * ext[3] * sizeof(float));
  fhs_mex_check(fhs_memcpy_d2h(ctx, h, dg, (size_t)ext[2] * ext[3] * sizeof(float)), "memcpy_d2h");
  py = ext[0] < 0 ? -ext[0] : 0;
  px = ext[1] < 0 ? -ext[1] : 0;
  plhs[0] = mxCreateDoubleMatrix((mwSize)(ext[2] - py), (mwSize)(ext[3] - px), mxREAL);
  o = mxGetDoubles(plhs[0]);
  for (c = px; c < ext[3]; ++c)
    for (r = py; r < ext[2]; ++r) o[(size_t)(r - py) + (size_t)(ext[2] - py) * (c - px)] = h[r + (size_t)ext[2] * c];
  mxFree(h);
  if (dz) fhs_device_free(ctx, dz);
  fhs_device_free(ctx, dg);
}
