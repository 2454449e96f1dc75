/*
 * Fx = apply_scaled_fourier_transform(x, A, No)
 *                                  (src/apply_scaled_fourier_transform.m:1-26)
 * Drop-in MEX over fhs_apply_scaled_fourier_transform: Fx(:, l) = A(x(:, :, l))
 * for every channel, the zero-padded scaled FFT on the GPU.  A is the
 * fhs_data_state handle (kind 'data': its scaled FFT) or an NUFFT handle
 * (kind 'nufft'); x real Ny x Nx x L; returns the complex [No, L] array
 * (No = Kty*Ktx) that update_dual_data_fidelity takes as Fx_old.  As in
 * src/hs/hyperSARA.m:446, where A is only used through this call,
 * update_dual_data_fidelity and compute_residual_images.
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  const fhs_nufft* plan = NULL;
  const mwSize* d;
  const mxArray* k;
  char kind[32];
  int Ny, Nx, L, Kty = 0, Ktx = 0;
  int64_t n, No;
  float* dx;
  void* dF = NULL;
  (void)nlhs;
  if (nrhs != 3 || mxIsComplex(prhs[0]) || !mxGetDoubles(prhs[0]))
    mexErrMsgIdAndTxt("fhs:args", "Fx = apply_scaled_fourier_transform(x, A, No), x real double");
  if (!mxIsStruct(prhs[1]) || !(k = mxGetField(prhs[1], 0, "fhs_kind")) || mxGetString(k, kind, sizeof(kind)))
    mexErrMsgIdAndTxt("fhs:handle", "A must be an fhs data or nufft handle");
  if (!strcmp(kind, "data"))
    fhs_mex_check(fhs_data_plan((const fhs_data*)fhs_mex_handle(prhs[1], "data"), &plan), "data_plan");
  else
    plan = (const fhs_nufft*)fhs_mex_handle(prhs[1], "nufft");
  fhs_mex_check(fhs_nufft_info(plan, NULL, NULL, &Kty, &Ktx), "nufft_info");
  No = (int64_t)Kty * Ktx;
  if ((int64_t)mxGetScalar(prhs[2]) != No) mexErrMsgIdAndTxt("fhs:args", "No must be %lld (Kty*Ktx)", (long long)No);
  d = mxGetDimensions(prhs[0]);
  Ny = (int)d[0];
  Nx = (int)d[1];
  L = mxGetNumberOfDimensions(prhs[0]) > 2 ? (int)d[2] : 1;
  {
    int pNy, pNx;
    fhs_mex_check(fhs_nufft_info(plan, &pNy, &pNx, NULL, NULL), "nufft_info");
    if (pNy != Ny || pNx != Nx) mexErrMsgIdAndTxt("fhs:args", "x must be %d x %d x L", pNy, pNx);
  }
  dx = fhs_mex_to_device_real(prhs[0], &n);
  fhs_mex_check(fhs_device_alloc(ctx, No * L * 2 * sizeof(float), &dF), "device_alloc");
  fhs_mex_check(fhs_apply_scaled_fourier_transform(ctx, plan, dx, L, dF), "apply_scaled_fourier_transform");
  {
    mxArray* out = mxCreateDoubleMatrix((mwSize)No, (mwSize)L, mxCOMPLEX);
    mxComplexDouble* z = mxGetComplexDoubles(out);
    float* h = (float*)mxMalloc(No * L * 2 * sizeof(float));
    int64_t i;
    fhs_mex_check(fhs_memcpy_d2h(ctx, h, dF, No * L * 2 * sizeof(float)), "memcpy_d2h");
    for (i = 0; i < No * L; ++i) {
      z[i].real = h[2 * i];
      z[i].imag = h[2 * i + 1];
    }
    mxFree(h);
    plhs[0] = out;
  }
  if (dx) fhs_device_free(ctx, dx);
  fhs_device_free(ctx, dF);
}
