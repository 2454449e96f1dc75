/*
 * f = nuclear_norm(x)                            (src/nuclear_norm.m:25-26)
 * Drop-in MEX over fhs_nuclear_norm: x real Ny x Nx x L (L <= 16); the sum
 * of the singular values of reshape(x, Ny*Nx, L), FP64 Gram and Jacobi
 * eigen-decomposition on the GPU.
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  const mwSize* d;
  mwSize nd;
  int Ny, Nx, L;
  int64_t n;
  float* dx;
  double f = 0.0;
  (void)nlhs;
  if (nrhs != 1 || mxIsComplex(prhs[0])) mexErrMsgIdAndTxt("fhs:args", "f = nuclear_norm(x), x real");
  nd = mxGetNumberOfDimensions(prhs[0]);
  d = mxGetDimensions(prhs[0]);
  Ny = (int)d[0];
  Nx = (int)d[1];
  L = nd > 2 ? (int)d[2] : 1;
  dx = fhs_mex_to_device_real(prhs[0], &n);   /* column-major Ny x Nx x L = the library's cube layout */
  fhs_mex_check(fhs_nuclear_norm(ctx, dx, Ny, Nx, L, &f), "nuclear_norm");
  if (dx) fhs_device_free(ctx, dx);
  plhs[0] = mxCreateDoubleScalar(f);
}
