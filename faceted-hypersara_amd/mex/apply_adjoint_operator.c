/*
 * Fx = apply_adjoint_operator(y, G, Lambda)      (src/apply_adjoint_operator.m:21)
 * Drop-in MEX over fhs_apply_adjoint_operator.  G is the struct returned by
 * fhs_gmatrix(G_sparse) (kind 'gmat'); Lambda may be [], a scalar or a real
 * vector of size(G, 1).
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  fhs_gmat* G;
  int64_t n, M, ncols, nnz, nl;
  void *dy, *dFx = NULL;
  float* dl;
  (void)nlhs;
  if (nrhs != 3) mexErrMsgIdAndTxt("fhs:args", "Fx = apply_adjoint_operator(y, G, Lambda)");
  G = (fhs_gmat*)fhs_mex_handle(prhs[1], "gmat");
  fhs_mex_check(fhs_gmat_info(G, &M, &ncols, &nnz), "gmat_info");
  dy = fhs_mex_to_device_complex(prhs[0], &n);
  if (n != M) mexErrMsgIdAndTxt("fhs:args", "y must have size(G, 1) entries");
  dl = fhs_mex_to_device_real(prhs[2], &nl);
  if (dl && nl != M && nl != 1) mexErrMsgIdAndTxt("fhs:args", "Lambda must be [], a scalar or size(G, 1) long");
  fhs_mex_check(fhs_device_alloc(ctx, 2 * ncols * sizeof(float), &dFx), "device_alloc");
  fhs_mex_check(fhs_apply_adjoint_operator(ctx, G, dy, nl == M ? dl : NULL, dFx), "apply_adjoint_operator");
  plhs[0] = fhs_mex_from_device_complex(dFx, ncols);
  if (dl && nl == 1) {   /* scalar Lambda */
    mxComplexDouble* z = mxGetComplexDoubles(plhs[0]);
    const double s = mxGetScalar(prhs[2]);
    int64_t i;
    for (i = 0; i < ncols; ++i) {
      z[i].real *= s;
      z[i].imag *= s;
    }
  }
  fhs_device_free(ctx, dy);
  fhs_device_free(ctx, dFx);
  if (dl) fhs_device_free(ctx, dl);
}
