/*
 * y = apply_direct_operator(Fx, G, Lambda)       (src/apply_direct_operator.m)
 * Drop-in MEX: placed next to (or ahead of) the .m on the MATLAB path it
 * shadows it.  G is the struct returned by fhs_gmatrix(G_sparse) (device
 * handle, kind 'gmat'); Lambda may be [] or a real vector.
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  fhs_gmat* G;
  int64_t n, M, ncols, nnz, nl;
  void *dFx, *dy = NULL;
  float* dl;
  (void)nlhs;
  if (nrhs != 3) mexErrMsgIdAndTxt("fhs:args", "y = apply_direct_operator(Fx, G, Lambda)");
  G = (fhs_gmat*)fhs_mex_handle(prhs[1], "gmat");
  fhs_mex_check(fhs_gmat_info(G, &M, &ncols, &nnz), "gmat_info");
  dFx = fhs_mex_to_device_complex(prhs[0], &n);
  if (n != ncols) mexErrMsgIdAndTxt("fhs:args", "Fx must have size(G, 2) entries");
  dl = fhs_mex_to_device_real(prhs[2], &nl);
  if (dl && nl != M && nl != 1) mexErrMsgIdAndTxt("fhs:args", "Lambda must be [] or have size(G, 1) entries");
  fhs_mex_check(fhs_device_alloc(ctx, 2 * M * sizeof(float), &dy), "device_alloc");
  fhs_mex_check(fhs_apply_direct_operator(ctx, G, dFx, nl == M ? dl : NULL, dy), "apply_direct_operator");
  plhs[0] = fhs_mex_from_device_complex(dy, M);
  if (dl && nl == 1) {   /* scalar Lambda */
    mxComplexDouble* z = mxGetComplexDoubles(plhs[0]);
    const double s = mxGetScalar(prhs[2]);
    int64_t i;
    for (i = 0; i < M; ++i) {
      z[i].real *= s;
      z[i].imag *= s;
    }
  }
  fhs_device_free(ctx, dFx);
  fhs_device_free(ctx, dy);
  if (dl) fhs_device_free(ctx, dl);
}
