/*
 * Gh = fhs_gmatrix(G)   — uploads a MATLAB sparse complex G (the G{i}{j} of
 * op_p_nufft, lib/operators/util_gen_measurement_operator_synth.m:85) to the
 * GPU and returns a handle struct (kind 'gmat') for apply_direct_operator /
 * apply_adjoint_operator / update_dual_data_fidelity.
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  const mxArray* G;
  const mwIndex *jc, *ir;
  int64_t M, ncols, nnz, e;
  int64_t *colptr, *rowidx;
  float* vals;
  fhs_gmat* h = NULL;
  const char* fields[] = {"fhs_handle", "fhs_kind"};
  mxArray* hv;
  (void)nlhs;
  if (nrhs != 1 || !mxIsSparse(prhs[0])) mexErrMsgIdAndTxt("fhs:args", "Gh = fhs_gmatrix(G_sparse)");
  G = prhs[0];
  M = (int64_t)mxGetM(G);
  ncols = (int64_t)mxGetN(G);
  jc = mxGetJc(G);
  ir = mxGetIr(G);
  nnz = (int64_t)jc[ncols];
  colptr = (int64_t*)mxMalloc((ncols + 1) * sizeof(int64_t));
  rowidx = (int64_t*)mxMalloc((nnz + 1) * sizeof(int64_t));
  vals = (float*)mxMalloc((2 * nnz + 2) * sizeof(float));
  for (e = 0; e <= ncols; ++e) colptr[e] = (int64_t)jc[e];
  if (mxIsComplex(G)) {
    const mxComplexDouble* z = mxGetComplexDoubles(G);
    for (e = 0; e < nnz; ++e) {
      rowidx[e] = (int64_t)ir[e];
      vals[2 * e] = (float)z[e].real;
      vals[2 * e + 1] = (float)z[e].imag;
    }
  } else {
    const double* r = mxGetDoubles(G);
    for (e = 0; e < nnz; ++e) {
      rowidx[e] = (int64_t)ir[e];
      vals[2 * e] = (float)r[e];
      vals[2 * e + 1] = 0.f;
    }
  }
  fhs_mex_check(fhs_gmat_create_csc_host(ctx, M, ncols, colptr, rowidx, vals, &h), "gmat_create_csc");
  mxFree(colptr);
  mxFree(rowidx);
  mxFree(vals);
  plhs[0] = mxCreateStructMatrix(1, 1, 2, fields);
  hv = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *(uint64_t*)mxGetData(hv) = (uint64_t)(uintptr_t)h;
  mxSetField(plhs[0], 0, "fhs_handle", hv);
  mxSetField(plhs[0], 0, "fhs_kind", mxCreateString("gmat"));
}
