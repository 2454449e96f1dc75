/*
 * Ph = fhs_sara_dictionary(Ny, Nx, nlevel, lo_r, with_dirac)
 * The handle (kind 'sara') that replaces the Psit / Psi closures of the SARA
 * dictionary (imaging/imaging.m:807-814): lo_r is a cell of the bases'
 * reconstruction low-pass filters (e.g. {wfilters('db1','r'), ...,
 * wfilters('db8','r')}), with_dirac adds the 'self' basis.  Coefficients
 * follow MATLAB's wavedec2 order with 'zpd' extension, scaled by
 * 1/sqrt(number of bases).
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  int Ny, Nx, nlevel, nb, b, total = 0, off = 0, with_dirac;
  int* flen;
  double* lo;
  fhs_sara* S = NULL;
  const char* fields[] = {"fhs_handle", "fhs_kind"};
  mxArray* hv;
  (void)nlhs;
  if (nrhs != 5 || !mxIsCell(prhs[3]))
    mexErrMsgIdAndTxt("fhs:args", "Ph = fhs_sara_dictionary(Ny, Nx, nlevel, lo_r_cell, with_dirac)");
  Ny = (int)mxGetScalar(prhs[0]);
  Nx = (int)mxGetScalar(prhs[1]);
  nlevel = (int)mxGetScalar(prhs[2]);
  with_dirac = mxGetScalar(prhs[4]) != 0.0;
  nb = (int)mxGetNumberOfElements(prhs[3]);
  flen = (int*)mxMalloc((nb + 1) * sizeof(int));
  for (b = 0; b < nb; ++b) {
    flen[b] = (int)mxGetNumberOfElements(mxGetCell(prhs[3], b));
    total += flen[b];
  }
  lo = (double*)mxMalloc((total + 1) * sizeof(double));
  for (b = 0; b < nb; ++b) {
    memcpy(lo + off, mxGetDoubles(mxGetCell(prhs[3], b)), flen[b] * sizeof(double));
    off += flen[b];
  }
  fhs_mex_check(fhs_sara_create_host(ctx, Ny, Nx, nlevel, nb, flen, lo, with_dirac, &S), "sara_create");
  mxFree(flen);
  mxFree(lo);
  plhs[0] = mxCreateStructMatrix(1, 1, 2, fields);
  hv = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *(uint64_t*)mxGetData(hv) = (uint64_t)(uintptr_t)S;
  mxSetField(plhs[0], 0, "fhs_handle", hv);
  mxSetField(plhs[0], 0, "fhs_kind", mxCreateString("sara"));
}
