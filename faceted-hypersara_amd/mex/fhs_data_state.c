/*
 * Dh = fhs_data_state(sy, sx, Kty, Ktx, G, block_sizes[, row_bits])
 * The device state (kind 'data') that update_dual_data_fidelity takes in
 * place of the A, At, G, W arguments of src/update_dual_data_fidelity.m:
 * sy (Ny), sx (Nx) are the NUFFT scaling factors of A (op_p_nufft), Kty x
 * Ktx the oversampled grid, G{l} the fhs_gmatrix handle of channel l with
 * the rows of its data blocks stacked in order (vertcat(G{l}{:}) of the
 * reference's per-block matrices restricted by W), block_sizes{l} the row
 * counts of those blocks.  row_bits (optional, 32 or 16) is the storage
 * precision of the G rows the state's kernels read (fhs_data_set_row_precision:
 * 16 keeps an FP16 copy and needs tile-gridded separable G on every channel).
 */
#include "fhs_mex_common.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  int L, l, b, nblocks = 0, k = 0;
  int64_t* bstart;
  int* bchan;
  const fhs_gmat** G;
  fhs_nufft* A = NULL;
  fhs_data* D = NULL;
  const char* fields[] = {"fhs_handle", "fhs_kind"};
  mxArray* hv;
  (void)nlhs;
  if ((nrhs != 6 && nrhs != 7) || !mxIsCell(prhs[4]) || !mxIsCell(prhs[5]))
    mexErrMsgIdAndTxt("fhs:args", "Dh = fhs_data_state(sy, sx, Kty, Ktx, G_cell, block_sizes_cell[, row_bits])");
  L = (int)mxGetNumberOfElements(prhs[4]);
  fhs_mex_check(fhs_nufft_create_host(ctx, (int)mxGetNumberOfElements(prhs[0]), (int)mxGetNumberOfElements(prhs[1]),
                                      (int)mxGetScalar(prhs[2]), (int)mxGetScalar(prhs[3]), mxGetDoubles(prhs[0]),
                                      mxGetDoubles(prhs[1]), &A),
                "nufft_create");
  for (l = 0; l < L; ++l) nblocks += (int)mxGetNumberOfElements(mxGetCell(prhs[5], l));
  bstart = (int64_t*)mxMalloc((nblocks + 1) * sizeof(int64_t));
  bchan = (int*)mxMalloc((nblocks + 1) * sizeof(int));
  G = (const fhs_gmat**)mxMalloc((L + 1) * sizeof(fhs_gmat*));
  bstart[0] = 0;
  for (l = 0; l < L; ++l) {
    const mxArray* bs = mxGetCell(prhs[5], l);
    const double* r = mxGetDoubles(bs);
    const int nb = (int)mxGetNumberOfElements(bs);
    G[l] = (const fhs_gmat*)fhs_mex_handle(mxGetCell(prhs[4], l), "gmat");
    for (b = 0; b < nb; ++b, ++k) {
      bstart[k + 1] = bstart[k] + (int64_t)r[b];
      bchan[k] = l;
    }
  }
  fhs_mex_check(fhs_data_create(ctx, A, L, G, nblocks, bstart, bchan, &D), "data_create");
  mxFree(bstart);
  mxFree(bchan);
  mxFree((void*)G);
  if (nrhs == 7) {
    const int rc = fhs_data_set_row_precision(ctx, D, (int)mxGetScalar(prhs[6]));
    if (rc != FHS_OK) fhs_data_destroy(D);
    fhs_mex_check(rc, "data_set_row_precision");
  }
  plhs[0] = mxCreateStructMatrix(1, 1, 2, fields);
  hv = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *(uint64_t*)mxGetData(hv) = (uint64_t)(uintptr_t)D;
  mxSetField(plhs[0], 0, "fhs_handle", hv);
  mxSetField(plhs[0], 0, "fhs_kind", mxCreateString("data"));
}
