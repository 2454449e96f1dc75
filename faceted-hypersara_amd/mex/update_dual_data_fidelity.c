/*
 * [v2, Ftx, Fx_old, proj, norm_res, global_norm_res, norm_epsilon] =
 *   update_dual_data_fidelity(v2, y, x, Fx_old, proj, A, At, G, W, pU, epsilon,
 *     elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps, sigma22,
 *     flag_dimensionality_reduction, Lambda)   (src/update_dual_data_fidelity.m:1-149)
 * Drop-in MEX over fhs_update_dual_data_fidelity.
 *  - A is the fhs_data_state handle (kind 'data'); At, G and W are accepted
 *    and ignored (the state holds the operator).
 *  - v2, y, proj, pU, epsilon, Lambda: cells {L}{nblocks} as in the
 *    reference (pU, Lambda may be []).
 *  - Fx_old carries G A(x_old) instead of A(x_old) (by linearity
 *    G (2 A x - A x_old) = 2 G A x - G A x_old): a cell {L} of complex
 *    columns of the channel's row count, returned by the previous call; []
 *    (or zeros) for x_old = 0, the reference's initial state.  Callers only
 *    pass it back (src/fhs/facetHyperSARA.m:772-775).
 *  - norm_res is returned as a cell {L}{nblocks} of scalars, Ftx as the
 *    real Ny x Nx x L array sigma22 .* real(At(G' v2)).
 */
#include "fhs_mex_common.h"

/* rows of cell {L}{nb} of columns, and its per-block sizes */
static int64_t cell_rows(const mxArray* c, int64_t* sizes, int* nblk) {
  int64_t tot = 0;
  int l, b, k = 0;
  const int L = (int)mxGetNumberOfElements(c);
  for (l = 0; l < L; ++l) {
    const mxArray* cl = mxGetCell(c, l);
    const int nb = (int)mxGetNumberOfElements(cl);
    for (b = 0; b < nb; ++b) {
      const int64_t m = (int64_t)mxGetNumberOfElements(mxGetCell(cl, b));
      if (sizes) sizes[k] = m;
      ++k;
      tot += m;
    }
  }
  if (nblk) *nblk = k;
  return tot;
}

/* stacked complex64 host copy of a {L}{nb} cell of complex columns */
static float* stack_complex(const mxArray* c, int64_t M) {
  float* h = (float*)mxMalloc(2 * M * sizeof(float) + 8);
  int64_t k = 0, i;
  int l, b;
  for (l = 0; l < (int)mxGetNumberOfElements(c); ++l) {
    const mxArray* cl = mxGetCell(c, l);
    for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) {
      const mxArray* a = mxGetCell(cl, b);
      const int64_t m = (int64_t)mxGetNumberOfElements(a);
      if (mxIsComplex(a)) {
        const mxComplexDouble* z = mxGetComplexDoubles(a);
        for (i = 0; i < m; ++i, ++k) { h[2 * k] = (float)z[i].real; h[2 * k + 1] = (float)z[i].imag; }
      } else {
        const double* r = mxGetDoubles(a);
        for (i = 0; i < m; ++i, ++k) { h[2 * k] = (float)r[i]; h[2 * k + 1] = 0.f; }
      }
    }
  }
  return h;
}

static void* to_device(const void* h, size_t bytes) {
  void* d = NULL;
  fhs_mex_check(fhs_device_alloc(g_fhs_ctx, bytes, &d), "device_alloc");
  fhs_mex_check(fhs_memcpy_h2d(g_fhs_ctx, d, h, bytes), "memcpy_h2d");
  return d;
}

/* device complex64 stacked -> cell shaped like `like` ({L}{nb} columns) */
static mxArray* unstack_complex(const void* d, int64_t M, const mxArray* like) {
  float* h = (float*)mxMalloc(2 * M * sizeof(float) + 8);
  mxArray* out = mxCreateCellMatrix(mxGetM(like), mxGetN(like));
  int64_t k = 0, i;
  int l, b;
  fhs_mex_check(fhs_memcpy_d2h(g_fhs_ctx, h, d, 2 * M * sizeof(float)), "memcpy_d2h");
  for (l = 0; l < (int)mxGetNumberOfElements(like); ++l) {
    const mxArray* cl = mxGetCell(like, l);
    mxArray* ol = mxCreateCellMatrix(mxGetM(cl), mxGetN(cl));
    for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) {
      const int64_t m = (int64_t)mxGetNumberOfElements(mxGetCell(cl, b));
      mxArray* a = mxCreateDoubleMatrix((mwSize)m, 1, mxCOMPLEX);
      mxComplexDouble* z = mxGetComplexDoubles(a);
      for (i = 0; i < m; ++i, ++k) { z[i].real = h[2 * k]; z[i].imag = h[2 * k + 1]; }
      mxSetCell(ol, b, a);
    }
    mxSetCell(out, l, ol);
  }
  mxFree(h);
  return out;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  fhs_data* D;
  const mxArray *v2c, *yc, *xa, *fxo, *projc, *pUc, *epsc, *s22, *lamc;
  int64_t M, k, i;
  int nblk, l, b, L, dr;
  const mwSize* dims;
  int Ny, Nx, Lx;
  float *h, *hp, *s22f, *dx, *dFtx, *dpU = NULL, *dlam = NULL;
  double *eps, *nres, gres = 0.0, neps = 0.0;
  void *dv2, *dy, *dgx, *dproj;
  int64_t n;
  (void)nlhs;
  if (nrhs != 17) mexErrMsgIdAndTxt("fhs:args", "update_dual_data_fidelity: 17 arguments as src/update_dual_data_fidelity.m");
  v2c = prhs[0]; yc = prhs[1]; xa = prhs[2]; fxo = prhs[3]; projc = prhs[4];
  D = (fhs_data*)fhs_mex_handle(prhs[5], "data");
  pUc = prhs[9]; epsc = prhs[10]; s22 = prhs[14]; lamc = prhs[16];
  dr = mxGetScalar(prhs[15]) != 0.0;
  if (!mxIsCell(v2c) || !mxIsCell(yc) || !mxIsCell(epsc))
    mexErrMsgIdAndTxt("fhs:args", "v2, y and epsilon must be cells {L}{nblocks}");
  M = cell_rows(yc, NULL, &nblk);
  L = (int)mxGetNumberOfElements(yc);
  dims = mxGetDimensions(xa);
  Ny = (int)dims[0];
  Nx = (int)dims[1];
  Lx = mxGetNumberOfDimensions(xa) > 2 ? (int)dims[2] : 1;
  if (Lx != L) mexErrMsgIdAndTxt("fhs:args", "x must have one slice per channel of y");
  /* stacked device copies */
  h = stack_complex(v2c, M); dv2 = to_device(h, 2 * M * sizeof(float)); mxFree(h);
  h = stack_complex(yc, M); dy = to_device(h, 2 * M * sizeof(float)); mxFree(h);
  dproj = NULL;
  if (!dr) { h = stack_complex(projc, M); dproj = to_device(h, 2 * M * sizeof(float)); mxFree(h); }
  /* Gx_old: a cell {L} of columns from the previous call, or [] (zeros) */
  hp = (float*)mxMalloc(2 * M * sizeof(float) + 8);
  memset(hp, 0, 2 * M * sizeof(float));
  if (mxIsCell(fxo)) {
    k = 0;
    for (l = 0; l < (int)mxGetNumberOfElements(fxo); ++l) {
      const mxArray* a = mxGetCell(fxo, l);
      const int64_t m = (int64_t)mxGetNumberOfElements(a);
      if (!mxIsComplex(a) || k + m > M) mexErrMsgIdAndTxt("fhs:args", "Fx_old: the state returned by a previous call");
      {
        const mxComplexDouble* z = mxGetComplexDoubles(a);
        for (i = 0; i < m; ++i, ++k) { hp[2 * k] = (float)z[i].real; hp[2 * k + 1] = (float)z[i].imag; }
      }
    }
  }
  dgx = to_device(hp, 2 * M * sizeof(float));
  mxFree(hp);
  /* pU and Lambda: optional real cells */
  if (mxIsCell(pUc) && mxGetNumberOfElements(pUc) > 0) {
    float* hr = (float*)mxMalloc(M * sizeof(float) + 4);
    k = 0;
    for (l = 0; l < (int)mxGetNumberOfElements(pUc); ++l) {
      const mxArray* cl = mxGetCell(pUc, l);
      for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) {
        const mxArray* a = mxGetCell(cl, b);
        const double* r = mxGetDoubles(a);
        for (i = 0; i < (int64_t)mxGetNumberOfElements(a); ++i) hr[k++] = (float)r[i];
      }
    }
    dpU = (float*)to_device(hr, M * sizeof(float));
    mxFree(hr);
  }
  if (mxIsCell(lamc) && mxGetNumberOfElements(lamc) > 0) {
    float* hr = (float*)mxMalloc(M * sizeof(float) + 4);
    k = 0;
    for (l = 0; l < (int)mxGetNumberOfElements(lamc); ++l) {
      const mxArray* a = mxGetCell(lamc, l);
      const double* r = mxGetDoubles(a);
      for (i = 0; i < (int64_t)mxGetNumberOfElements(a); ++i) hr[k++] = (float)r[i];
    }
    dlam = (float*)to_device(hr, M * sizeof(float));
    mxFree(hr);
  }
  /* epsilon per block, sigma22 per channel, x */
  eps = (double*)mxMalloc((nblk + 1) * sizeof(double));
  nres = (double*)mxMalloc((nblk + 1) * sizeof(double));
  k = 0;
  for (l = 0; l < (int)mxGetNumberOfElements(epsc); ++l) {
    const mxArray* cl = mxGetCell(epsc, l);
    for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) eps[k++] = mxGetScalar(mxGetCell(cl, b));
  }
  s22f = (float*)mxMalloc((L + 1) * sizeof(float));
  for (l = 0; l < L; ++l) s22f[l] = (float)(mxGetNumberOfElements(s22) > 1 ? mxGetDoubles(s22)[l] : mxGetScalar(s22));
  dx = fhs_mex_to_device_real(xa, &n);
  fhs_mex_check(fhs_device_alloc(ctx, (size_t)Ny * Nx * L * sizeof(float), (void**)&dFtx), "device_alloc");
  fhs_mex_check(fhs_update_dual_data_fidelity(ctx, D, dv2, dy, dx, dgx, dproj, dpU, eps,
                                              (int)mxGetScalar(prhs[11]), (int)mxGetScalar(prhs[12]),
                                              mxGetScalar(prhs[13]), s22f, dr, dlam, dFtx, nres, &gres, &neps),
                "update_dual_data_fidelity");
  /* outputs */
  plhs[0] = unstack_complex(dv2, M, v2c);
  {
    const mwSize od[3] = {(mwSize)Ny, (mwSize)Nx, (mwSize)L};
    mxArray* F = mxCreateNumericArray(3, od, mxDOUBLE_CLASS, mxREAL);
    float* hf = (float*)mxMalloc((size_t)Ny * Nx * L * sizeof(float));
    double* fd = mxGetDoubles(F);
    fhs_mex_check(fhs_memcpy_d2h(ctx, hf, dFtx, (size_t)Ny * Nx * L * sizeof(float)), "memcpy_d2h");
    for (i = 0; i < (int64_t)Ny * Nx * L; ++i) fd[i] = hf[i];
    mxFree(hf);
    plhs[1] = F;
  }
  {
    /* the Gx_old state: one complex column per channel */
    float* hg = (float*)mxMalloc(2 * M * sizeof(float) + 8);
    mxArray* st = mxCreateCellMatrix(L, 1);
    fhs_mex_check(fhs_memcpy_d2h(ctx, hg, dgx, 2 * M * sizeof(float)), "memcpy_d2h");
    k = 0;
    for (l = 0; l < L; ++l) {
      const mxArray* cl = mxGetCell(yc, l);
      int64_t m = 0;
      mxArray* a;
      mxComplexDouble* z;
      for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) m += (int64_t)mxGetNumberOfElements(mxGetCell(cl, b));
      a = mxCreateDoubleMatrix((mwSize)m, 1, mxCOMPLEX);
      z = mxGetComplexDoubles(a);
      for (i = 0; i < m; ++i, ++k) { z[i].real = hg[2 * k]; z[i].imag = hg[2 * k + 1]; }
      mxSetCell(st, l, a);
    }
    mxFree(hg);
    plhs[2] = st;
  }
  plhs[3] = dr ? mxDuplicateArray(projc) : unstack_complex(dproj, M, projc);
  {
    mxArray* nr = mxCreateCellMatrix(mxGetM(yc), mxGetN(yc));
    k = 0;
    for (l = 0; l < L; ++l) {
      const mxArray* cl = mxGetCell(yc, l);
      mxArray* ol = mxCreateCellMatrix(mxGetM(cl), mxGetN(cl));
      for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) mxSetCell(ol, b, mxCreateDoubleScalar(nres[k++]));
      mxSetCell(nr, l, ol);
    }
    plhs[4] = nr;
  }
  plhs[5] = mxCreateDoubleScalar(gres);
  plhs[6] = mxCreateDoubleScalar(neps);
  fhs_device_free(ctx, dv2); fhs_device_free(ctx, dy); fhs_device_free(ctx, dgx);
  if (dproj) fhs_device_free(ctx, dproj);
  if (dpU) fhs_device_free(ctx, dpU);
  if (dlam) fhs_device_free(ctx, dlam);
  if (dx) fhs_device_free(ctx, dx);
  fhs_device_free(ctx, dFtx);
  mxFree(eps); mxFree(nres); mxFree(s22f);
}
