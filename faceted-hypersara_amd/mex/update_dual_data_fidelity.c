/*
 * [v2, Ftx, Fx_old, proj, norm_res, global_norm_res, norm_epsilon] =
 *   update_dual_data_fidelity(v2, y, x, Fx_old, proj, A, At, G, W, pU, epsilon,
 *     elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps, sigma22,
 *     flag_dimensionality_reduction, Lambda)   (src/update_dual_data_fidelity.m:1-149)
 * Drop-in MEX over fhs_update_dual_data_fidelity.
 *  - The device operator is the fhs_data_state handle (kind 'data'), passed
 *    in the G slot (or in the A slot).  A stays the reference's callable
 *    function handle, as the solvers call A(xi) themselves
 *    (src/fhs/facetHyperSARA.m:574-576); At, W and the other slot are
 *    accepted and ignored (the state holds the operator).
 *  - v2, y, proj, pU, epsilon, Lambda: cells {L}{nblocks} as in the
 *    reference; every block of v2, proj (standard form), pU and Lambda must
 *    have the row count of the same block of y.  pU and Lambda may be [].
 *  - Fx_old, input: the reference's own [No, L] array A(x_old)
 *    (src/update_dual_data_fidelity.m:92,132; G Fx_old is formed on the
 *    device), [] or zeros for x_old = 0, or the cell returned by a previous
 *    call.  Output: a cell {L} of complex columns holding G_l A(x) (the rows
 *    of channel l), which the next call takes back: by linearity
 *    G (2 A x - A x_old) = 2 G A x - G A x_old, so M values per channel
 *    travel instead of No.  The solvers only pass it back
 *    (src/fhs/facetHyperSARA.m:772-775).
 *  - norm_res is returned as a cell {L}{nblocks} of scalars, Ftx as the
 *    real Ny x Nx x L array sigma22 .* real(At(G' v2)).
 */
#include "fhs_mex_common.h"

/* block structure of the {L}{nb} data cell y */
typedef struct {
  int L, nblk;
  int64_t M;
  int64_t* size;       /* rows of block k (blocks in channel order) */
  int64_t* chan_rows;  /* rows of channel l */
} Blocks;

static Blocks blocks_of(const mxArray* yc, int L_expected) {
  Blocks B;
  int l, b, k = 0;
  if (!mxIsCell(yc)) mexErrMsgIdAndTxt("fhs:args", "y must be a cell {L}{nblocks}");
  B.L = (int)mxGetNumberOfElements(yc);
  if (B.L != L_expected) mexErrMsgIdAndTxt("fhs:args", "y has %d channels, the data state %d", B.L, L_expected);
  B.nblk = 0;
  for (l = 0; l < B.L; ++l) {
    const mxArray* cl = mxGetCell(yc, l);
    if (!cl || !mxIsCell(cl)) mexErrMsgIdAndTxt("fhs:args", "y{%d} must be a cell of blocks", l + 1);
    B.nblk += (int)mxGetNumberOfElements(cl);
  }
  B.size = (int64_t*)mxMalloc((B.nblk + 1) * sizeof(int64_t));
  B.chan_rows = (int64_t*)mxMalloc((B.L + 1) * sizeof(int64_t));
  B.M = 0;
  for (l = 0; l < B.L; ++l) {
    const mxArray* cl = mxGetCell(yc, l);
    B.chan_rows[l] = 0;
    for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b, ++k) {
      const mxArray* a = mxGetCell(cl, b);
      if (!a || mxIsCell(a) || mxIsStruct(a)) mexErrMsgIdAndTxt("fhs:args", "y{%d}{%d} must be numeric", l + 1, b + 1);
      B.size[k] = (int64_t)mxGetNumberOfElements(a);
      B.chan_rows[l] += B.size[k];
      B.M += B.size[k];
    }
  }
  return B;
}

/* c must be a {L}{nb} cell whose block k has B->size[k] numeric elements */
static void check_like(const mxArray* c, const Blocks* B, const char* name) {
  int l, b, k = 0;
  if (!mxIsCell(c) || (int)mxGetNumberOfElements(c) != B->L)
    mexErrMsgIdAndTxt("fhs:args", "%s must be a cell {L}{nblocks} shaped like y", name);
  for (l = 0; l < B->L; ++l) {
    const mxArray* cl = mxGetCell(c, l);
    if (!cl || !mxIsCell(cl)) mexErrMsgIdAndTxt("fhs:args", "%s{%d} must be a cell of blocks", name, l + 1);
    for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b, ++k) {
      const mxArray* a = mxGetCell(cl, b);
      if (k >= B->nblk || !a || mxIsCell(a) || mxIsStruct(a) || (int64_t)mxGetNumberOfElements(a) != B->size[k])
        mexErrMsgIdAndTxt("fhs:args", "%s{%d}{%d} must have the rows of y{%d}{%d}", name, l + 1, b + 1, l + 1,
                          b + 1);
    }
  }
  if (k != B->nblk) mexErrMsgIdAndTxt("fhs:args", "%s must have the blocks of y", name);
}

/* stacked complex64 host copy of a {L}{nb} cell (checked by check_like) */
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

/* stacked float32 host copy of a real {L}{nb} cell (checked by check_like) */
static float* stack_real(const mxArray* c, int64_t M, const char* name) {
  float* h = (float*)mxMalloc(M * sizeof(float) + 4);
  int64_t k = 0, i;
  int l, b;
  for (l = 0; l < (int)mxGetNumberOfElements(c); ++l) {
    const mxArray* cl = mxGetCell(c, l);
    for (b = 0; b < (int)mxGetNumberOfElements(cl); ++b) {
      const mxArray* a = mxGetCell(cl, b);
      const double* r = mxGetDoubles(a);
      if (!r) mexErrMsgIdAndTxt("fhs:args", "%s{%d}{%d} must be real double", name, l + 1, b + 1);
      for (i = 0; i < (int64_t)mxGetNumberOfElements(a); ++i) h[k++] = (float)r[i];
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

/* the data state from the G slot, else from the A slot */
static fhs_data* data_handle(const mxArray* G, const mxArray* A) {
  char buf[32];
  const mxArray* k;
  if (mxIsStruct(G) && (k = mxGetField(G, 0, "fhs_kind")) && !mxGetString(k, buf, sizeof(buf)) && !strcmp(buf, "data"))
    return (fhs_data*)fhs_mex_handle(G, "data");
  return (fhs_data*)fhs_mex_handle(A, "data");
}

static int all_zero(const mxArray* a) {
  const int64_t n = (int64_t)mxGetNumberOfElements(a);
  int64_t i;
  if (mxIsComplex(a)) {
    const mxComplexDouble* z = mxGetComplexDoubles(a);
    for (i = 0; i < n; ++i)
      if (z[i].real != 0.0 || z[i].imag != 0.0) return 0;
  } else {
    const double* r = mxGetDoubles(a);
    if (!r) mexErrMsgIdAndTxt("fhs:args", "Fx_old must be double");
    for (i = 0; i < n; ++i)
      if (r[i] != 0.0) return 0;
  }
  return 1;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  fhs_data* D;
  const mxArray *v2c, *yc, *xa, *fxo, *projc, *pUc, *epsc, *s22, *lamc;
  Blocks B;
  int64_t k, i, No_d, M_d;
  int l, b, L, dr, Ny_d, Nx_d, L_d, nb_d;
  const mwSize* dims;
  int Ny, Nx, Lx;
  float *h, *s22f, *dx, *dFtx, *dpU = NULL, *dlam = NULL;
  double *eps, *nres, gres = 0.0, neps = 0.0;
  void *dv2, *dy, *dgx, *dproj;
  int64_t n;
  (void)nlhs;
  if (nrhs != 17) mexErrMsgIdAndTxt("fhs:args", "update_dual_data_fidelity: 17 arguments as src/update_dual_data_fidelity.m");
  v2c = prhs[0]; yc = prhs[1]; xa = prhs[2]; fxo = prhs[3]; projc = prhs[4];
  D = data_handle(prhs[7], prhs[5]);
  pUc = prhs[9]; epsc = prhs[10]; s22 = prhs[14]; lamc = prhs[16];
  dr = mxGetScalar(prhs[15]) != 0.0;
  fhs_mex_check(fhs_data_info(D, &Ny_d, &Nx_d, &L_d, &No_d, &M_d, &nb_d), "data_info");
  B = blocks_of(yc, L_d);
  L = B.L;
  if (B.M != M_d || B.nblk != nb_d)
    mexErrMsgIdAndTxt("fhs:args", "y has %lld rows in %d blocks, the data state %lld in %d", (long long)B.M, B.nblk,
                      (long long)M_d, nb_d);
  check_like(v2c, &B, "v2");
  if (!dr) check_like(projc, &B, "proj");
  if (mxIsCell(pUc) && mxGetNumberOfElements(pUc) > 0) check_like(pUc, &B, "pU");
  if (mxIsCell(lamc) && mxGetNumberOfElements(lamc) > 0) check_like(lamc, &B, "Lambda");
  if (!mxIsCell(epsc) || (int)mxGetNumberOfElements(epsc) != L)
    mexErrMsgIdAndTxt("fhs:args", "epsilon must be a cell {L}{nblocks}");
  dims = mxGetDimensions(xa);
  Ny = (int)dims[0];
  Nx = (int)dims[1];
  Lx = mxGetNumberOfDimensions(xa) > 2 ? (int)dims[2] : 1;
  if (Ny != Ny_d || Nx != Nx_d || Lx != L || !mxGetDoubles(xa))
    mexErrMsgIdAndTxt("fhs:args", "x must be a real %d x %d x %d array", Ny_d, Nx_d, L);
  if ((int64_t)mxGetNumberOfElements(s22) != 1 && (int64_t)mxGetNumberOfElements(s22) != L)
    mexErrMsgIdAndTxt("fhs:args", "sigma22 must be a scalar or have one value per channel");
  /* stacked device copies */
  h = stack_complex(v2c, B.M); dv2 = to_device(h, 2 * B.M * sizeof(float)); mxFree(h);
  h = stack_complex(yc, B.M); dy = to_device(h, 2 * B.M * sizeof(float)); mxFree(h);
  dproj = NULL;
  if (!dr) { h = stack_complex(projc, B.M); dproj = to_device(h, 2 * B.M * sizeof(float)); mxFree(h); }
  if (mxIsCell(pUc) && mxGetNumberOfElements(pUc) > 0) {
    h = stack_real(pUc, B.M, "pU"); dpU = (float*)to_device(h, B.M * sizeof(float)); mxFree(h);
  }
  if (mxIsCell(lamc) && mxGetNumberOfElements(lamc) > 0) {
    h = stack_real(lamc, B.M, "Lambda"); dlam = (float*)to_device(h, B.M * sizeof(float)); mxFree(h);
  }
  /* the G A(x_old) state */
  fhs_mex_check(fhs_device_alloc(ctx, 2 * B.M * sizeof(float) + 8, &dgx), "device_alloc");
  if (mxIsCell(fxo)) {
    /* the cell returned by a previous call: channel l holds its rows */
    float* hp = (float*)mxMalloc(2 * B.M * sizeof(float) + 8);
    if ((int)mxGetNumberOfElements(fxo) != L) mexErrMsgIdAndTxt("fhs:args", "Fx_old: the cell returned by a previous call");
    k = 0;
    for (l = 0; l < L; ++l) {
      const mxArray* a = mxGetCell(fxo, l);
      if (!a || !mxIsComplex(a) || (int64_t)mxGetNumberOfElements(a) != B.chan_rows[l])
        mexErrMsgIdAndTxt("fhs:args", "Fx_old{%d}: the state returned by a previous call (%lld complex rows)", l + 1,
                          (long long)B.chan_rows[l]);
      {
        const mxComplexDouble* z = mxGetComplexDoubles(a);
        for (i = 0; i < B.chan_rows[l]; ++i, ++k) { hp[2 * k] = (float)z[i].real; hp[2 * k + 1] = (float)z[i].imag; }
      }
    }
    fhs_mex_check(fhs_memcpy_h2d(ctx, dgx, hp, 2 * B.M * sizeof(float)), "memcpy_h2d");
    mxFree(hp);
  } else if (mxGetNumberOfElements(fxo) == 0 || all_zero(fxo)) {
    float* hp = (float*)mxMalloc(2 * B.M * sizeof(float) + 8);
    memset(hp, 0, 2 * B.M * sizeof(float));
    fhs_mex_check(fhs_memcpy_h2d(ctx, dgx, hp, 2 * B.M * sizeof(float)), "memcpy_h2d");
    mxFree(hp);
  } else {
    /* the reference's A(x_old): [No, L], G Fx_old formed on the device */
    void* dfx;
    int64_t nf;
    if (mxGetM(fxo) != (size_t)No_d || mxGetN(fxo) != (size_t)L)
      mexErrMsgIdAndTxt("fhs:args", "Fx_old must be [No, L] = [%lld, %d] (A(x_old) per channel), [] or a previous state",
                        (long long)No_d, L);
    dfx = fhs_mex_to_device_complex(fxo, &nf);
    fhs_mex_check(fhs_data_fx_old_to_state(ctx, D, dfx, dlam, dr, dgx), "data_fx_old_to_state");
    fhs_device_free(ctx, dfx);
  }
  /* epsilon per block, sigma22 per channel, x */
  eps = (double*)mxMalloc((B.nblk + 1) * sizeof(double));
  nres = (double*)mxMalloc((B.nblk + 1) * sizeof(double));
  k = 0;
  for (l = 0; l < L; ++l) {
    const mxArray* cl = mxGetCell(epsc, l);
    const mxArray* yl = mxGetCell(yc, l);
    if (!cl || !mxIsCell(cl) || mxGetNumberOfElements(cl) != mxGetNumberOfElements(yl))
      mexErrMsgIdAndTxt("fhs:args", "epsilon{%d} must have one value per block of y{%d}", l + 1, l + 1);
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
  plhs[0] = unstack_complex(dv2, B.M, v2c);
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
    /* the G A(x) state: one complex column per channel */
    float* hg = (float*)mxMalloc(2 * B.M * sizeof(float) + 8);
    mxArray* st = mxCreateCellMatrix(L, 1);
    fhs_mex_check(fhs_memcpy_d2h(ctx, hg, dgx, 2 * B.M * sizeof(float)), "memcpy_d2h");
    k = 0;
    for (l = 0; l < L; ++l) {
      mxArray* a = mxCreateDoubleMatrix((mwSize)B.chan_rows[l], 1, mxCOMPLEX);
      mxComplexDouble* z = mxGetComplexDoubles(a);
      for (i = 0; i < B.chan_rows[l]; ++i, ++k) { z[i].real = hg[2 * k]; z[i].imag = hg[2 * k + 1]; }
      mxSetCell(st, l, a);
    }
    mxFree(hg);
    plhs[2] = st;
  }
  plhs[3] = dr ? mxDuplicateArray(projc) : unstack_complex(dproj, B.M, projc);
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
}
