/*
 * [epsilon, t_block] = update_epsilon(epsilon, t, t_block, norm_res,
 *     adapt_eps_tol_in, adapt_eps_tol_out, adapt_eps_steps,
 *     adapt_eps_change_percentage)              (src/update_epsilon.m:52-71)
 * Drop-in MEX over fhs_update_epsilon (host logic).  epsilon, t_block and
 * norm_res are cells {L}{nblocks} of scalars (or flat cells / arrays); the
 * outputs keep the input's cell structure.
 */
#include "fhs_mex_common.h"

/* number of scalars in a (nested) cell or array */
static int64_t count_leaves(const mxArray* a) {
  if (mxIsCell(a)) {
    int64_t n = 0, i, m = (int64_t)mxGetNumberOfElements(a);
    for (i = 0; i < m; ++i) n += count_leaves(mxGetCell(a, i));
    return n;
  }
  return (int64_t)mxGetNumberOfElements(a);
}

static void gather(const mxArray* a, double* out, int64_t* k) {
  if (mxIsCell(a)) {
    int64_t i, m = (int64_t)mxGetNumberOfElements(a);
    for (i = 0; i < m; ++i) gather(mxGetCell(a, i), out, k);
    return;
  } else {
    int64_t i, m = (int64_t)mxGetNumberOfElements(a);
    const double* r = mxGetDoubles(a);
    if (!r && m > 0) mexErrMsgIdAndTxt("fhs:args", "update_epsilon: cells of real double scalars");
    for (i = 0; i < m; ++i) out[(*k)++] = r[i];
  }
}

/* a copy of the structure of `like` filled from vals */
static mxArray* scatter(const mxArray* like, const double* vals, int64_t* k) {
  if (mxIsCell(like)) {
    const mwSize m = mxGetNumberOfElements(like);
    mxArray* c = mxCreateCellMatrix(mxGetM(like), mxGetN(like));
    mwSize i;
    for (i = 0; i < m; ++i) mxSetCell(c, i, scatter(mxGetCell(like, i), vals, k));
    return c;
  } else {
    mxArray* r = mxDuplicateArray(like);
    double* d = mxGetDoubles(r);
    const int64_t m = (int64_t)mxGetNumberOfElements(like);
    int64_t i;
    for (i = 0; i < m; ++i) d[i] = vals[(*k)++];
    return r;
  }
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  int64_t n, k, i;
  double *eps, *tb_d, *res;
  int *tb, n_updated = 0;
  (void)nlhs;
  /* host logic only: no GPU context */
  if (nrhs != 8) mexErrMsgIdAndTxt("fhs:args", "[epsilon, t_block] = update_epsilon(epsilon, t, t_block, norm_res, "
                                              "tol_in, tol_out, steps, change_percentage)");
  n = count_leaves(prhs[0]);
  if (count_leaves(prhs[2]) != n || count_leaves(prhs[3]) != n)
    mexErrMsgIdAndTxt("fhs:args", "epsilon, t_block and norm_res must have the same structure");
  eps = (double*)mxMalloc((n + 1) * sizeof(double));
  tb_d = (double*)mxMalloc((n + 1) * sizeof(double));
  res = (double*)mxMalloc((n + 1) * sizeof(double));
  tb = (int*)mxMalloc((n + 1) * sizeof(int));
  k = 0; gather(prhs[0], eps, &k);
  k = 0; gather(prhs[2], tb_d, &k);
  k = 0; gather(prhs[3], res, &k);
  for (i = 0; i < n; ++i) tb[i] = (int)tb_d[i];
  if (fhs_update_epsilon(eps, (int)mxGetScalar(prhs[1]), tb, res, (int)n, mxGetScalar(prhs[4]), mxGetScalar(prhs[5]),
                         (int)mxGetScalar(prhs[6]), mxGetScalar(prhs[7]), &n_updated) != FHS_OK)
    mexErrMsgIdAndTxt("fhs:args", "update_epsilon: invalid arguments");
  for (i = 0; i < n; ++i) tb_d[i] = (double)tb[i];
  k = 0; plhs[0] = scatter(prhs[0], eps, &k);
  k = 0; plhs[1] = scatter(prhs[2], tb_d, &k);
  mxFree(eps); mxFree(tb_d); mxFree(res); mxFree(tb);
}
