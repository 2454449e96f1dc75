/*
 * [I_overlap_ref, dims_overlap_ref, I_overlap, dims_overlap, status, offset,
 *  offsetL, offsetR, Ncoefs, temLIdxs, temRIdxs] =
 *      sdwt2_setup(imsize, I, dims, nlevel, wavelet, filter_length)
 * (called at src/fhs/facetHyperSARA.m:325-326; the faceted-wavelet submodule
 * that defines it is absent from the reference tree).  Drop-in over
 * fhs_sara_facets_create: the facets are the grid of I (0-based origins,
 * Q x 2) and dims (Q x 2), facet q = (qx-1)*Qy + qy; wavelet a cell of
 * 'db1'..'db8' and 'self' (last); the filters come from wfilters(name, 'r').
 * Outputs in the solver's own terms: dims_overlap_ref(q, :) and
 * I_overlap_ref(q, :) the in-image part of facet q's extended image (its
 * pixels plus the top/left border its neighbours supply), offsetL(q, :) the
 * zero rows / columns that complete it at the image's top / left edge,
 * offsetR zeros, Ncoefs{q} the (rows, cols) of every level and of the Dirac
 * part (so 3 sum(p) - 2 sum(p(nlevel+1:nlevel+1:end)) - 2 p(end), p =
 * prod(Ncoefs{q}, 2), is the facet's coefficient count), I_overlap{q} and
 * dims_overlap{q} the same region as the _ref outputs.  offset is the
 * device handle (kind 'sara_facets'), status(q) the facet index and
 * temLIdxs{q} the handle with the facet index (field fhs_facet): the three
 * values sdwt2_sara_faceting and isdwt2_sara_faceting receive back.
 */
#include "fhs_mex_common.h"

static mxArray* facet_handle(fhs_sara_facets* F, int q) {
  const char* fields[] = {"fhs_handle", "fhs_kind", "fhs_facet"};
  mxArray* s = mxCreateStructMatrix(1, 1, q >= 0 ? 3 : 2, fields);
  mxArray* hv = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *(uint64_t*)mxGetData(hv) = (uint64_t)(uintptr_t)F;
  mxSetField(s, 0, "fhs_handle", hv);
  mxSetField(s, 0, "fhs_kind", mxCreateString("sara_facets"));
  if (q >= 0) mxSetField(s, 0, "fhs_facet", mxCreateDoubleScalar((double)q));
  return s;
}

static int sorted_unique(const double* v, int n, int* out) {
  int k = 0, i, j;
  for (i = 0; i < n; ++i) {
    const int x = (int)v[i];
    for (j = 0; j < k && out[j] != x; ++j) {
    }
    if (j == k) out[k++] = x;
  }
  for (i = 1; i < k; ++i) {   /* insertion sort */
    const int x = out[i];
    for (j = i - 1; j >= 0 && out[j] > x; --j) out[j + 1] = out[j];
    out[j + 1] = x;
  }
  return k;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  fhs_ctx* ctx = fhs_mex_ctx();
  int M, N, Q, Qy, Qx, nlevel, nw, nb = 0, with_dirac = 0, q, b, total = 0, off = 0, nrows;
  const double *I, *dims;
  int *yb, *xb, *flen, *rows;
  double* lo;
  fhs_sara_facets* F = NULL;
  mxArray* out[11];
  char name[32];
  if (nrhs != 6 || mxGetNumberOfElements(prhs[0]) != 2 || !mxIsCell(prhs[4]) || mxGetN(prhs[1]) != 2 ||
      mxGetN(prhs[2]) != 2 || mxGetM(prhs[1]) != mxGetM(prhs[2]))
    mexErrMsgIdAndTxt("fhs:args", "sdwt2_setup(imsize, I, dims, nlevel, wavelet, filter_length): I and dims Q x 2, "
                                  "wavelet a cell");
  M = (int)mxGetDoubles(prhs[0])[0];
  N = (int)mxGetDoubles(prhs[0])[1];
  Q = (int)mxGetM(prhs[1]);
  I = mxGetDoubles(prhs[1]);
  dims = mxGetDoubles(prhs[2]);
  nlevel = (int)mxGetScalar(prhs[3]);
  yb = (int*)mxMalloc((Q + 1) * sizeof(int));
  xb = (int*)mxMalloc((Q + 1) * sizeof(int));
  Qy = sorted_unique(I, Q, yb);          /* column 1: row origins */
  Qx = sorted_unique(I + Q, Q, xb);      /* column 2: column origins */
  if (Qy * Qx != Q) mexErrMsgIdAndTxt("fhs:args", "sdwt2_setup: the facets must form a grid");
  yb[Qy] = M;
  xb[Qx] = N;
  for (q = 0; q < Q; ++q) {   /* MATLAB order q = (qx - 1) Qy + qy, facets tiling the image */
    const int qy = q % Qy, qx = q / Qy;
    if ((int)I[q] != yb[qy] || (int)I[Q + q] != xb[qx] || (int)dims[q] != yb[qy + 1] - yb[qy] ||
        (int)dims[Q + q] != xb[qx + 1] - xb[qx])
      mexErrMsgIdAndTxt("fhs:args", "sdwt2_setup: facet %d is not the (%d, %d) tile of the grid", q + 1, qy + 1,
                        qx + 1);
  }
  nw = (int)mxGetNumberOfElements(prhs[4]);
  flen = (int*)mxMalloc((nw + 1) * sizeof(int));
  lo = (double*)mxMalloc((16 * nw + 1) * sizeof(double));
  for (b = 0; b < nw; ++b) {
    mxArray *args[2], *f = NULL;
    if (mxGetString(mxGetCell(prhs[4], b), name, sizeof(name)))
      mexErrMsgIdAndTxt("fhs:args", "sdwt2_setup: wavelet names must be strings");
    if (!strcmp(name, "self")) {
      if (b != nw - 1) mexErrMsgIdAndTxt("fhs:args", "sdwt2_setup: 'self' must be the last basis");
      with_dirac = 1;
      continue;
    }
    args[0] = mxCreateString(name);
    args[1] = mxCreateString("r");
    mexCallMATLAB(1, &f, 2, args, "wfilters");
    flen[nb] = (int)mxGetNumberOfElements(f);
    if (flen[nb] < 2 || flen[nb] > 16) mexErrMsgIdAndTxt("fhs:args", "sdwt2_setup: %s: filters of 2..16 taps", name);
    memcpy(lo + off, mxGetDoubles(f), flen[nb] * sizeof(double));
    off += flen[nb];
    total += flen[nb];
    ++nb;
  }
  fhs_mex_check(fhs_sara_facets_create(ctx, M, N, nlevel, nb, flen, lo, with_dirac, Qy, Qx, yb, xb, &F),
                "sara_facets_create");
  nrows = nb * (nlevel + 1) + with_dirac;
  rows = (int*)mxMalloc(2 * nrows * sizeof(int));
  out[0] = mxCreateDoubleMatrix(Q, 2, mxREAL);   /* I_overlap_ref */
  out[1] = mxCreateDoubleMatrix(Q, 2, mxREAL);   /* dims_overlap_ref */
  out[2] = mxCreateCellMatrix(Q, 1);             /* I_overlap */
  out[3] = mxCreateCellMatrix(Q, 1);             /* dims_overlap */
  out[4] = mxCreateDoubleMatrix(Q, 1, mxREAL);   /* status: facet index */
  out[5] = facet_handle(F, -1);                  /* offset: the handle */
  out[6] = mxCreateDoubleMatrix(Q, 2, mxREAL);   /* offsetL */
  out[7] = mxCreateDoubleMatrix(Q, 2, mxREAL);   /* offsetR */
  out[8] = mxCreateCellMatrix(Q, 1);             /* Ncoefs */
  out[9] = mxCreateCellMatrix(Q, 1);             /* temLIdxs: handle + facet */
  out[10] = mxCreateCellMatrix(Q, 1);            /* temRIdxs */
  for (q = 0; q < Q; ++q) {
    int ext[4], r;
    int64_t n = 0;
    double *o, *d, *nc;
    mxArray *io, *dd, *ncm;
    fhs_mex_check(fhs_sara_facets_info(F, q, ext, &n, rows), "sara_facets_info");
    {
      const int py = ext[0] < 0 ? -ext[0] : 0, px = ext[1] < 0 ? -ext[1] : 0;
      const double oy = ext[0] + py, ox = ext[1] + px, ny = ext[2] - py, nx = ext[3] - px;
      mxGetDoubles(out[0])[q] = oy;
      mxGetDoubles(out[0])[Q + q] = ox;
      mxGetDoubles(out[1])[q] = ny;
      mxGetDoubles(out[1])[Q + q] = nx;
      mxGetDoubles(out[6])[q] = py;
      mxGetDoubles(out[6])[Q + q] = px;
      io = mxCreateDoubleMatrix(1, 2, mxREAL);
      o = mxGetDoubles(io);
      o[0] = oy;
      o[1] = ox;
      dd = mxCreateDoubleMatrix(1, 2, mxREAL);
      d = mxGetDoubles(dd);
      d[0] = ny;
      d[1] = nx;
    }
    mxSetCell(out[2], q, io);
    mxSetCell(out[3], q, dd);
    mxGetDoubles(out[4])[q] = q;
    ncm = mxCreateDoubleMatrix(nrows, 2, mxREAL);
    nc = mxGetDoubles(ncm);
    for (r = 0; r < nrows; ++r) {
      nc[r] = rows[2 * r];
      nc[nrows + r] = rows[2 * r + 1];
    }
    mxSetCell(out[8], q, ncm);
    mxSetCell(out[9], q, facet_handle(F, q));
    mxSetCell(out[10], q, mxCreateDoubleMatrix(0, 0, mxREAL));
  }
  for (q = 0; q < 11; ++q) {
    if (q < nlhs || (q == 0 && nlhs == 0))
      plhs[q] = out[q];
    else
      mxDestroyArray(out[q]);
  }
  (void)total;
}
