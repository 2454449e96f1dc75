// Interpolation matrix G: construction (host -> device layouts) and the
// apply_direct_operator / apply_adjoint_operator entry points.
#include <algorithm>
#include <numeric>

#include "gmat.hpp"

namespace fhs {

// column ranges of ~ADJ_ENTRIES nonzeros (<= ADJ_COLS columns)
std::vector<int> adjoint_units(const std::vector<int>& colptr) {
  const int64_t ncols = (int64_t)colptr.size() - 1;
  std::vector<int> u{0};
  int64_t cnt = 0, cols = 0;
  for (int64_t c = 0; c < ncols; ++c) {
    const int64_t len = colptr[c + 1] - colptr[c];
    if (cols > 0 && (cnt + len > ADJ_ENTRIES || cols >= ADJ_COLS)) {
      u.push_back((int)c);
      cnt = 0;
      cols = 0;
    }
    cnt += len;
    ++cols;
  }
  u.push_back((int)ncols);
  return u;
}

// CSC of a separable G from its anchors: colptr over the Kty*Ktx grid and per
// nonzero (row << 6) | (jx*Ky + jy), rows ascending inside every column.
void sep_csc(int64_t M, const int32_t* anchor, int Kty, int Ktx, int Ky, int Kx, std::vector<int>& colptr,
             std::vector<unsigned>& pk) {
  const int64_t ncols = (int64_t)Kty * Ktx;
  colptr.assign(ncols + 1, 0);
  for (int64_t m = 0; m < M; ++m) {
    const int a = anchor[m];
    const int ky0 = a % Kty, kx0 = a / Kty;
    for (int jx = 0; jx < Kx; ++jx)
      for (int jy = 0; jy < Ky; ++jy) colptr[(ky0 + jy) % Kty + Kty * ((kx0 + jx) % Ktx) + 1]++;
  }
  for (int64_t k = 0; k < ncols; ++k) colptr[k + 1] += colptr[k];
  std::vector<int> pos(colptr.begin(), colptr.end() - 1);
  pk.resize((size_t)M * Ky * Kx);
  for (int64_t m = 0; m < M; ++m) {
    const int a = anchor[m];
    const int ky0 = a % Kty, kx0 = a / Kty;
    for (int jx = 0; jx < Kx; ++jx)
      for (int jy = 0; jy < Ky; ++jy) {
        const int col = (ky0 + jy) % Kty + Kty * ((kx0 + jx) % Ktx);
        pk[pos[col]++] = ((unsigned)m << 6) | (unsigned)(jx * Ky + jy);
      }
  }
}

namespace {

template <class T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  FHS_HIP(hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) FHS_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// CSC (column pointers, rows, values) of a COO list; rows ascending per column.
void build_csc(int64_t ncols, const std::vector<int>& rows, const std::vector<int>& cols,
               const std::vector<float2>& vals, std::vector<int>& colptr,
               std::vector<int>& crow, std::vector<float2>& cval) {
  const size_t nnz = rows.size();
  colptr.assign(ncols + 1, 0);
  for (size_t e = 0; e < nnz; ++e) colptr[cols[e] + 1]++;
  for (int64_t k = 0; k < ncols; ++k) colptr[k + 1] += colptr[k];
  std::vector<int> pos(colptr.begin(), colptr.end() - 1);
  crow.resize(nnz);
  cval.resize(nnz);
  // entries arrive row-major (rows ascending) so each column stays sorted
  for (size_t e = 0; e < nnz; ++e) {
    const int p = pos[cols[e]]++;
    crow[p] = rows[e];
    cval[p] = vals[e];
  }
}

struct TileLists {
  int tiles_y = 0, tiles_x = 0;
  std::vector<int> rows, cell, empty;
  std::vector<int4> items, red;
  int slots = 0;
  std::vector<int> prim_rows;
  std::vector<int4> fitems;
};

// every row under each tile its Ky x Kx window touches, sorted by local cell
TileLists tile_lists(int64_t M, const int32_t* anchor, int Kty, int Ktx, int Ky, int Kx) {
  TileLists T;
  T.tiles_y = Kty / TILE_Y;
  T.tiles_x = Ktx / TILE_X;
  const int nt = T.tiles_y * T.tiles_x, ncy = TILE_Y + Ky - 1;
  auto visit = [&](int a, auto&& f) {
    const int ay = a % Kty, ax = a / Kty;
    const int ys[2] = {ay / TILE_Y, ((ay + Ky - 1) % Kty) / TILE_Y};
    const int xs[2] = {ax / TILE_X, ((ax + Kx - 1) % Ktx) / TILE_X};
    for (int i = 0; i < 2; ++i) {
      if (i == 1 && xs[1] == xs[0]) break;
      for (int j = 0; j < 2; ++j) {
        if (j == 1 && ys[1] == ys[0]) break;
        const int ty = ys[j], tx = xs[i];
        const int cyl = ((ay - ty * TILE_Y + Ky - 1) % Kty + Kty) % Kty;
        const int cxl = ((ax - tx * TILE_X + Kx - 1) % Ktx + Ktx) % Ktx;
        f(tx * T.tiles_y + ty, cxl * ncy + cyl);
      }
    }
  };
  std::vector<int64_t> cnt(nt + 1, 0);
  for (int64_t m = 0; m < M; ++m) visit(anchor[m], [&](int t, int) { cnt[t + 1]++; });
  for (int t = 0; t < nt; ++t) cnt[t + 1] += cnt[t];
  FHS_CHECK(cnt[nt] < (1ll << 31), "gmat_create_separable: tile lists too long");
  std::vector<std::pair<int, int>> ent(cnt[nt]);   // (cell, row)
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int64_t m = 0; m < M; ++m)
    visit(anchor[m], [&](int t, int c) { ent[pos[t]++] = {c, (int)m}; });
  T.rows.resize(cnt[nt]);
  T.cell.resize(cnt[nt]);
  for (int t = 0; t < nt; ++t) std::sort(ent.begin() + cnt[t], ent.begin() + cnt[t + 1]);   // by cell, then row
  for (size_t e = 0; e < ent.size(); ++e) {
    T.cell[e] = ent[e].first;
    T.rows[e] = ent[e].second;
  }
  // work items, largest first: whole tiles of at most TILE_ITEM listed rows,
  // larger tiles cut into TILE_ITEM pieces (partial slots), empty tiles in groups
  std::vector<int> order(nt);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return cnt[a + 1] - cnt[a] > cnt[b + 1] - cnt[b]; });
  for (int t : order) {
    const int b = (int)cnt[t], e = (int)cnt[t + 1], n = e - b;
    if (n == 0) {
      T.empty.push_back(t);
      continue;
    }
    const int k = (n + TILE_ITEM - 1) / TILE_ITEM;
    if (k == 1) {
      T.items.push_back(make_int4(t, b, e, -1));
      continue;
    }
    T.red.push_back(make_int4(t, T.slots, k, 0));
    for (int j = 0; j < k; ++j)
      T.items.push_back(make_int4(t, b + j * TILE_ITEM, std::min(e, b + (j + 1) * TILE_ITEM), T.slots++));
  }
  std::stable_sort(T.items.begin(), T.items.end(), [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
  for (size_t e = 0; e < T.empty.size(); e += TILE_EMPTY_GROUP)
    T.items.push_back(make_int4(-1, (int)e, (int)std::min(T.empty.size(), e + TILE_EMPTY_GROUP), -2));
  // forward: every row once, under the tile of its anchor, rows ascending
  std::vector<int64_t> pc(nt + 1, 0);
  auto prim = [&](int a) { return (a / Kty / TILE_X) * T.tiles_y + (a % Kty) / TILE_Y; };
  for (int64_t m = 0; m < M; ++m) pc[prim(anchor[m]) + 1]++;
  for (int t = 0; t < nt; ++t) pc[t + 1] += pc[t];
  T.prim_rows.resize(M);
  {
    std::vector<int64_t> pp(pc.begin(), pc.end() - 1);
    for (int64_t m = 0; m < M; ++m) T.prim_rows[pp[prim(anchor[m])]++] = (int)m;
  }
  for (int t = 0; t < nt; ++t)
    for (int64_t b = pc[t]; b < pc[t + 1]; b += FWD_ITEM)
      T.fitems.push_back(make_int4(t, (int)b, (int)std::min<int64_t>(pc[t + 1], b + FWD_ITEM), 0));
  std::stable_sort(T.fitems.begin(), T.fitems.end(),
                   [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
  return T;
}

__global__ __launch_bounds__(ADJ_THREADS) void k_gmat_adjoint_reduce(GDev g, float2* __restrict__ out) {
  g_adjoint_tile_reduce(g, blockIdx.x, out);
}

template <int KM>
__global__ __launch_bounds__(ADJ_THREADS) void k_gmat_forward_tiled(GDev g, const float2* __restrict__ Fx,
                                                                  const float* __restrict__ lam,
                                                                  float2* __restrict__ y) {
  __shared__ float2 win[fwd_win<KM>()];
  g_forward_tile<KM>(
      g, blockIdx.x, Fx, win, [&](int r) { return lam ? lam[r] : 1.f; },
      [&](int r, float2 v, float s) { y[r] = lam ? cscale(v, s) : v; });
}

__global__ void k_gmat_forward(GDev g, const float2* __restrict__ Fx,
                               const float* __restrict__ lam, float2* __restrict__ y) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  float2 acc = g_row_dot(g, m, Fx);
  if (lam) acc = cscale(acc, lam[m]);
  y[m] = acc;
}

template <int KM>
__global__ __launch_bounds__(ADJ_THREADS) void k_gmat_adjoint(GDev g, const float2* __restrict__ y,
                                                               const float* __restrict__ lam,
                                                               float2* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char lds[];
  auto none = [](int) { return make_float2(0.f, 0.f); };
  auto wl = [&](int r) { return cscale(y[r], lam[r]); };
  auto wp = [&](int r) { return y[r]; };
  if (g.tiled) {
    if (lam) g_adjoint_tile<KM>(g, blockIdx.x, wl, out, lds);
    else g_adjoint_tile<KM>(g, blockIdx.x, wp, out, lds);
  } else {
    float2* prod = reinterpret_cast<float2*>(lds);
    if (lam) g_adjoint_unit(g, blockIdx.x, wl, none, out, prod);
    else g_adjoint_unit(g, blockIdx.x, wp, none, out, prod);
  }
}

}  // namespace

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_gmat_create_csr_host(fhs_ctx* ctx, int64_t M, int64_t ncols, const int64_t* rowptr,
                             const int64_t* colidx, const float* vals, fhs_gmat** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(M > 0 && ncols > 0 && M < (1ll << 31) && ncols < (1ll << 31),
              "gmat_create_csr: M and ncols must be in [1, 2^31)");
    FHS_CHECK(rowptr[0] == 0, "gmat_create_csr: rowptr[0] must be 0");
    const int64_t nnz = rowptr[M];
    FHS_CHECK(nnz >= 0 && nnz < (1ll << 31), "gmat_create_csr: nnz out of range");
    int K = 0;
    for (int64_t m = 0; m < M; ++m) {
      FHS_CHECK(rowptr[m + 1] >= rowptr[m], "gmat_create_csr: rowptr not monotone");
      K = std::max<int>(K, (int)(rowptr[m + 1] - rowptr[m]));
    }
    K = std::max(K, 1);
    std::vector<int> ell_c((size_t)K * M, 0);
    std::vector<float2> ell_v((size_t)K * M, make_float2(0.f, 0.f));
    std::vector<int> rows(nnz), cols(nnz);
    std::vector<float2> v(nnz);
    for (int64_t m = 0; m < M; ++m) {
      for (int64_t e = rowptr[m]; e < rowptr[m + 1]; ++e) {
        const int64_t c = colidx[e];
        FHS_CHECK(c >= 0 && c < ncols, "gmat_create_csr: column index out of range");
        const int j = (int)(e - rowptr[m]);
        const float2 val = make_float2(vals[2 * e], vals[2 * e + 1]);
        ell_c[(size_t)j * M + m] = (int)c;
        ell_v[(size_t)j * M + m] = val;
        rows[e] = (int)m;
        cols[e] = (int)c;
        v[e] = val;
      }
    }
    std::vector<int> colptr, crow;
    std::vector<float2> cval;
    build_csc(ncols, rows, cols, v, colptr, crow, cval);
    const std::vector<int> units = adjoint_units(colptr);
    FHS_HIP(hipSetDevice(ctx->device));
    fhs_gmat* g = new fhs_gmat();
    g->device = ctx->device;
    g->M = M;
    g->ncols = ncols;
    g->nnz = nnz;
    g->mem[6] = upload(units);
    g->d.ucol = (const int*)g->mem[6];
    g->d.nunits = (int)units.size() - 1;
    g->mem[0] = upload(ell_c);
    g->mem[1] = upload(ell_v);
    g->mem[2] = upload(colptr);
    g->mem[3] = upload(crow);
    g->mem[4] = upload(cval);
    GDev& d = g->d;
    d.layout = G_ELL;
    d.M = (int)M;
    d.K = K;
    d.ncols = (int)ncols;
    d.cols = (const int*)g->mem[0];
    d.vals = (const float2*)g->mem[1];
    d.colptr = (const int*)g->mem[2];
    d.csc_rows = (const int*)g->mem[3];
    d.csc_vals = (const float2*)g->mem[4];
    *out = g;
  });
}

int fhs_gmat_create_window_host(fhs_ctx* ctx, int64_t M, int Kty, int Ktx, int Ky, int Kx,
                                const int32_t* anchor, const float* coef, fhs_gmat** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(M > 0 && M < (1ll << 31), "gmat_create_window: M out of range");
    FHS_CHECK(Ky >= 1 && Kx >= 1 && Ky <= Kty && Kx <= Ktx, "gmat_create_window: bad window");
    const int64_t ncols = (int64_t)Kty * Ktx;
    FHS_CHECK(ncols < (1ll << 31), "gmat_create_window: grid too large");
    const int K = Ky * Kx;
    const int64_t nnz = M * K;
    FHS_CHECK(nnz < (1ll << 31), "gmat_create_window: nnz out of range");
    std::vector<float2> wv((size_t)K * M);
    std::vector<int> rows(nnz), cols(nnz);
    std::vector<float2> v(nnz);
    std::vector<int> anc(anchor, anchor + M);
    for (int64_t m = 0; m < M; ++m) {
      const int a = anchor[m];
      FHS_CHECK(a >= 0 && a < ncols, "gmat_create_window: anchor out of range");
      const int ky0 = a % Kty, kx0 = a / Kty;
      for (int jx = 0; jx < Kx; ++jx) {
        for (int jy = 0; jy < Ky; ++jy) {
          const int j = jx * Ky + jy;
          const float2 val = make_float2(coef[2 * (m * K + j)], coef[2 * (m * K + j) + 1]);
          wv[(size_t)j * M + m] = val;
          const int ky = (ky0 + jy) % Kty, kx = (kx0 + jx) % Ktx;
          const int64_t e = m * K + j;
          rows[e] = (int)m;
          cols[e] = ky + Kty * kx;
          v[e] = val;
        }
      }
    }
    std::vector<int> colptr, crow;
    std::vector<float2> cval;
    build_csc(ncols, rows, cols, v, colptr, crow, cval);
    const std::vector<int> units = adjoint_units(colptr);
    FHS_HIP(hipSetDevice(ctx->device));
    fhs_gmat* g = new fhs_gmat();
    g->device = ctx->device;
    g->M = M;
    g->ncols = ncols;
    g->nnz = nnz;
    g->mem[6] = upload(units);
    g->d.ucol = (const int*)g->mem[6];
    g->d.nunits = (int)units.size() - 1;
    g->mem[0] = upload(anc);
    g->mem[1] = upload(wv);
    g->mem[2] = upload(colptr);
    g->mem[3] = upload(crow);
    g->mem[4] = upload(cval);
    GDev& d = g->d;
    d.layout = G_WINDOW;
    d.M = (int)M;
    d.K = K;
    d.Kty = Kty;
    d.Ktx = Ktx;
    d.Ky = Ky;
    d.Kx = Kx;
    d.ncols = (int)ncols;
    d.anchor = (const int*)g->mem[0];
    d.vals = (const float2*)g->mem[1];
    d.colptr = (const int*)g->mem[2];
    d.csc_rows = (const int*)g->mem[3];
    d.csc_vals = (const float2*)g->mem[4];
    *out = g;
  });
}

int fhs_gmat_create_separable_host(fhs_ctx* ctx, int64_t M, int Kty, int Ktx, int Ky, int Kx,
                                   const int32_t* anchor, const float* cy, const float* cx,
                                   fhs_gmat** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(M > 0 && M < (1ll << 26), "gmat_create_separable: M must be in [1, 2^26)");
    FHS_CHECK(Ky >= 1 && Kx >= 1 && Ky <= SEP_KMAX && Kx <= SEP_KMAX && Ky <= Kty && Kx <= Ktx,
              "gmat_create_separable: 1 <= Ky, Kx <= 16");
    const int64_t ncols = (int64_t)Kty * Ktx;
    FHS_CHECK(ncols < (1ll << 31), "gmat_create_separable: grid too large");
    const int K = Ky * Kx;
    const int64_t nnz = M * K;
    FHS_CHECK(nnz < (1ll << 31), "gmat_create_separable: nnz out of range");
    const int km = sep_kmax_for(Ky, Kx), row = 2 * km;
    std::vector<float2> sep((size_t)row * M, make_float2(0.f, 0.f));
    std::vector<int> anc(anchor, anchor + M);
    for (int64_t m = 0; m < M; ++m) {
      const int a = anchor[m];
      FHS_CHECK(a >= 0 && a < ncols, "gmat_create_separable: anchor out of range");
      for (int jy = 0; jy < Ky; ++jy)
        sep[(size_t)m * row + jy] = make_float2(cy[2 * (m * Ky + jy)], cy[2 * (m * Ky + jy) + 1]);
      for (int jx = 0; jx < Kx; ++jx)
        sep[(size_t)m * row + km + jx] = make_float2(cx[2 * (m * Kx + jx)], cx[2 * (m * Kx + jx) + 1]);
    }
    const bool tiled = tile_grid_ok(Kty, Ktx, Ky, Kx);
    // the CSC of small grids packs (row << 6) | tap: 64 taps at most
    FHS_CHECK(tiled || K <= 64, "gmat_create_separable: more than 64 taps per row needs a grid of whole 32 x 32 "
                                "tiles (the tiled adjoint)");
    std::vector<int> colptr, units;
    std::vector<unsigned> pk;
    TileLists tl;
    if (tiled) {
      tl = tile_lists(M, anchor, Kty, Ktx, Ky, Kx);
    } else {
      sep_csc(M, anchor, Kty, Ktx, Ky, Kx, colptr, pk);
      units = adjoint_units(colptr);
    }
    FHS_HIP(hipSetDevice(ctx->device));
    fhs_gmat* g = new fhs_gmat();
    g->device = ctx->device;
    g->M = M;
    g->ncols = ncols;
    g->nnz = nnz;
    g->mem[0] = upload(anc);
    g->mem[1] = upload(sep);
    GDev& d = g->d;
    d.layout = G_SEP;
    d.kmax = km;
    d.M = (int)M;
    d.K = K;
    d.Kty = Kty;
    d.Ktx = Ktx;
    d.Ky = Ky;
    d.Kx = Kx;
    d.ncols = (int)ncols;
    d.anchor = (const int*)g->mem[0];
    d.vals = (const float2*)g->mem[1];
    if (tiled) {
      g->mem[8] = upload(tl.rows);
      g->mem[9] = upload(tl.cell);
      g->mem[4] = upload(tl.items);
      g->mem[5] = upload(tl.empty);
      g->mem[7] = upload(tl.red);
      FHS_HIP(hipMalloc(&g->mem[10], std::max<size_t>(1, (size_t)tl.slots * TILE_Y * TILE_X) * sizeof(float2)));
      g->mem[11] = upload(tl.prim_rows);
      g->mem[12] = upload(tl.fitems);
      d.tiled = 1;
      d.tiles_y = tl.tiles_y;
      d.tiles_x = tl.tiles_x;
      d.tile_rows = (const int*)g->mem[8];
      d.tile_cell = (const int*)g->mem[9];
      d.items = (const int4*)g->mem[4];
      d.empty_tiles = (const int*)g->mem[5];
      d.red = (const int4*)g->mem[7];
      d.nred = (int)tl.red.size();
      d.partial = (float2*)g->mem[10];
      d.prim_rows = (const int*)g->mem[11];
      d.fitems = (const int4*)g->mem[12];
      d.nfitems = (int)tl.fitems.size();
      d.nunits = (int)tl.items.size();
    } else {
      g->mem[2] = upload(colptr);
      g->mem[3] = upload(pk);
      g->mem[6] = upload(units);
      d.colptr = (const int*)g->mem[2];
      d.csc_pk = (const unsigned*)g->mem[3];
      d.ucol = (const int*)g->mem[6];
      d.nunits = (int)units.size() - 1;
    }
    *out = g;
  });
}

int fhs_gmat_destroy(fhs_gmat* g) {
  if (!g) return FHS_OK;
  (void)hipSetDevice(g->device);
  for (void* p : g->mem)
    if (p) (void)hipFree(p);
  delete g;
  return FHS_OK;
}

int fhs_gmat_info(const fhs_gmat* g, int64_t* M, int64_t* ncols, int64_t* nnz) {
  if (!g) return FHS_ERR_ARG;
  if (M) *M = g->M;
  if (ncols) *ncols = g->ncols;
  if (nnz) *nnz = g->nnz;
  return FHS_OK;
}

int fhs_apply_direct_operator(fhs_ctx* ctx, const fhs_gmat* G, const void* Fx,
                              const float* Lambda, void* y) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(G && Fx && y, "apply_direct_operator: null argument");
    Scope sc(ctx, "gmat_forward");
    if (G->d.tiled) {
      if (G->d.nfitems > 0)
        (G->d.kmax > 8 ? k_gmat_forward_tiled<16> : k_gmat_forward_tiled<8>)<<<G->d.nfitems, ADJ_THREADS, 0,
                                                                               ctx->stream>>>(
            G->d, (const float2*)Fx, Lambda, (float2*)y);
    } else {
      k_gmat_forward<<<cdiv(G->M, 256), 256, 0, ctx->stream>>>(G->d, (const float2*)Fx, Lambda, (float2*)y);
    }
    check_launch();
  });
}

int fhs_apply_adjoint_operator(fhs_ctx* ctx, const fhs_gmat* G, const void* y,
                               const float* Lambda, void* Fx) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(G && Fx && y, "apply_adjoint_operator: null argument");
    Scope sc(ctx, "gmat_adjoint");
    const size_t lds = G->d.tiled ? tile_lds_bytes(G->d.Ky, G->d.Kx) : ADJ_CSC_LDS_BYTES;
    auto kern = G->d.kmax > 8 ? k_gmat_adjoint<16> : k_gmat_adjoint<8>;
    allow_dynamic_lds((const void*)kern, lds);
    kern<<<G->d.nunits, ADJ_THREADS, lds, ctx->stream>>>(G->d, (const float2*)y, Lambda, (float2*)Fx);
    check_launch();
    if (G->d.tiled && G->d.nred > 0) {
      k_gmat_adjoint_reduce<<<G->d.nred, ADJ_THREADS, 0, ctx->stream>>>(G->d, (float2*)Fx);
      check_launch();
    }
  });
}

}  // extern "C"
