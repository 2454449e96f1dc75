// Data-fidelity dual step of all channels and blocks
// (src/update_dual_data_fidelity.m:126-147, non-DR branch; DR branch :85-125).
//
// Per iteration:
//   Fx   = A(x_l)                           (batched scaled FFT)
//   gx   = G_l Fx ;  r2 = 2 gx - gx_old ;  gx_old <- gx
//        (by linearity G(2 Fx - Fx_old) = 2 G Fx - G Fx_old: the state keeps
//         G Fx_old, M complex per channel, instead of Fx_old, No per channel)
//   per block: proj = ellipsoid projection (forward-backward, warm start),
//              v2 += pU (r2 - proj), norm_res = ||gx - y||      (FP64)
//   g2_l = G_l' v2_l (atomic-free gather per grid point)
//   Ftx_l = sigma22_l real(At(g2_l))
// DR branch: G is square lower-triangular (holographic H = G + G'), the
// dual uses the closed-form l2 prox, no ellipsoid.
#include <algorithm>

#include "gmat.hpp"
#include "kernels.hpp"
#include "nufft.hpp"

struct fhs_data {
  int device = 0;
  int L = 0;
  int nblocks = 0;
  int64_t Mtot = 0;
  int maxM = 0;
  int adj_grid = 0;                    // max adjoint units + grid-tail units over CSC channels
  int tile_grid = 0;                   // max tile work items over tile-gridded channels
  int red_grid = 0;                    // max split tiles over tile-gridded channels
  size_t tile_lds = 0;                 // LDS bytes of a tile workgroup (max over channels)
  int fwd_grid = 0;                    // max forward work items over tile-gridded channels
  bool any_rows = false;               // some channel without tile gridding (row kernels)
  int64_t No = 0;
  const fhs_nufft* A = nullptr;
  std::vector<int64_t> row_off;        // [L+1]
  std::vector<int64_t> blk_start;      // [nblocks+1]
  std::vector<int> blk_chan;           // [nblocks]
  bool square = false;                 // all G square (DR holographic form)
  fhs::GDev* gdev = nullptr;           // device [L]
  int64_t* d_row_off = nullptr;        // device [L+1]
  int64_t* d_blk_start = nullptr;      // device [nblocks+1]
  double* d_eps = nullptr;             // device [nblocks]
  float* d_sigma22 = nullptr;          // device [L]
  double* d_norm_res = nullptr;        // device [nblocks] (squared residual norms)
  float2* Fx = nullptr;                // [L][No]
  float2* g2 = nullptr;                // [L][No]
  float2* r2 = nullptr;                // [Mtot]
};

namespace fhs {
namespace {

constexpr int FB_THREADS = 1024;
constexpr int NXCD = 8;   // MI355X: 8 XCDs, each with its own L2

// Workgroups of a launch are dispatched to the XCDs round-robin (linear id
// mod 8).  The per-channel kernels below are launched as a 1-D grid of
// NXCD * ceil(L / NXCD) * units workgroups and remapped so that XCD x only
// runs channels x, x + 8, ...: one channel's v2 / coefficients / grid
// region is then shared by one L2 instead of being pulled into all eight.
// Units are interleaved over those channels (unit-major), so the first
// units of every channel (the heaviest adjoint tiles) start first.
__device__ __forceinline__ bool xcd_map(int L, int units, int& chan, int& unit) {
  const int b = blockIdx.x;
  const int xcd = b % NXCD, s = b / NXCD;
  const int cpx = (L + NXCD - 1) / NXCD;   // channels per XCD
  chan = xcd + NXCD * (s % cpx);
  unit = s / cpx;
  return chan < L && unit < units;
}

inline unsigned xcd_grid(int L, int units) {
  return (unsigned)(NXCD * ((L + NXCD - 1) / NXCD) * units);
}

__global__ __launch_bounds__(256) void k_data_forward(const GDev* __restrict__ gd,
                                                      const int64_t* __restrict__ row_off,
                                                      const float2* __restrict__ Fx, int64_t No,
                                                      float2* __restrict__ gx_old,
                                                      float2* __restrict__ r2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (g.tiled) return;   // k_data_forward_tiled
  const int m = u * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const float2 gx = g_row_dot(g, m, Fx + (size_t)l * No);
  const int64_t i = row_off[l] + m;
  const float2 old = gx_old[i];
  r2[i] = make_float2(2.f * gx.x - old.x, 2.f * gx.y - old.y);
  gx_old[i] = gx;
}

// Holographic (DR) forward: Lambda (G + G') applied to (2 Fx - Fx_old), and
// the residual operator on Fx, G square M x M.
__global__ __launch_bounds__(256) void k_data_forward_dr(const GDev* __restrict__ gd,
                                                         const int64_t* __restrict__ row_off,
                                                         const float2* __restrict__ Fx, int64_t No,
                                                         const float* __restrict__ lam,
                                                         float2* __restrict__ gx_old,
                                                         float2* __restrict__ r2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  const int m = u * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const float2* F = Fx + (size_t)l * No;
  float2 gx = cadd(g_row_dot(g, m, F), g_col_dot(g, m, [&](int r) { return F[r]; }));
  const int64_t i = row_off[l] + m;
  const float lm = lam ? lam[i] : 1.f;
  gx = cscale(gx, lm);
  const float2 old = gx_old[i];
  r2[i] = make_float2(2.f * gx.x - old.x, 2.f * gx.y - old.y);
  gx_old[i] = gx;
}

// One workgroup per data block: forward-backward projection onto
// B(y, eps) in the metric diag(pU), then the dual update and the residual.
// z <- P_B(z - mu pU (z - w)), w = v2/pU + r2, mu = 1/max(pU), warm start
// z = proj; stop after min_it once ||dz|| <= tol ||z|| or at max_it.
// w is formed once and kept in the v2 buffer during the iterations (28 B
// read + 8 B written per row and iteration instead of 44 B); the dual
// update v2 + pU (r2 - z) is then pU (w - z).
__global__ __launch_bounds__(FB_THREADS) void k_fb_dual(
    const int64_t* __restrict__ blk_start, const double* __restrict__ eps_arr,
    float2* __restrict__ v2, const float2* __restrict__ y, const float* __restrict__ pU,
    const float2* __restrict__ r2, float2* __restrict__ z, const float2* __restrict__ gx,
    int max_it, int min_it, double tol, double* __restrict__ norm_res) {
  __shared__ double sh[4 * FB_THREADS / 64];
  __shared__ float shm[FB_THREADS / 64];
  const int64_t a = blk_start[blockIdx.x], b = blk_start[blockIdx.x + 1];
  const double eps = eps_arr[blockIdx.x];
  // max(pU) -> step
  float pmax = 0.f;
#pragma unroll 4
  for (int64_t i = a + threadIdx.x; i < b; i += FB_THREADS) pmax = fmaxf(pmax, pU ? pU[i] : 1.f);
  pmax = wave_max(pmax);
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = pmax;
  __syncthreads();
  pmax = 0.f;
  for (int w = 0; w < FB_THREADS / 64; ++w) pmax = fmaxf(pmax, shm[w]);
  const float pm = pmax > 0.f ? pmax : 1.f;
  const float inv_pm = 1.f / pm;
  // a = pU / max pU as one multiply; exactly 1 where pU is maximal
  auto step = [&](float p) { return p == pm ? 1.f : p * inv_pm; };
  // gradient step z - mu pU (z - w), mu = 1/max pU, as (1-a) z + a w with
  // a = pU / max pU (a == 1 exactly where pU is maximal, so a constant pU
  // reproduces w bit for bit and the iteration stops at its fixed point).
  // Pass 1: w into v2, first gradient-step distance to y.
  double s1[1] = {0.0};
#pragma unroll 4
  for (int64_t i = a + threadIdx.x; i < b; i += FB_THREADS) {
    const float p = pU ? pU[i] : 1.f;
    const float2 vv = v2[i], rr = r2[i], zz = z[i], yy = y[i];
    const float2 w = make_float2(vv.x / p + rr.x, vv.y / p + rr.y);
    v2[i] = w;
    const float aa = step(p);
    const float2 g = make_float2((1.f - aa) * zz.x + aa * w.x, (1.f - aa) * zz.y + aa * w.y);
    const double dx = (double)g.x - yy.x, dy = (double)g.y - yy.y;
    s1[0] += dx * dx + dy * dy;
  }
  block_sum<1>(s1, sh);
  double s_cur = s1[0];
  for (int it = 1; it <= max_it; ++it) {
    const double nrm = sqrt(s_cur);
    const float scale = nrm > 0.0 ? (float)fmin(eps / nrm, 1.0) : 1.f;
    double acc[3] = {0.0, 0.0, 0.0};  // ||dz||^2, ||z||^2, next ||g - y||^2
#pragma unroll 4
    for (int64_t i = a + threadIdx.x; i < b; i += FB_THREADS) {
      const float p = pU ? pU[i] : 1.f;
      const float2 w = v2[i], zz = z[i], yy = y[i];
      const float aa = step(p);
      const float2 g = make_float2((1.f - aa) * zz.x + aa * w.x, (1.f - aa) * zz.y + aa * w.y);
      const float2 zn = make_float2(yy.x + (g.x - yy.x) * scale, yy.y + (g.y - yy.y) * scale);
      z[i] = zn;
      const double ddx = (double)zn.x - zz.x, ddy = (double)zn.y - zz.y;
      acc[0] += ddx * ddx + ddy * ddy;
      acc[1] += (double)zn.x * zn.x + (double)zn.y * zn.y;
      const float2 g2 = make_float2((1.f - aa) * zn.x + aa * w.x, (1.f - aa) * zn.y + aa * w.y);
      const double ex = (double)g2.x - yy.x, ey = (double)g2.y - yy.y;
      acc[2] += ex * ex + ey * ey;
    }
    block_sum<3>(acc, sh);
    const double crit = acc[1] > 0.0 ? sqrt(acc[0] / acc[1]) : 0.0;
    if (it >= min_it && crit <= tol) break;
    s_cur = acc[2];
  }
  // dual update v2 = pU (w - z) and residual ||G x - y||
  double res[1] = {0.0};
#pragma unroll 4
  for (int64_t i = a + threadIdx.x; i < b; i += FB_THREADS) {
    const float p = pU ? pU[i] : 1.f;
    const float2 w = v2[i], zz = z[i], yy = y[i], gg = gx[i];
    v2[i] = make_float2(p * (w.x - zz.x), p * (w.y - zz.y));
    const double ex = (double)gg.x - yy.x, ey = (double)gg.y - yy.y;
    res[0] += ex * ex + ey * ey;
  }
  block_sum<1>(res, sh);
  if (threadIdx.x == 0) norm_res[blockIdx.x] = res[0];
}

// DR dual step: v2 <- sc(v2 + r2 - y, eps) (closed-form l2 prox), residual
// ||gx - y||; sc(z, r) = z (1 - min(r/||z||, 1)).
__global__ __launch_bounds__(FB_THREADS) void k_l2_dual(const int64_t* __restrict__ blk_start,
                                                        const double* __restrict__ eps_arr,
                                                        float2* __restrict__ v2,
                                                        const float2* __restrict__ y,
                                                        const float2* __restrict__ r2,
                                                        const float2* __restrict__ gx,
                                                        double* __restrict__ norm_res) {
  __shared__ double sh[2 * FB_THREADS / 64];
  const int64_t a = blk_start[blockIdx.x], b = blk_start[blockIdx.x + 1];
  const double eps = eps_arr[blockIdx.x];
  double acc[2] = {0.0, 0.0};
  for (int64_t i = a + threadIdx.x; i < b; i += FB_THREADS) {
    const float2 vv = v2[i], rr = r2[i], yy = y[i], gg = gx[i];
    const double zx = (double)vv.x + rr.x - yy.x, zy = (double)vv.y + rr.y - yy.y;
    acc[0] += zx * zx + zy * zy;
    const double ex = (double)gg.x - yy.x, ey = (double)gg.y - yy.y;
    acc[1] += ex * ex + ey * ey;
  }
  block_sum<2>(acc, sh);
  const double nz = sqrt(acc[0]);
  const float f = nz > 0.0 ? (float)(1.0 - fmin(eps / nz, 1.0)) : 0.f;
  for (int64_t i = a + threadIdx.x; i < b; i += FB_THREADS) {
    const float2 vv = v2[i], rr = r2[i], yy = y[i];
    v2[i] = make_float2((vv.x + rr.x - yy.x) * f, (vv.y + rr.y - yy.y) * f);
  }
  if (threadIdx.x == 0) norm_res[blockIdx.x] = acc[1];
}

// g2_l = G_l' (lam .* v2_l) (+ G_l (lam .* v2_l) in the DR form) for the
// channels whose G keeps a CSC copy: one workgroup per adjoint work unit;
// grid points past a channel's columns are zero.
__global__ __launch_bounds__(ADJ_THREADS) void k_data_adjoint(const GDev* __restrict__ gd,
                                                              const int64_t* __restrict__ row_off,
                                                              const float2* __restrict__ v2,
                                                              const float* __restrict__ lam, int dr,
                                                              int64_t No, float2* __restrict__ g2, int L,
                                                              int units) {
  __shared__ float2 prod[ADJ_LDS];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (g.tiled) return;
  float2* out = g2 + (size_t)l * No;
  if (u >= g.nunits) {
    // zero the grid tail beyond the matrix columns (only when ncols < No)
    const int64_t extra_units = (No - g.ncols + ADJ_COLS - 1) / ADJ_COLS;
    const int64_t t = u - g.nunits;
    if (t < extra_units) {
      const int64_t c0 = g.ncols + t * ADJ_COLS;
      for (int64_t c = c0 + threadIdx.x; c < c0 + ADJ_COLS && c < No; c += ADJ_THREADS)
        out[c] = make_float2(0.f, 0.f);
    }
    return;
  }
  const float2* v = v2 + row_off[l];
  const float* lm = lam ? lam + row_off[l] : nullptr;
  auto none = [](int) { return make_float2(0.f, 0.f); };
  auto wl = [&](int r) { return cscale(v[r], lm[r]); };
  auto wp = [&](int r) { return v[r]; };
  if (lm) {
    if (dr)
      g_adjoint_unit(g, u, wl, [&](int c) { return g_row_apply(g, c, wl); }, out, prod);
    else
      g_adjoint_unit(g, u, wl, none, out, prod);
  } else {
    if (dr)
      g_adjoint_unit(g, u, wp, [&](int c) { return g_row_apply(g, c, wp); }, out, prod);
    else
      g_adjoint_unit(g, u, wp, none, out, prod);
  }
}

// g2_l = G_l' (lam .* v2_l) for the channels with a tile-gridded separable G
// (one workgroup per grid tile; never the DR form, whose H is square).
template <bool LAM>
__global__ __launch_bounds__(ADJ_THREADS) void k_data_adjoint_tiled(const GDev* __restrict__ gd,
                                                                    const int64_t* __restrict__ row_off,
                                                                    const float2* __restrict__ v2,
                                                                    const float* __restrict__ lam, int64_t No,
                                                                    float2* __restrict__ g2, int L, int units) {
  extern __shared__ __align__(16) unsigned char t[];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nunits) return;
  const float2* __restrict__ v = v2 + row_off[l];
  const float* __restrict__ lm = LAM ? lam + row_off[l] : nullptr;
  if (LAM)
    g_adjoint_tile(g, u, [=](int r) { return cscale(v[r], lm[r]); }, g2 + (size_t)l * No, t);
  else
    g_adjoint_tile(g, u, [=](int r) { return v[r]; }, g2 + (size_t)l * No, t);
}

// sums of the partial tiles of the tiles split over several work items
__global__ __launch_bounds__(ADJ_THREADS) void k_data_adjoint_reduce(const GDev* __restrict__ gd, int64_t No,
                                                                     float2* __restrict__ g2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nred) return;
  g_adjoint_tile_reduce(g, u, g2 + (size_t)l * No);
}

void launch_data_adjoint(fhs_ctx* ctx, const fhs_data* D, const float2* v2, const float* lam, int dr) {
  const int L = D->L;
  if (D->tile_grid > 0) {
    if (lam)
      k_data_adjoint_tiled<true><<<xcd_grid(L, D->tile_grid), ADJ_THREADS, D->tile_lds, ctx->stream>>>(
          D->gdev, D->d_row_off, v2, lam, D->No, D->g2, L, D->tile_grid);
    else
      k_data_adjoint_tiled<false><<<xcd_grid(L, D->tile_grid), ADJ_THREADS, D->tile_lds, ctx->stream>>>(
          D->gdev, D->d_row_off, v2, nullptr, D->No, D->g2, L, D->tile_grid);
    check_launch();
    if (D->red_grid > 0) {
      k_data_adjoint_reduce<<<xcd_grid(L, D->red_grid), ADJ_THREADS, 0, ctx->stream>>>(D->gdev, D->No, D->g2, L,
                                                                                       D->red_grid);
      check_launch();
    }
  }
  if (D->adj_grid > 0) {
    k_data_adjoint<<<xcd_grid(L, D->adj_grid), ADJ_THREADS, 0, ctx->stream>>>(D->gdev, D->d_row_off, v2, lam, dr,
                                                                              D->No, D->g2, L, D->adj_grid);
    check_launch();
  }
}

// y = lam .* (G_l A x_l) for all channels (stacked rows)
__global__ __launch_bounds__(256) void k_data_apply(const GDev* __restrict__ gd,
                                                    const int64_t* __restrict__ row_off,
                                                    const float2* __restrict__ Fx, int64_t No,
                                                    const float* __restrict__ lam,
                                                    float2* __restrict__ y, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (g.tiled) return;   // k_data_apply_tiled
  const int m = u * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const int64_t i = row_off[l] + m;
  float2 v = g_row_dot(g, m, Fx + (size_t)l * No);
  if (lam) v = cscale(v, lam[i]);
  y[i] = v;
}

// the same two operations for the channels with a tile-gridded separable G:
// one workgroup per forward work item, grid window staged in LDS
__global__ __launch_bounds__(ADJ_THREADS) void k_data_forward_tiled(const GDev* __restrict__ gd,
                                                                    const int64_t* __restrict__ row_off,
                                                                    const float2* __restrict__ Fx, int64_t No,
                                                                    float2* __restrict__ gx_old,
                                                                    float2* __restrict__ r2, int L, int units) {
  __shared__ float2 win[FWD_WIN_MAX];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nfitems) return;
  float2* __restrict__ go = gx_old + row_off[l];
  float2* __restrict__ rr = r2 + row_off[l];
  g_forward_tile(
      g, u, Fx + (size_t)l * No, win, [=](int r) { return go[r]; },
      [=](int r, float2 gx, float2 old) {
        rr[r] = make_float2(2.f * gx.x - old.x, 2.f * gx.y - old.y);
        go[r] = gx;
      });
}

__global__ __launch_bounds__(ADJ_THREADS) void k_data_apply_tiled(const GDev* __restrict__ gd,
                                                                  const int64_t* __restrict__ row_off,
                                                                  const float2* __restrict__ Fx, int64_t No,
                                                                  const float* __restrict__ lam,
                                                                  float2* __restrict__ y, int L, int units) {
  __shared__ float2 win[FWD_WIN_MAX];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nfitems) return;
  float2* __restrict__ yo = y + row_off[l];
  const float* __restrict__ lm = lam ? lam + row_off[l] : nullptr;
  g_forward_tile(
      g, u, Fx + (size_t)l * No, win, [=](int r) { return lm ? lm[r] : 1.f; },
      [=](int r, float2 v, float s) { yo[r] = lm ? cscale(v, s) : v; });
}

// G Fx -> r2 = 2 G Fx - gx_old, gx_old <- G Fx (non-DR form) for all channels
void launch_data_forward(fhs_ctx* ctx, const fhs_data* D, float2* gx_old, unsigned row_grid, int row_units) {
  if (D->fwd_grid > 0) {
    k_data_forward_tiled<<<xcd_grid(D->L, D->fwd_grid), ADJ_THREADS, 0, ctx->stream>>>(
        D->gdev, D->d_row_off, D->Fx, D->No, gx_old, D->r2, D->L, D->fwd_grid);
    check_launch();
  }
  if (D->any_rows) {
    k_data_forward<<<row_grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, gx_old, D->r2, D->L,
                                                      row_units);
    check_launch();
  }
}


__global__ void k_sum_blocks(const double* __restrict__ norm_res, const double* __restrict__ eps,
                             int nb, double* __restrict__ out2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0, e = 0.0;
    for (int i = 0; i < nb; ++i) {
      s += norm_res[i];
      e += eps[i] * eps[i];
    }
    out2[0] = s;
    out2[1] = e;
  }
}

template <class T>
T* dev_copy(const T* h, size_t n) {
  T* d = nullptr;
  FHS_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) FHS_HIP(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

// Device-resident step: all inputs/outputs on the device; norms (squared) in
// data->d_norm_res, sums in d_out2 = {sum norm_res^2, sum eps^2}.
void data_step(fhs_ctx* ctx, fhs_data* D, float2* v2, const float2* y, const float* x,
               float2* gx_old, float2* proj, const float* pU, int max_it, int min_it, double tol,
               int dr, const float* lam, float* Ftx, double* d_out2) {
  const int L = D->L;
  nufft_forward(ctx, D->A, x, L, D->Fx);
  {
    Scope sc(ctx, "data_forward");
    const int units = (int)cdiv(D->maxM, 256);
    const unsigned grid = xcd_grid(L, units);
    if (dr)
      k_data_forward_dr<<<grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, gx_old, D->r2, L,
                                                       units);
    else
      launch_data_forward(ctx, D, gx_old, grid, units);
    check_launch();
  }
  {
    Scope sc(ctx, dr ? "l2_dual" : "fb_dual");
    if (dr)
      k_l2_dual<<<D->nblocks, FB_THREADS, 0, ctx->stream>>>(D->d_blk_start, D->d_eps, v2, y, D->r2, gx_old,
                                                            D->d_norm_res);
    else
      k_fb_dual<<<D->nblocks, FB_THREADS, 0, ctx->stream>>>(D->d_blk_start, D->d_eps, v2, y, pU, D->r2,
                                                            proj, gx_old, max_it, min_it, tol, D->d_norm_res);
    check_launch();
  }
  {
    Scope sc(ctx, "data_adjoint");
    launch_data_adjoint(ctx, D, v2, lam, dr);
  }
  nufft_adjoint(ctx, D->A, D->g2, L, D->d_sigma22, Ftx);
  if (d_out2) {
    k_sum_blocks<<<1, 64, 0, ctx->stream>>>(D->d_norm_res, D->d_eps, D->nblocks, d_out2);
    check_launch();
  }
}

void data_set_sigma22(fhs_ctx* ctx, fhs_data* D, const float* h_sigma22) {
  FHS_HIP(hipMemcpyAsync(D->d_sigma22, h_sigma22, D->L * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
}

void data_set_epsilon(fhs_ctx* ctx, fhs_data* D, const double* h_eps) {
  FHS_HIP(hipMemcpyAsync(D->d_eps, h_eps, D->nblocks * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
}

void data_init_gx(fhs_ctx* ctx, fhs_data* D, const float* x, float2* gx_old, int dr, const float* lam) {
  // gx_old = G A(x): run the forward with a zero previous state, keep gx
  nufft_forward(ctx, D->A, x, D->L, D->Fx);
  FHS_HIP(hipMemsetAsync(gx_old, 0, D->Mtot * sizeof(float2), ctx->stream));
  const int units = (int)cdiv(D->maxM, 256);
  const unsigned grid = xcd_grid(D->L, units);
  if (dr)
    k_data_forward_dr<<<grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, gx_old, D->r2,
                                                     D->L, units);
  else
    launch_data_forward(ctx, D, gx_old, grid, units);
  check_launch();
}

// y = Lambda .* G A x (all channels)
void data_forward(fhs_ctx* ctx, fhs_data* D, const float* x, const float* lam, float2* y) {
  nufft_forward(ctx, D->A, x, D->L, D->Fx);
  Scope sc(ctx, "data_forward");
  const int units = (int)cdiv(D->maxM, 256);
  if (D->fwd_grid > 0) {
    k_data_apply_tiled<<<xcd_grid(D->L, D->fwd_grid), ADJ_THREADS, 0, ctx->stream>>>(
        D->gdev, D->d_row_off, D->Fx, D->No, lam, y, D->L, D->fwd_grid);
    check_launch();
  }
  if (D->any_rows)
    k_data_apply<<<xcd_grid(D->L, units), 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, y, D->L,
                                                                 units);
  check_launch();
}

// out_l = scale_l real(At(G_l' (lam .* v)))
void data_adjoint(fhs_ctx* ctx, fhs_data* D, const float2* v, const float* lam, const float* scale_dev,
                  float* out) {
  {
    Scope sc(ctx, "data_adjoint");
    launch_data_adjoint(ctx, D, v, lam, 0);
  }
  nufft_adjoint(ctx, D->A, D->g2, D->L, scale_dev, out);
}

void data_get_norm_res(fhs_ctx* ctx, fhs_data* D, double* h_out) {
  std::vector<double> nr(D->nblocks);
  FHS_HIP(hipMemcpyAsync(nr.data(), D->d_norm_res, D->nblocks * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  FHS_HIP(hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < D->nblocks; ++b) h_out[b] = std::sqrt(nr[b]);
}

int data_nblocks(const fhs_data* D) { return D->nblocks; }
bool data_square(const fhs_data* D) { return D->square; }
int data_channels(const fhs_data* D) { return D->L; }

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_data_create(fhs_ctx* ctx, const fhs_nufft* A, int L, const fhs_gmat* const* G,
                    int nblocks, const int64_t* block_start, const int* block_channel,
                    fhs_data** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(A && G && L >= 1 && L <= 65535 && nblocks >= L, "data_create: bad argument");
    FHS_HIP(hipSetDevice(ctx->device));
    fhs_data* D = new fhs_data();
    D->device = ctx->device;
    D->L = L;
    D->A = A;
    D->nblocks = nblocks;
    D->No = (int64_t)A->Kty * A->Ktx;
    D->row_off.assign(L + 1, 0);
    std::vector<GDev> gd(L);
    bool square = true;
    for (int l = 0; l < L; ++l) {
      FHS_CHECK(G[l] != nullptr, "data_create: null G");
      FHS_CHECK(G[l]->ncols <= D->No, "data_create: G has more columns than the FFT grid");
      gd[l] = G[l]->d;
      D->row_off[l + 1] = D->row_off[l] + G[l]->M;
      D->maxM = std::max<int>(D->maxM, (int)G[l]->M);
      if (G[l]->d.tiled) {
        D->tile_grid = std::max<int>(D->tile_grid, G[l]->d.nunits);
        D->red_grid = std::max<int>(D->red_grid, G[l]->d.nred);
        D->tile_lds = std::max(D->tile_lds, tile_lds_bytes(G[l]->d.Ky, G[l]->d.Kx));
        D->fwd_grid = std::max<int>(D->fwd_grid, G[l]->d.nfitems);
      } else {
        D->any_rows = true;
        const int64_t tail = (D->No - G[l]->ncols + ADJ_COLS - 1) / ADJ_COLS;
        D->adj_grid = std::max<int>(D->adj_grid, G[l]->d.nunits + (int)tail);
      }
      square = square && (G[l]->ncols == G[l]->M) && !G[l]->d.tiled;   // H of the DR form is not tiled
    }
    D->square = square;
    D->Mtot = D->row_off[L];
    D->blk_start.assign(block_start, block_start + nblocks + 1);
    D->blk_chan.assign(block_channel, block_channel + nblocks);
    FHS_CHECK(D->blk_start[0] == 0 && D->blk_start[nblocks] == D->Mtot,
              "data_create: blocks must cover all rows of all channels");
    for (int b = 0; b < nblocks; ++b) {
      const int c = D->blk_chan[b];
      FHS_CHECK(c >= 0 && c < L, "data_create: block channel out of range");
      FHS_CHECK(D->blk_start[b] < D->blk_start[b + 1], "data_create: empty block");
      FHS_CHECK(D->blk_start[b] >= D->row_off[c] && D->blk_start[b + 1] <= D->row_off[c + 1],
                "data_create: a block straddles channels");
    }
    D->gdev = dev_copy(gd.data(), L);
    D->d_row_off = dev_copy(D->row_off.data(), L + 1);
    D->d_blk_start = dev_copy(D->blk_start.data(), nblocks + 1);
    FHS_HIP(hipMalloc(&D->d_eps, nblocks * sizeof(double)));
    FHS_HIP(hipMemset(D->d_eps, 0, nblocks * sizeof(double)));
    FHS_HIP(hipMalloc(&D->d_norm_res, nblocks * sizeof(double)));
    FHS_HIP(hipMalloc(&D->d_sigma22, L * sizeof(float)));
    FHS_HIP(hipMalloc(&D->Fx, (size_t)L * D->No * sizeof(float2)));
    FHS_HIP(hipMalloc(&D->g2, (size_t)L * D->No * sizeof(float2)));
    FHS_HIP(hipMalloc(&D->r2, (size_t)D->Mtot * sizeof(float2)));
    *out = D;
  });
}

int fhs_data_destroy(fhs_data* D) {
  if (!D) return FHS_OK;
  (void)hipSetDevice(D->device);
  void* ps[] = {D->gdev, D->d_row_off, D->d_blk_start, D->d_eps, D->d_norm_res,
                D->d_sigma22, D->Fx, D->g2, D->r2};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete D;
  return FHS_OK;
}

int fhs_data_forward(fhs_ctx* ctx, fhs_data* D, const float* x, const float* Lambda, void* y) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && x && y, "data_forward: null argument");
    data_forward(ctx, D, x, Lambda, (float2*)y);
  });
}

int fhs_data_adjoint(fhs_ctx* ctx, fhs_data* D, const void* y, const float* Lambda, const float* scale,
                     float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && y && out, "data_adjoint: null argument");
    float* sd = nullptr;
    if (scale) {
      sd = (float*)ctx->ws.get("data_adj_scale", D->L * sizeof(float));
      FHS_HIP(hipMemcpyAsync(sd, scale, D->L * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    }
    data_adjoint(ctx, D, (const float2*)y, Lambda, sd, out);
  });
}

int fhs_data_init_fx_old(fhs_ctx* ctx, fhs_data* D, const float* x, const float* Lambda,
                         int flag_dimensionality_reduction, void* Gx_old) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && x && Gx_old, "data_init_fx_old: bad argument");
    FHS_CHECK(!flag_dimensionality_reduction || D->square, "data_init_fx_old: DR needs square G");
    data_init_gx(ctx, D, x, (float2*)Gx_old, flag_dimensionality_reduction, Lambda);
  });
}

int fhs_update_dual_data_fidelity(fhs_ctx* ctx, fhs_data* D, void* v2, const void* y,
                                  const float* x, void* Gx_old, void* proj, const float* pU,
                                  const double* epsilon, int elipse_proj_max_iter,
                                  int elipse_proj_min_iter, double elipse_proj_eps,
                                  const float* sigma22, int flag_dimensionality_reduction,
                                  const float* Lambda, float* Ftx, double* norm_res,
                                  double* global_norm_res, double* norm_epsilon) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && v2 && y && x && Gx_old && Ftx && epsilon && sigma22,
              "update_dual_data_fidelity: null argument");
    const int dr = flag_dimensionality_reduction != 0;
    FHS_CHECK(dr || proj, "update_dual_data_fidelity: proj required without DR");
    FHS_CHECK(!dr || D->square, "update_dual_data_fidelity: DR needs square (holographic) G");
    FHS_CHECK(elipse_proj_max_iter >= 1 && elipse_proj_min_iter >= 0,
              "update_dual_data_fidelity: bad ellipsoid iteration bounds");
    data_set_epsilon(ctx, D, epsilon);
    data_set_sigma22(ctx, D, sigma22);
    double* d2 = (double*)ctx->ws.get("data_out2", 2 * sizeof(double));
    data_step(ctx, D, (float2*)v2, (const float2*)y, x, (float2*)Gx_old, (float2*)proj, pU,
              elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps, dr, Lambda, Ftx, d2);
    std::vector<double> nr(D->nblocks);
    double h2[2];
    FHS_HIP(hipMemcpyAsync(nr.data(), D->d_norm_res, D->nblocks * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    FHS_HIP(hipMemcpyAsync(h2, d2, sizeof(h2), hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
    if (norm_res)
      for (int b = 0; b < D->nblocks; ++b) norm_res[b] = std::sqrt(nr[b]);
    if (global_norm_res) *global_norm_res = h2[0];
    if (norm_epsilon) *norm_epsilon = h2[1];
  });
}

}  // extern "C"
