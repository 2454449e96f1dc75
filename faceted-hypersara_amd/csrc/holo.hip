// Holographic (dimensionality-reduced) data of one channel:
// lib/operators/util_load_data_gen_dr_measurement_operator.m:170-225.
//
//   H = G'G over the oversampled grid (G: M visibilities x Kty*Ktx grid
//       points, separable min-max rows; the reference's G_curr * G_curr' with
//       G_curr = G', :167-168, :196)
//   W = {p : H_pp > 0}  (:203),  Lambda = 1 ./ sqrt(H_pp) on W  (:206)
//   stored: tril(H) restricted to W with its diagonal halved (:215-222)
//
// W is ascending, so the lower triangle in W's order is the lower triangle
// in grid order: H is kept as a square Kty*Ktx matrix whose rows and columns
// off W are empty, and Lambda on the whole grid (0 off W).  The data path
// applies H + H' (k_data_forward_dr) to the whole grid: every entry off W of
// its vectors is zero and stays zero, and the entries on W are the
// reference's compact vectors.
//
// H is banded: H_pq != 0 only if one window of Ky x Kx taps holds both p and
// q, i.e. q - p in [-(Ky-1), Ky-1] x [-(Kx-1), Kx-1] (with the grid's wrap).
// Per grid point p a wavefront evaluates that neighbourhood,
//   H[p, q] = sum over the rows m of column p of conj(G[m,p]) G[m,q]
// (G[m,q] = 0 when q is outside m's window), in FP64, rows ascending.  A
// count pass sizes row p of the lower triangle (q <= p) and column p of it
// (q >= p); the host scans the counts; a fill pass recomputes the same sums
// and writes the ELL rows and the CSC columns that apply_direct_operator /
// apply_adjoint_operator read, plus the full Hermitian stencil of every
// active mode, (2Ky-1)(2Kx-1) values offset-major, that the DR data kernels
// read: H + H' = G'G applied as a stencil, coalesced and without column
// indices.  No atomics: the result is deterministic.
#include <algorithm>
#include <numeric>

#include "gmat.hpp"

namespace fhs {
namespace {

constexpr int HOLO_WAVES = 4;                 // grid points per 256-thread workgroup
constexpr int HOLO_SLOTS = 4;                 // neighbourhood offsets per lane: (2*8-1)^2 = 225 <= 256

struct HoloGeom {
  int Kty, Ktx, Ky, Kx;
  int ny, nx;        // neighbourhood extent 2Ky-1, 2Kx-1
  int64_t No;
};

// H[p, q] for the lane's neighbourhood slots of grid point p (FP64 sums).
// nz[s]: slot s holds a nonzero (after rounding to FP32); q[s]: its grid
// index; diag: H_pp in the lane that holds the centre slot, 0 elsewhere.
__device__ __forceinline__ void holo_point(const GDev& g, const HoloGeom& h, int p, int lane, float2* val,
                                           int* q, bool* nz, double& diag) {
  const int py = p % h.Kty, px = p / h.Kty;
  const int no = h.ny * h.nx;
  double ar[HOLO_SLOTS], ai[HOLO_SLOTS];
  int dy[HOLO_SLOTS], dx[HOLO_SLOTS];
#pragma unroll
  for (int s = 0; s < HOLO_SLOTS; ++s) {
    ar[s] = ai[s] = 0.0;
    const int o = lane + 64 * s;
    dy[s] = o % h.ny - (h.Ky - 1);
    dx[s] = o / h.ny - (h.Kx - 1);
    const int qy = (py + dy[s] + h.Kty) % h.Kty, qx = (px + dx[s] + h.Ktx) % h.Ktx;
    q[s] = (o < no) ? qy + h.Kty * qx : -1;
  }
  const int e0 = g.colptr[p], e1 = g.colptr[p + 1];
  for (int e = e0; e < e1; ++e) {
    const unsigned pk = g.csc_pk[e];
    const int m = (int)(pk >> 6), t = (int)(pk & 63u);
    const int jx = t / g.Ky, jy = t - jx * g.Ky;
    const float2* row = g.vals + (size_t)m * 2 * g.kmax;
    const float2 gp = cmul(row[jy], row[g.kmax + jx]);
#pragma unroll
    for (int s = 0; s < HOLO_SLOTS; ++s) {
      const int ty = jy + dy[s], tx = jx + dx[s];
      if (q[s] >= 0 && ty >= 0 && ty < g.Ky && tx >= 0 && tx < g.Kx) {
        const float2 gq = cmul(row[ty], row[g.kmax + tx]);
        ar[s] += (double)gp.x * gq.x + (double)gp.y * gq.y;     // conj(gp) * gq
        ai[s] += (double)gp.x * gq.y - (double)gp.y * gq.x;
      }
    }
  }
  diag = 0.0;
#pragma unroll
  for (int s = 0; s < HOLO_SLOTS; ++s) {
    val[s] = make_float2((float)ar[s], (float)ai[s]);
    nz[s] = q[s] >= 0 && (val[s].x != 0.f || val[s].y != 0.f);
    if (q[s] == p) diag = ar[s];
  }
}

// count pass: entries of lower row p (q <= p) and of lower column p (q >= p);
// Lambda_p = 1/sqrt(H_pp) (0 off W)
__global__ __launch_bounds__(256) void k_holo_count(GDev g, HoloGeom h, int* __restrict__ nlo,
                                                    int* __restrict__ nup, float* __restrict__ lam) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * HOLO_WAVES + (threadIdx.x >> 6);
  if (p >= h.No) return;
  float2 val[HOLO_SLOTS];
  int q[HOLO_SLOTS];
  bool nz[HOLO_SLOTS];
  double d;
  holo_point(g, h, (int)p, lane, val, q, nz, d);
  int lo = 0, up = 0;
#pragma unroll
  for (int s = 0; s < HOLO_SLOTS; ++s) {
    lo += __popcll(__ballot(nz[s] && q[s] <= p));
    up += __popcll(__ballot(nz[s] && q[s] >= p));
  }
  // the centre slot (dy = dx = 0) sits in one lane: sum it over the wave
  for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
  if (lane == 0) {
    nlo[p] = lo;
    nup[p] = up;
    lam[p] = d > 0.0 ? (float)(1.0 / sqrt(d)) : 0.f;
  }
}

// fill pass: ELL row p of tril(H) (slots [nlo_p, K) padded with column p and
// value 0) and CSC column p (rows q >= p, values conj(H[p, q]) = H[q, p]);
// the diagonal entry is halved in both.  Active p also get their full
// stencil row (every neighbourhood slot, zeros included, diagonal whole).
__global__ __launch_bounds__(256) void k_holo_fill(GDev g, HoloGeom h, int K, const int* __restrict__ colptr,
                                                   int* __restrict__ ell_c, float2* __restrict__ ell_v,
                                                   int* __restrict__ csc_r, float2* __restrict__ csc_v,
                                                   const int* __restrict__ act, int nact,
                                                   float2* __restrict__ stencil) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * HOLO_WAVES + (threadIdx.x >> 6);
  if (p >= h.No) return;
  float2 val[HOLO_SLOTS];
  int q[HOLO_SLOTS];
  bool nz[HOLO_SLOTS];
  double d;
  holo_point(g, h, (int)p, lane, val, q, nz, d);
  const int ia = act[p];
  if (ia >= 0) {
#pragma unroll
    for (int s = 0; s < HOLO_SLOTS; ++s)
      if (q[s] >= 0) stencil[(size_t)(lane + 64 * s) * nact + ia] = val[s];
  }
  const int c1 = colptr[p + 1];
  const unsigned long long below = (1ull << lane) - 1ull;
  int lo = 0, up = colptr[p];
#pragma unroll
  for (int s = 0; s < HOLO_SLOTS; ++s) {
    const bool fl = nz[s] && q[s] <= p, fu = nz[s] && q[s] >= p;
    const unsigned long long bl = __ballot(fl), bu = __ballot(fu);
    float2 v = val[s];
    if (q[s] == p) v = make_float2(0.5f * v.x, 0.f);
    // both passes evaluate the same sums; the bounds only guard the buffers
    const int j = lo + __popcll(bl & below);
    if (fl && j < K) {
      ell_c[(size_t)j * h.No + p] = q[s];
      ell_v[(size_t)j * h.No + p] = v;
    }
    const int k = up + __popcll(bu & below);
    if (fu && k < c1) {
      csc_r[k] = q[s];
      csc_v[k] = make_float2(v.x, -v.y);
    }
    lo += __popcll(bl);
    up += __popcll(bu);
  }
  for (int j = lo + lane; j < K; j += 64) {
    ell_c[(size_t)j * h.No + p] = (int)p;
    ell_v[(size_t)j * h.No + p] = make_float2(0.f, 0.f);
  }
}

__global__ void k_holo_scale(const float* __restrict__ lam, float2* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = cscale(v[i], lam[i]);
}

template <class T>
T* dev_alloc(size_t n) {
  T* d = nullptr;
  FHS_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  return d;
}

template <class T>
T* dev_upload(const std::vector<T>& v) {
  T* d = dev_alloc<T>(v.size());
  if (!v.empty()) FHS_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

}  // namespace
}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_gmat_create_holographic(fhs_ctx* ctx, const fhs_gmat* G, fhs_gmat** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(G, "gmat_create_holographic: null G");
    FHS_CHECK(G->d.layout == G_SEP, "gmat_create_holographic: G must have separable (min-max) rows");
    FHS_CHECK(G->d.Ky * G->d.Kx <= 64, "gmat_create_holographic: at most 64 taps per row (8 x 8 windows)");
    FHS_CHECK(G->device == ctx->device, "gmat_create_holographic: G lives on another device");
    const GDev& gd = G->d;
    HoloGeom h{gd.Kty, gd.Ktx, gd.Ky, gd.Kx, 2 * gd.Ky - 1, 2 * gd.Kx - 1, (int64_t)gd.Kty * gd.Ktx};
    FHS_CHECK(gd.Kty >= h.ny && gd.Ktx >= h.nx,
              "gmat_create_holographic: the grid must be at least 2K-1 points on each axis");
    FHS_CHECK(h.ny * h.nx <= 64 * HOLO_SLOTS, "gmat_create_holographic: window too large");
    FHS_HIP(hipSetDevice(ctx->device));
    const int64_t M = G->M, No = h.No;
    std::vector<int32_t> anchor(M);
    FHS_HIP(hipMemcpy(anchor.data(), gd.anchor, M * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int> gcolptr;
    std::vector<unsigned> pk;
    sep_csc(M, anchor.data(), gd.Kty, gd.Ktx, gd.Ky, gd.Kx, gcolptr, pk);
    std::vector<void*> tmp;
    fhs_gmat* H = nullptr;
    try {
      GDev g = gd;
      g.colptr = dev_upload(gcolptr);
      tmp.push_back((void*)g.colptr);
      g.csc_pk = dev_upload(pk);
      tmp.push_back((void*)g.csc_pk);
      int* d_nlo = dev_alloc<int>(No);
      tmp.push_back(d_nlo);
      int* d_nup = dev_alloc<int>(No);
      tmp.push_back(d_nup);
      H = new fhs_gmat();
      H->device = ctx->device;
      float* d_lam = dev_alloc<float>(No);
      H->mem[13] = d_lam;
      const unsigned blocks = cdiv(No, HOLO_WAVES);
      k_holo_count<<<blocks, 64 * HOLO_WAVES, 0, ctx->stream>>>(g, h, d_nlo, d_nup, d_lam);
      check_launch();
      std::vector<int> nlo(No), nup(No);
      FHS_HIP(hipMemcpyAsync(nlo.data(), d_nlo, No * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      FHS_HIP(hipMemcpyAsync(nup.data(), d_nup, No * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      FHS_HIP(hipStreamSynchronize(ctx->stream));
      std::vector<int> colptr(No + 1, 0);
      int64_t nnz = 0, nnz_up = 0, modes = 0;
      int K = 1;
      for (int64_t p = 0; p < No; ++p) {
        nnz += nlo[p];
        nnz_up += nup[p];
        FHS_CHECK(nnz_up < (1ll << 31), "gmat_create_holographic: H has too many nonzeros");
        colptr[p + 1] = (int)nnz_up;
        K = std::max(K, nlo[p]);
        modes += nlo[p] > 0;
      }
      FHS_CHECK(nnz == nnz_up, "gmat_create_holographic: lower/upper counts disagree");
      FHS_CHECK(modes > 0, "gmat_create_holographic: G touches no grid point");
      std::vector<int> act(No);
      {
        int i = 0;
        for (int64_t p = 0; p < No; ++p) act[p] = nlo[p] > 0 ? i++ : -1;
      }
      int* d_act = dev_upload(act);
      H->mem[8] = d_act;
      float2* sten = dev_alloc<float2>((size_t)h.ny * h.nx * modes);
      H->mem[5] = sten;
      int* ell_c = dev_alloc<int>((size_t)K * No);
      H->mem[0] = ell_c;
      float2* ell_v = dev_alloc<float2>((size_t)K * No);
      H->mem[1] = ell_v;
      int* d_colptr = dev_upload(colptr);
      H->mem[2] = d_colptr;
      int* csc_r = dev_alloc<int>(nnz);
      H->mem[3] = csc_r;
      float2* csc_v = dev_alloc<float2>(nnz);
      H->mem[4] = csc_v;
      k_holo_fill<<<blocks, 64 * HOLO_WAVES, 0, ctx->stream>>>(g, h, K, d_colptr, ell_c, ell_v, csc_r, csc_v,
                                                                d_act, (int)modes, sten);
      check_launch();
      const std::vector<int> units = adjoint_units(colptr);
      H->mem[6] = dev_upload(units);
      FHS_HIP(hipStreamSynchronize(ctx->stream));
      H->M = No;
      H->ncols = No;
      H->nnz = nnz;
      H->n_modes = modes;
      GDev& d = H->d;
      d.layout = G_ELL;
      d.M = (int)No;
      d.K = K;
      d.ncols = (int)No;
      d.cols = ell_c;
      d.vals = ell_v;
      d.colptr = d_colptr;
      d.csc_rows = csc_r;
      d.csc_vals = csc_v;
      d.ucol = (const int*)H->mem[6];
      d.nunits = (int)units.size() - 1;
      d.Kty = h.Kty;
      d.Ktx = h.Ktx;
      d.Ky = h.Ky;
      d.Kx = h.Kx;
      d.stencil = sten;
      d.act = d_act;
      d.nact = (int)modes;
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      for (void* p : tmp) (void)hipFree(p);
      if (H) fhs_gmat_destroy(H);
      throw;
    }
    for (void* p : tmp) FHS_HIP(hipFree(p));
    *out = H;
  });
}

int fhs_gmat_holographic_info(fhs_ctx* ctx, const fhs_gmat* H, int64_t* n_modes, float* Lambda) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(H && H->n_modes > 0, "gmat_holographic_info: not a holographic matrix");
    if (n_modes) *n_modes = H->n_modes;
    if (Lambda)
      FHS_HIP(hipMemcpyAsync(Lambda, H->mem[13], H->M * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  });
}

int fhs_holographic_grid(fhs_ctx* ctx, const fhs_gmat* G, const fhs_gmat* H, const void* y, void* yh) {
  if (!ctx) return FHS_ERR_ARG;
  int st = guard(ctx, [&] {
    FHS_CHECK(G && H && y && yh, "holographic_grid: null argument");
    FHS_CHECK(H->n_modes > 0 && H->M == G->ncols, "holographic_grid: H is not the holographic form of G");
  });
  if (st != FHS_OK) return st;
  st = fhs_apply_adjoint_operator(ctx, G, y, nullptr, yh);
  if (st != FHS_OK) return st;
  return guard(ctx, [&] {
    k_holo_scale<<<cdiv(H->M, 256), 256, 0, ctx->stream>>>((const float*)H->mem[13], (float2*)yh, H->M);
    check_launch();
  });
}

}  // extern "C"
