// Faceted SARA transform: Psit / Psi of one non-overlapping facet, exact
// pieces of the whole-image transform (the role of sdwt2_setup,
// sdwt2_sara_faceting and isdwt2_sara_faceting in the reference,
// src/fhs/facetHyperSARA.m:325-326, src/fhs/fhs_update_dual_sparsity.m:71-93;
// those live in a submodule absent from the reference tree).
//
// Ownership (per axis, per filter length lf): the level-0 ranges are the
// facets' pixels; at level j a facet owns the coefficients k whose last input
// sample 2k + 1 lies in its level-(j - 1) range, so the boundaries are
// t = floor(s / 2) of the level-(j - 1) boundaries s, and the last facet also
// owns the zpd right-border coefficients.  The owned coefficients at level j
// read level-(j - 1) samples from 2 t + 2 - lf on, so a facet reads the
// pixels from m_0 with m_J = t_J, m_(j-1) = 2 m_j + 2 - lf: its extended
// image starts there (the smallest over the bases), rounded down to a multiple
// of 2^J so that the extended image decimates in phase with the whole one.
// The transform of the extended image (zpd outside it) then gives every owned
// coefficient exactly, shifted by origin / 2^j at level j; its adjoint on the
// owned coefficients gives the facet's share of the whole-image synthesis.
//
// Each facet keeps a dictionary of its extended size (the same bases and
// levels) and the index of every owned coefficient in that dictionary's
// vector; analysis = extended-image analysis + gather, synthesis = scatter
// into zeros + extended-image synthesis.  Facet coefficient order: per basis
// the owned tiles of A_J, then H_j, V_j, D_j for j = J .. 1 (column-major),
// the Dirac part (the owned pixels) last; scaled by 1/sqrt(P) like the
// whole-image dictionary.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace fhs {
namespace {

__global__ void k_facet_gather(const float* __restrict__ full, int64_t s_full, const int64_t* __restrict__ idx,
                               int64_t n, int L, float* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * L; i += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(i / n);
    const int64_t k = i - (int64_t)l * n;
    out[i] = full[(int64_t)l * s_full + idx[k]];
  }
}

__global__ void k_facet_zero(float* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.f;
}

__global__ void k_facet_scatter(const float* __restrict__ in, int64_t n, const int64_t* __restrict__ idx,
                                int64_t s_full, int L, float* __restrict__ full) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * L; i += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(i / n);
    const int64_t k = i - (int64_t)l * n;
    full[(int64_t)l * s_full + idx[k]] = in[i];
  }
}

// w = v1 + z; v1 <- w - max(||w||_l - beta weights, 0) / (||w||_l + eps) w
// (src/fhs/fhs_update_dual_sparsity.m:83-87), row norms over the L channels
// in FP64
__global__ __launch_bounds__(256) void k_facet_l21(float* __restrict__ v1, const float* __restrict__ z, int64_t n,
                                                   int L, const float* __restrict__ weights, double beta) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    double ss = 0.0;
    for (int l = 0; l < L; ++l) {
      const float w = v1[(int64_t)l * n + c] + z[(int64_t)l * n + c];
      ss += (double)w * w;
    }
    const double l2 = sqrt(ss);
    const double t = beta * (weights ? (double)weights[c] : 1.0);
    const float f = (float)(fmax(l2 - t, 0.0) / (l2 + 2.220446049250313e-16));
    for (int l = 0; l < L; ++l) {
      const float w = v1[(int64_t)l * n + c] + z[(int64_t)l * n + c];
      v1[(int64_t)l * n + c] = w - f * w;
    }
  }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>(cdiv(n, 256), 1), 4096); }

int level_len(int n, int lf) { return (n + lf - 1) / 2; }

}  // namespace
}  // namespace fhs

struct fhs_sara_facets {
  int device = 0;
  int Qy = 0, Qx = 0, nlevel = 0, nb = 0, dirac = 0;
  struct Facet {
    int ey0 = 0, ex0 = 0, eny = 0, enx = 0;   // extended image (origin may be negative: zeros)
    int64_t n = 0;                            // owned coefficients
    int64_t s_sub = 0;                        // coefficients of the extended-size dictionary
    fhs_sara* sub = nullptr;
    int64_t* d_idx = nullptr;
    std::vector<int> ncoefs;                  // (nb (J + 1) + dirac) rows of (rows, cols)
  };
  std::vector<Facet> f;
};

using namespace fhs;

extern "C" {

int fhs_sara_facets_create(fhs_ctx* ctx, int Ny, int Nx, int nlevel, int nbases, const int* filter_len,
                           const double* lo_r, int with_dirac, int Qy, int Qx, const int* ybounds,
                           const int* xbounds, fhs_sara_facets** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  fhs_sara_facets* F = new fhs_sara_facets();
  const int st = guard(ctx, [&] {
    FHS_CHECK(Qy >= 1 && Qx >= 1 && ybounds && xbounds, "sara_facets_create: bad facet grid");
    FHS_CHECK(ybounds[0] == 0 && ybounds[Qy] == Ny && xbounds[0] == 0 && xbounds[Qx] == Nx,
              "sara_facets_create: the bounds must run from 0 to the image size");
    for (int i = 0; i < Qy; ++i) FHS_CHECK(ybounds[i] < ybounds[i + 1], "sara_facets_create: empty facet row");
    for (int i = 0; i < Qx; ++i) FHS_CHECK(xbounds[i] < xbounds[i + 1], "sara_facets_create: empty facet column");
    FHS_CHECK(nlevel >= 1 && nlevel <= 6 && nbases >= 0 && nbases <= 8, "sara_facets_create: bad dictionary");
    F->device = ctx->device;
    F->Qy = Qy;
    F->Qx = Qx;
    F->nlevel = nlevel;
    F->nb = nbases;
    F->dirac = with_dirac ? 1 : 0;
    const int J = nlevel;
    // per basis, per axis: boundaries of every level
    auto bounds = [&](int N, const int* b0, int Qn, int lf) {
      std::vector<std::vector<int>> B(J + 1);
      B[0].assign(b0, b0 + Qn + 1);
      int n = N;
      for (int j = 1; j <= J; ++j) {
        n = level_len(n, lf);
        B[j].resize(Qn + 1);
        B[j][0] = 0;
        for (int i = 1; i < Qn; ++i) B[j][i] = B[j - 1][i] / 2;   // ceil((s - 1) / 2)
        B[j][Qn] = n;
      }
      return B;
    };
    std::vector<std::vector<std::vector<int>>> by(nbases), bx(nbases);
    for (int b = 0; b < nbases; ++b) {
      by[b] = bounds(Ny, ybounds, Qy, filter_len[b]);
      bx[b] = bounds(Nx, xbounds, Qx, filter_len[b]);
    }
    auto ext_start = [&](int tJ, int lf) {
      long long m = tJ;
      for (int j = 0; j < J; ++j) m = 2 * m + 2 - lf;
      const long long step = 1LL << J;
      long long q = m >= 0 ? m / step : -((-m + step - 1) / step);
      return (int)(q * step);
    };
    F->f.resize((size_t)Qy * Qx);
    for (int qx = 0; qx < Qx; ++qx)
      for (int qy = 0; qy < Qy; ++qy) {
        fhs_sara_facets::Facet& fc = F->f[(size_t)qx * Qy + qy];
        int ey0 = ybounds[qy], ex0 = xbounds[qx];
        for (int b = 0; b < nbases; ++b) {
          ey0 = std::min(ey0, ext_start(by[b][J][qy], filter_len[b]));
          ex0 = std::min(ex0, ext_start(bx[b][J][qx], filter_len[b]));
        }
        fc.ey0 = ey0;
        fc.ex0 = ex0;
        fc.eny = ybounds[qy + 1] - ey0;
        fc.enx = xbounds[qx + 1] - ex0;
        const int st2 = fhs_sara_create_host(ctx, fc.eny, fc.enx, nlevel, nbases, filter_len, lo_r, with_dirac,
                                             &fc.sub);
        if (st2 != FHS_OK) throw Error(st2, ctx->err);
        fhs_sara_info(fc.sub, &fc.s_sub);
        // owned coefficients: positions in the extended-size vector
        std::vector<int64_t> idx;
        int64_t boff = 0;
        for (int b = 0; b < nbases; ++b) {
          const int lf = filter_len[b];
          std::vector<int> sny(J + 1), snx(J + 1);
          sny[0] = fc.eny;
          snx[0] = fc.enx;
          for (int j = 1; j <= J; ++j) {
            sny[j] = level_len(sny[j - 1], lf);
            snx[j] = level_len(snx[j - 1], lf);
          }
          int64_t so = 0;
          auto tile = [&](int j) {
            const int r0 = by[b][j][qy], r1 = by[b][j][qy + 1];
            const int c0 = bx[b][j][qx], c1 = bx[b][j][qx + 1];
            const int dy = ey0 >> j, dx = ex0 >> j;   // exact: ey0, ex0 are multiples of 2^J
            for (int c = c0; c < c1; ++c)
              for (int r = r0; r < r1; ++r) idx.push_back(boff + so + (int64_t)(r - dy) + (int64_t)sny[j] * (c - dx));
            so += (int64_t)sny[j] * snx[j];
          };
          tile(J);   // A_J
          for (int j = J; j >= 1; --j)
            for (int sb = 0; sb < 3; ++sb) tile(j);   // H_j, V_j, D_j
          boff += so;
          for (int j = 1; j <= J; ++j) {
            fc.ncoefs.push_back(by[b][j][qy + 1] - by[b][j][qy]);
            fc.ncoefs.push_back(bx[b][j][qx + 1] - bx[b][j][qx]);
          }
          fc.ncoefs.push_back(by[b][J][qy + 1] - by[b][J][qy]);
          fc.ncoefs.push_back(bx[b][J][qx + 1] - bx[b][J][qx]);
        }
        if (with_dirac) {
          for (int c = xbounds[qx]; c < xbounds[qx + 1]; ++c)
            for (int r = ybounds[qy]; r < ybounds[qy + 1]; ++r)
              idx.push_back(boff + (int64_t)(r - ey0) + (int64_t)fc.eny * (c - ex0));
          fc.ncoefs.push_back(ybounds[qy + 1] - ybounds[qy]);
          fc.ncoefs.push_back(xbounds[qx + 1] - xbounds[qx]);
        }
        fc.n = (int64_t)idx.size();
        FHS_HIP(hipSetDevice(ctx->device));
        FHS_HIP(hipMalloc(&fc.d_idx, std::max<size_t>(1, idx.size()) * sizeof(int64_t)));
        FHS_HIP(hipMemcpy(fc.d_idx, idx.data(), idx.size() * sizeof(int64_t), hipMemcpyHostToDevice));
      }
  });
  if (st != FHS_OK) {
    for (auto& fc : F->f) {
      if (fc.sub) fhs_sara_destroy(fc.sub);
      if (fc.d_idx) (void)hipFree(fc.d_idx);
    }
    delete F;
    return st;
  }
  *out = F;
  return FHS_OK;
}

int fhs_sara_facets_destroy(fhs_sara_facets* F) {
  if (!F) return FHS_OK;
  (void)hipSetDevice(F->device);
  for (auto& fc : F->f) {
    if (fc.sub) fhs_sara_destroy(fc.sub);
    if (fc.d_idx) (void)hipFree(fc.d_idx);
  }
  delete F;
  return FHS_OK;
}

int fhs_sara_facets_info(const fhs_sara_facets* F, int q, int* ext, int64_t* n_coefs, int* ncoefs_rows) {
  if (!F || q < 0 || q >= (int)F->f.size()) return FHS_ERR_ARG;
  const auto& fc = F->f[q];
  if (ext) {
    ext[0] = fc.ey0;
    ext[1] = fc.ex0;
    ext[2] = fc.eny;
    ext[3] = fc.enx;
  }
  if (n_coefs) *n_coefs = fc.n;
  if (ncoefs_rows) std::copy(fc.ncoefs.begin(), fc.ncoefs.end(), ncoefs_rows);
  return FHS_OK;
}

int fhs_sara_facet_analysis(fhs_ctx* ctx, const fhs_sara_facets* F, int q, const float* x_ext, int L, float* z) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(F && q >= 0 && q < (int)F->f.size() && x_ext && z && L >= 1, "sara_facet_analysis: bad argument");
    const auto& fc = F->f[q];
    float* full = (float*)ctx->ws.get("sara_facet_full", (size_t)fc.s_sub * L * sizeof(float));
    const int st = fhs_sara_analysis(ctx, fc.sub, x_ext, L, full, 0);
    if (st != FHS_OK) throw Error(st, ctx->err);
    k_facet_gather<<<grid_for(fc.n * L), 256, 0, ctx->stream>>>(full, fc.s_sub, fc.d_idx, fc.n, L, z);
    check_launch();
  });
}

int fhs_sara_facet_synthesis(fhs_ctx* ctx, const fhs_sara_facets* F, int q, const float* z, int L, float scale,
                             float* g_ext) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(F && q >= 0 && q < (int)F->f.size() && z && g_ext && L >= 1, "sara_facet_synthesis: bad argument");
    const auto& fc = F->f[q];
    float* full = (float*)ctx->ws.get("sara_facet_full", (size_t)fc.s_sub * L * sizeof(float));
    k_facet_zero<<<grid_for(fc.s_sub * L), 256, 0, ctx->stream>>>(full, fc.s_sub * L);
    check_launch();
    k_facet_scatter<<<grid_for(fc.n * L), 256, 0, ctx->stream>>>(z, fc.n, fc.d_idx, fc.s_sub, L, full);
    check_launch();
    const int st = fhs_sara_synthesis(ctx, fc.sub, full, L, scale, g_ext);
    if (st != FHS_OK) throw Error(st, ctx->err);
  });
}

int fhs_update_dual_sparsity_facet(fhs_ctx* ctx, const fhs_sara_facets* F, int q, float* v1, const float* x_ext,
                                   int L, const float* weights1, double beta1, float* g_ext) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(F && q >= 0 && q < (int)F->f.size() && v1 && x_ext && g_ext && L >= 1,
              "update_dual_sparsity_facet: bad argument");
    const auto& fc = F->f[q];
    float* full = (float*)ctx->ws.get("sara_facet_full", (size_t)fc.s_sub * L * sizeof(float));
    float* z = (float*)ctx->ws.get("sara_facet_z", (size_t)std::max<int64_t>(fc.n, 1) * L * sizeof(float));
    int st = fhs_sara_analysis(ctx, fc.sub, x_ext, L, full, 0);
    if (st != FHS_OK) throw Error(st, ctx->err);
    k_facet_gather<<<grid_for(fc.n * L), 256, 0, ctx->stream>>>(full, fc.s_sub, fc.d_idx, fc.n, L, z);
    check_launch();
    k_facet_l21<<<grid_for(fc.n), 256, 0, ctx->stream>>>(v1, z, fc.n, L, weights1, beta1);
    check_launch();
    k_facet_zero<<<grid_for(fc.s_sub * L), 256, 0, ctx->stream>>>(full, fc.s_sub * L);
    check_launch();
    k_facet_scatter<<<grid_for(fc.n * L), 256, 0, ctx->stream>>>(v1, fc.n, fc.d_idx, fc.s_sub, L, full);
    check_launch();
    st = fhs_sara_synthesis(ctx, fc.sub, full, L, 1.f, g_ext);
    if (st != FHS_OK) throw Error(st, ctx->err);
  });
}

}  // extern "C"
