// W-projected rows of G (the w-term of wide-field interferometry) on the
// GPU: the per-visibility part of op_p_nufft_wproj_dde
// (lib/operators/util_load_data_gen_measurement_operator.m:119-121).
//
// With the Fresnel form of the w-phase, c(l, m) = exp(i pi w (l^2 + m^2)),
// the w-term multiplies the image by a separable chirp, so its Fourier
// kernel is an outer product and a w-projected row stays separable:
//   cy'[j] = (1/Kty) sum_i cy[i] Chat_y(w)[i + S - j],   j = 0..Jy + 2S - 1
// (and the same for x), anchored S grid points before the min-max row.
// Chat(w)[d] (|d| <= S) is the oversampled-grid DFT of the chirp extended
// smoothly over the padding (the host builds it for a fine grid of w values
// and truncates it to its energy fraction); a visibility's kernel is the
// linear interpolation between its two neighbouring w values (conjugate and
// mirrored for w < 0: Chat(-w)[d] = conj(Chat(w)[-d])).  Each axis' row is
// then cut to the largest taps holding `keep` of its energy.  FP64
// arithmetic, complex64 output.
#include <cmath>

#include "common.hpp"

namespace fhs {
namespace {

constexpr int WJ_MAX = 16;   // taps per axis of a w-projected row

struct WAxis {
  int K, J, S, nb;             // grid, min-max taps, halo taps, w values in the table
  double dw, keep;             // table step, energy fraction kept per axis
  const double2* tab;          // [nb][2S + 1]: Chat(b dw)[d], d = -S..S
};

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// one axis of one visibility: returns the new first grid index
__device__ int axis_wrow(const WAxis& A, const float2* __restrict__ c, int a_in, double w, float2* out) {
  const bool neg = w < 0.0;
  const double pos = fabs(w) / A.dw;
  int b = (int)pos;
  if (b > A.nb - 2) b = A.nb - 2;
  const double f = pos - b;
  const int Kw = A.J + 2 * A.S, W = 2 * A.S + 1;
  const double2* t0 = A.tab + (size_t)b * W;
  const double2* t1 = t0 + W;
  double2 g[WJ_MAX];
  double e[WJ_MAX];
  double etot = 0.0;
  for (int j = 0; j < Kw; ++j) {
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < A.J; ++i) {
      int d = i + A.S - j;               // kernel offset, zero outside [-S, S]
      if (d < -A.S || d > A.S) continue;
      if (neg) d = -d;
      const double2 k0 = t0[d + A.S], k1 = t1[d + A.S];
      double2 k = make_double2(k0.x + f * (k1.x - k0.x), k0.y + f * (k1.y - k0.y));
      if (neg) k.y = -k.y;
      const float2 ci = c[i];
      s = make_double2(s.x + ci.x * k.x - ci.y * k.y, s.y + ci.x * k.y + ci.y * k.x);
    }
    s.x /= A.K;
    s.y /= A.K;
    g[j] = s;
    e[j] = s.x * s.x + s.y * s.y;
    etot += e[j];
  }
  // the largest taps until `keep` of the energy (ties in index order)
  bool kept[WJ_MAX];
  for (int j = 0; j < Kw; ++j) kept[j] = A.keep >= 1.0;
  if (A.keep < 1.0) {
    double acc = 0.0;
    for (int n = 0; n < Kw && acc < A.keep * etot; ++n) {
      int best = -1;
      for (int j = 0; j < Kw; ++j)
        if (!kept[j] && (best < 0 || e[j] > e[best])) best = j;
      kept[best] = true;
      acc += e[best];
    }
  }
  for (int j = 0; j < Kw; ++j) out[j] = kept[j] ? make_float2((float)g[j].x, (float)g[j].y) : make_float2(0.f, 0.f);
  int k = a_in - A.S;
  k %= A.K;
  if (k < 0) k += A.K;
  return k;
}

__global__ __launch_bounds__(256) void k_wproj(int M, const int* __restrict__ anchor_in, const float2* __restrict__ cy,
                                               const float2* __restrict__ cx, const double* __restrict__ w, WAxis ay,
                                               WAxis ax, int* __restrict__ anchor_out, float2* __restrict__ cy_out,
                                               float2* __restrict__ cx_out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int a = anchor_in[m];
  const int wy = ay.J + 2 * ay.S, wx = ax.J + 2 * ax.S;
  const int iy = axis_wrow(ay, cy + (int64_t)m * ay.J, a % ay.K, w[m], cy_out + (int64_t)m * wy);
  const int ix = axis_wrow(ax, cx + (int64_t)m * ax.J, a / ay.K, w[m], cx_out + (int64_t)m * wx);
  anchor_out[m] = iy + ay.K * ix;
}

WAxis waxis(int K, int J, int S, int nb, double dw, double keep, const double2* tab) {
  FHS_CHECK(K >= 1 && J >= 1 && S >= 0 && J + 2 * S <= WJ_MAX && J + 2 * S <= K,
            "nufft_wproj_rows: need 1 <= J, 0 <= S and J + 2 S <= 16");
  FHS_CHECK(nb >= 2 && dw > 0.0 && keep > 0.0 && keep <= 1.0 && tab,
            "nufft_wproj_rows: need >= 2 table rows, dw > 0, 0 < keep <= 1");
  return WAxis{K, J, S, nb, dw, keep, tab};
}

}  // namespace
}  // namespace fhs

using namespace fhs;

extern "C" int fhs_nufft_wproj_rows(fhs_ctx* ctx, int Kty, int Ktx, int Jy, int Jx, int Sy, int Sx, int nbins,
                                    double dw, const double* table_y, const double* table_x, double keep_y,
                                    double keep_x, const double* w, int64_t M, const int32_t* anchor_in,
                                    const float* cy_in, const float* cx_in, int32_t* anchor_out, float* cy_out,
                                    float* cx_out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(M >= 0 && M < (int64_t)1 << 31, "nufft_wproj_rows: 0 <= M < 2^31");
    FHS_CHECK(table_y && table_x, "nufft_wproj_rows: null table");
    if (M == 0) return;
    FHS_CHECK(w && anchor_in && cy_in && cx_in && anchor_out && cy_out && cx_out, "nufft_wproj_rows: null pointer");
    // tables: host FP64 complex, [nbins][2S + 1] per axis, uploaded per call
    const size_t ny = (size_t)nbins * (2 * Sy + 1), nx = (size_t)nbins * (2 * Sx + 1);
    double2* d = (double2*)ctx->ws.get("wproj_tables", (ny + nx) * sizeof(double2));
    FHS_HIP(hipMemcpyAsync(d, table_y, ny * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
    FHS_HIP(hipMemcpyAsync(d + ny, table_x, nx * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
    const WAxis Ay = waxis(Kty, Jy, Sy, nbins, dw, keep_y, d);
    const WAxis Ax = waxis(Ktx, Jx, Sx, nbins, dw, keep_x, d + ny);
    Scope sc(ctx, "nufft_wproj");
    k_wproj<<<cdiv(M, 256), 256, 0, ctx->stream>>>((int)M, anchor_in, (const float2*)cy_in, (const float2*)cx_in,
                                                   w, Ay, Ax, anchor_out, (float2*)cy_out, (float2*)cx_out);
    check_launch();
    // the tables are read by the kernel: wait before the call returns (the
    // host arrays may be freed, and the workspace reused, right after)
    FHS_HIP(hipStreamSynchronize(ctx->stream));
  });
}
