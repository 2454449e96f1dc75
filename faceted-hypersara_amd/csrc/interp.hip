// Min-max NUFFT interpolation coefficients on the GPU (FP64 arithmetic,
// complex64 output), the per-visibility part of building G.
//
// Fessler & Sutton (IEEE TSP 2003) min-max interpolator with a cosine-series
// scaling s_n = sum_l a_l cos(gamma l (n - n0)): for frequency w on an axis
// of N pixels and K grid points, the J taps start at k0 + 1 with
// k0 = floor(w / gamma - J / 2), t_j = gamma (k0 + j) - w and
//   r_j = a_0 D(t_j) + sum_{l>=1} a_l / 2 [D(t_j + gamma l) + D(t_j - gamma l)],
//   D(t) = sin(N t / 2) / sin(t / 2)   (N where sin(t / 2) = 0),
//   c = T r  (T symmetric, the inverse of the J x J scaling Gram),
//   coef_j = c_j exp(i (gamma (k0 + j) n0 - w (n0 - floor(N / 2)))).
// The host supplies a_l and T (a least-squares fit and a 7x7 inverse, done
// once per axis); G row m is the Kronecker product cy[m] (x) cx[m] anchored at
// grid point (k0y + 1 mod Kty) + Kty (k0x + 1 mod Ktx).  Replaces the
// per-visibility loop of op_p_nufft (util_gen_measurement_operator_synth.m:85).
#include <cmath>

#include "common.hpp"

namespace fhs {
namespace {

constexpr int IJ_MAX = 8;       // taps per axis
constexpr int IA_MAX = 8;       // cosine-series terms

struct AxisDev {
  int N, K, J, na;
  double gamma, n0, shift;      // shift = n0 - floor(N / 2)
  double a[IA_MAX];
  double T[IJ_MAX * IJ_MAX];
};

__device__ __forceinline__ double dirichlet(double t, int N) {
  const double h = sin(0.5 * t);
  return fabs(h) < 1e-12 ? (double)N : sin(0.5 * (double)N * t) / h;
}

__device__ int axis_coef(const AxisDev& A, double w, double scale, float2* out) {
  const double k0 = floor(w / A.gamma - 0.5 * A.J);
  double r[IJ_MAX];
#pragma unroll
  for (int j = 0; j < IJ_MAX; ++j) {
    if (j >= A.J) break;
    const double t = A.gamma * (k0 + 1 + j) - w;
    double s = A.a[0] * dirichlet(t, A.N);
    for (int l = 1; l < A.na; ++l)
      s += 0.5 * A.a[l] * (dirichlet(t + A.gamma * l, A.N) + dirichlet(t - A.gamma * l, A.N));
    r[j] = s;
  }
#pragma unroll
  for (int i = 0; i < IJ_MAX; ++i) {
    if (i >= A.J) break;
    double c = 0.0;
    for (int j = 0; j < A.J; ++j) c += r[j] * A.T[j * A.J + i];
    c *= scale;
    const double ph = A.gamma * (k0 + 1 + i) * A.n0 - w * A.shift;
    double sp, cp;
    sincos(ph, &sp, &cp);
    out[i] = make_float2((float)(c * cp), (float)(c * sp));
  }
  long long k = (long long)k0 + 1;
  k %= A.K;
  if (k < 0) k += A.K;
  return (int)k;
}

// one thread per visibility: anchor, cy (J_y taps), cx (J_x taps, times row_scale)
__global__ __launch_bounds__(256) void k_interp(int M, const double* __restrict__ u, const double* __restrict__ v,
                                                const double* __restrict__ row_scale, AxisDev ay, AxisDev ax,
                                                int* __restrict__ anchor, float2* __restrict__ cy,
                                                float2* __restrict__ cx) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  float2 by[IJ_MAX], bx[IJ_MAX];
  const int iy = axis_coef(ay, v[m], 1.0, by);
  const int ix = axis_coef(ax, u[m], row_scale ? row_scale[m] : 1.0, bx);
  anchor[m] = iy + ay.K * ix;
  for (int j = 0; j < ay.J; ++j) cy[(int64_t)m * ay.J + j] = by[j];
  for (int j = 0; j < ax.J; ++j) cx[(int64_t)m * ax.J + j] = bx[j];
}

AxisDev axis_params(int N, int K, int J, int na, const double* a, const double* T) {
  FHS_CHECK(N >= 1 && K >= N && J >= 1 && J <= IJ_MAX && na >= 1 && na <= IA_MAX && a && T,
            "nufft_interp_coefficients: bad axis parameters (J <= 8, 1 <= na <= 8, K >= N)");
  AxisDev A{};
  A.N = N;
  A.K = K;
  A.J = J;
  A.na = na;
  A.gamma = 2.0 * M_PI / K;
  A.n0 = 0.5 * (N - 1);
  A.shift = A.n0 - (double)(N / 2);
  for (int l = 0; l < na; ++l) A.a[l] = a[l];
  for (int e = 0; e < J * J; ++e) A.T[e] = T[e];
  return A;
}

}  // namespace
}  // namespace fhs

using namespace fhs;

extern "C" int fhs_nufft_interp_coefficients(fhs_ctx* ctx, int Ny, int Kty, int Jy, int nay, const double* ay,
                                             const double* Ty, int Nx, int Ktx, int Jx, int nax, const double* ax,
                                             const double* Tx, const double* u, const double* v,
                                             const double* row_scale, int64_t M, int32_t* anchor, float* cy,
                                             float* cx) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(M >= 0 && M < (int64_t)1 << 31, "nufft_interp_coefficients: 0 <= M < 2^31");
    FHS_CHECK((int64_t)Kty * Ktx < (int64_t)1 << 31, "nufft_interp_coefficients: grid too large");
    const AxisDev Ay = axis_params(Ny, Kty, Jy, nay, ay, Ty);
    const AxisDev Ax = axis_params(Nx, Ktx, Jx, nax, ax, Tx);
    if (M == 0) return;
    FHS_CHECK(u && v && anchor && cy && cx, "nufft_interp_coefficients: null pointer");
    Scope sc(ctx, "nufft_interp");
    k_interp<<<cdiv(M, 256), 256, 0, ctx->stream>>>((int)M, u, v, row_scale, Ay, Ax, anchor, (float2*)cy,
                                                    (float2*)cx);
    check_launch();
  });
}
