// Uniform-weighting preconditioner on the GPU: aW = 1 / (number of uv
// points falling in the same cell of the oversampled Kty x Ktx grid), the
// diagonal of util_gen_preconditioning_matrix with gen_uniform_weight_matrix
// (imaging/default_parameters.json:47-51), computed per channel as in
// lib/operators/util_gen_measurement_operator_synth.m:73-74 (per data set in
// lib/operators/util_load_data_gen_measurement_operator.m:116).
// Cell of a point: floor((v + pi) Kty / (2 pi)) mod Kty along y, the same
// with u and Ktx along x, in FP64 in that order of operations.  The counts
// are integer atomics, so the result does not depend on their order.
#include <cmath>

#include "common.hpp"

namespace fhs {
namespace {

__global__ __launch_bounds__(256) void k_uw_count(const double* __restrict__ u, const double* __restrict__ v,
                                                  int64_t M, int Kty, int Ktx, int* __restrict__ cnt,
                                                  int* __restrict__ cell) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= M) return;
  int64_t iy = (int64_t)floor((v[m] + M_PI) * (double)Kty / (2.0 * M_PI)) % Kty;
  int64_t ix = (int64_t)floor((u[m] + M_PI) * (double)Ktx / (2.0 * M_PI)) % Ktx;
  if (iy < 0) iy += Kty;
  if (ix < 0) ix += Ktx;
  const int c = (int)(iy + (int64_t)Kty * ix);
  cell[m] = c;
  atomicAdd(cnt + c, 1);
}

__global__ __launch_bounds__(256) void k_uw_apply(const int* __restrict__ cell, const int* __restrict__ cnt,
                                                  int64_t M, float* __restrict__ aW) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= M) return;
  aW[m] = (float)(1.0 / (double)cnt[cell[m]]);
}

}  // namespace
}  // namespace fhs

using namespace fhs;

extern "C" int fhs_precond_uniform_weights(fhs_ctx* ctx, const double* u, const double* v, int64_t M, int Kty,
                                           int Ktx, float* aW) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(M >= 0 && Kty >= 1 && Ktx >= 1 && (int64_t)Kty * Ktx < ((int64_t)1 << 31),
              "precond_uniform_weights: 0 <= M, 1 <= Kty, Ktx and Kty*Ktx < 2^31");
    if (M == 0) return;
    FHS_CHECK(u && v && aW, "precond_uniform_weights: null pointer");
    const size_t ncell = (size_t)Kty * Ktx;
    int* cnt = (int*)ctx->ws.get("uw_count", ncell * sizeof(int));
    int* cell = (int*)ctx->ws.get("uw_cell", (size_t)M * sizeof(int));
    FHS_HIP(hipMemsetAsync(cnt, 0, ncell * sizeof(int), ctx->stream));
    Scope sc(ctx, "precond_weights");
    k_uw_count<<<cdiv(M, 256), 256, 0, ctx->stream>>>(u, v, M, Kty, Ktx, cnt, cell);
    check_launch();
    k_uw_apply<<<cdiv(M, 256), 256, 0, ctx->stream>>>(cell, cnt, M, aW);
    check_launch();
  });
}
