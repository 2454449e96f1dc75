// Elementwise steps with FP64 reductions: proj_l2ball and the primal update.
//
// Reductions are two-level and deterministic: a fixed grid of RED_BLOCKS
// workgroups writes one partial per workgroup, a single workgroup sums them
// in a fixed order.
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace fhs {
namespace {

constexpr int RED_BLOCKS = 2048;
constexpr int RED_THREADS = 256;

__global__ __launch_bounds__(RED_THREADS) void k_sqdiff_partial(const float2* __restrict__ x,
                                                                const float2* __restrict__ y,
                                                                int64_t n, double* partial) {
  __shared__ double sh[RED_THREADS / 64];
  double acc[1] = {0.0};
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 d = csub(x[i], y[i]);
    acc[0] += (double)d.x * d.x + (double)d.y * d.y;
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_proj_apply(const float2* __restrict__ x,
                                                            const float2* __restrict__ y,
                                                            int64_t n, double eps,
                                                            const double* __restrict__ partial,
                                                            int npart, float2* __restrict__ p) {
  __shared__ double sh[RED_THREADS / 64];
  double s[1] = {0.0};
  for (int i = threadIdx.x; i < npart; i += blockDim.x) s[0] += partial[i];
  block_sum<1>(s, sh);
  const double nrm = sqrt(s[0]);
  const float scale = nrm > 0.0 ? (float)fmin(eps / nrm, 1.0) : 1.f;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 yy = y[i];
    const float2 d = csub(x[i], yy);
    p[i] = make_float2(yy.x + d.x * scale, yy.y + d.y * scale);
  }
}

// With `truth` (a reference cube of the same layout) the kernel also sums
// (truth - x)^2, so a solver can follow the reconstruction SNR without a
// separate pass over the cube.
__global__ __launch_bounds__(RED_THREADS) void k_primal(float* __restrict__ x,
                                                        const float* __restrict__ ga,
                                                        const float* __restrict__ gb,
                                                        const float* __restrict__ gc,
                                                        float* __restrict__ xhat, int64_t n,
                                                        const float* __restrict__ truth,
                                                        double* partial) {
  __shared__ double sh[3 * RED_THREADS / 64];
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float prev = x[i];
    float g = ga[i];
    if (gb) g += gb[i];
    if (gc) g += gc[i];
    const float xn = fmaxf(prev - g, 0.f);
    x[i] = xn;
    xhat[i] = 2.f * xn - prev;
    const double d = (double)prev - (double)xn;
    acc[0] += d * d;
    acc[1] += (double)xn * xn;
    if (truth) {
      const double e = (double)truth[i] - (double)xn;
      acc[2] += e * e;
    }
  }
  block_sum<3>(acc, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = acc[0];
    partial[gridDim.x + blockIdx.x] = acc[1];
    partial[2 * gridDim.x + blockIdx.x] = acc[2];
  }
}

// out[v] = sum of partial vector v, for v < nvec; vector 2 goes to out2
__global__ __launch_bounds__(RED_THREADS) void k_sum_partials(const double* __restrict__ partial,
                                                              int npart, int nvec,
                                                              double* __restrict__ out,
                                                              double* __restrict__ out2) {
  __shared__ double sh[RED_THREADS / 64];
  for (int v = 0; v < nvec; ++v) {
    double s[1] = {0.0};
    for (int i = threadIdx.x; i < npart; i += blockDim.x) s[0] += partial[v * npart + i];
    block_sum<1>(s, sh);
    if (threadIdx.x == 0) {
      if (v == 2) *out2 = s[0];
      else out[v] = s[0];
    }
  }
}

unsigned red_grid(int64_t n) {
  const int64_t b = (n + RED_THREADS - 1) / RED_THREADS;
  return (unsigned)(b < RED_BLOCKS ? (b < 1 ? 1 : b) : RED_BLOCKS);
}

constexpr int NOALLOC_PARTIALS = RED_BLOCKS;
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
double* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};   // per device: the address of `scratch`

}  // namespace

// Scratch of fhs_proj_l2ball_noalloc: the FP64 partial sums of its
// reduction, one per workgroup.  The reduction grid is capped at RED_BLOCKS
// workgroups whatever n is, so this fixed array (16 KiB per device, held
// until the process ends) serves every size.  Host code never passes the
// array's name to a kernel: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_register_device).
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

// Called by fhs_ctx_create with `device` current: looks up this device's copy
// of `scratch` once and keeps its address.
void noalloc_register_device(int device) {
  FHS_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, "fhs_ctx_create: device index too large");
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_ptr[device]) return;
  void* ptr = nullptr;
  FHS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  FHS_CHECK(ptr != nullptr, "fhs_ctx_create: no device address for the proj_l2ball scratch");
  noalloc_ptr[device] = (double*)ptr;
}

void primal_update(fhs_ctx* ctx, float* x, const float* ga, const float* gb, const float* gc,
                   float* xhat, int64_t n, double* d_out2, const float* truth, double* d_truth) {
  const unsigned G = red_grid(n);
  double* part = (double*)ctx->ws.get("primal_partial", 3 * G * sizeof(double));
  Scope sc(ctx, "primal");
  k_primal<<<G, RED_THREADS, 0, ctx->stream>>>(x, ga, gb, gc, xhat, n, truth, part);
  check_launch();
  const bool t = truth && d_truth;
  k_sum_partials<<<1, RED_THREADS, 0, ctx->stream>>>(part, G, t ? 3 : 2, d_out2, d_truth);
  check_launch();
}

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_proj_l2ball(fhs_ctx* ctx, const void* x, double eps, const void* y, int64_t n, void* p) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(x && y && p && n >= 0, "proj_l2ball: bad argument");
    FHS_CHECK(eps >= 0.0, "proj_l2ball: eps must be >= 0");
    if (n == 0) return;
    const unsigned G = red_grid(n);
    double* part = (double*)ctx->ws.get("proj_partial", G * sizeof(double));
    Scope sc(ctx, "proj_l2ball");
    k_sqdiff_partial<<<G, RED_THREADS, 0, ctx->stream>>>((const float2*)x, (const float2*)y, n, part);
    check_launch();
    k_proj_apply<<<G, RED_THREADS, 0, ctx->stream>>>((const float2*)x, (const float2*)y, n, eps, part,
                                                     (int)G, (float2*)p);
    check_launch();
  });
}

int fhs_proj_l2ball_noalloc(fhs_ctx* ctx, const void* x, double eps, const void* y, int64_t n, void* p) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(x && y && p && n >= 0, "proj_l2ball_noalloc: bad argument");
    FHS_CHECK(eps >= 0.0, "proj_l2ball_noalloc: eps must be >= 0");
    if (n == 0) return;
    FHS_CHECK(ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES && noalloc_ptr[ctx->device],
              "proj_l2ball_noalloc: scratch not registered for this device (context not created by fhs_ctx_create)");
    double* part = noalloc_ptr[ctx->device];
    const unsigned G = red_grid(n);
    FHS_CHECK(G <= (unsigned)NOALLOC_PARTIALS, "proj_l2ball_noalloc: reduction grid exceeds the scratch capacity");
    // every partial the second kernel reads is written by the first one on
    // the same stream, so nothing needs clearing between calls
    k_sqdiff_partial<<<G, RED_THREADS, 0, ctx->stream>>>((const float2*)x, (const float2*)y, n, part);
    check_launch();
    k_proj_apply<<<G, RED_THREADS, 0, ctx->stream>>>((const float2*)x, (const float2*)y, n, eps, part,
                                                     (int)G, (float2*)p);
    check_launch();
  });
}

int fhs_update_primal(fhs_ctx* ctx, float* xsol, const float* g_a, const float* g_b,
                      const float* g_c, float* xhat, int64_t n, double* rel_x, double* norm_x) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(xsol && g_a && xhat && n > 0, "update_primal: bad argument");
    double* d = (double*)ctx->ws.get("primal_out", 2 * sizeof(double));
    primal_update(ctx, xsol, g_a, g_b, g_c, xhat, n, d, nullptr, nullptr);
    double h[2];
    FHS_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
    if (rel_x) *rel_x = h[0];
    if (norm_x) *norm_x = h[1];
  });
}

}  // extern "C"
