// Kernels of the learned warm start (a small U-Net): BF16 GEMM on the
// matrix cores with FP32 accumulation, im2col / col2im for 3x3 convolutions,
// pooling, upsampling, ReLU, bias reductions, MSE and Adam.
//
// Activations are FP32 in [C][N][H][W] order (channel planes over the whole
// batch), so a 3x3 convolution is one GEMM:
//   forward      out[co][q]  = sum_k W[co][k] col[k][q] (+ bias, ReLU)
//   grad input   dcol[k][q]  = sum_co W[co][k] dout[co][q]  -> col2im gather
//   grad weight  dW[co][k]   = sum_q dout[co][q] col[k][q]
// with k = ci*9 + ky*3 + kx and q = (n, y, x).  Operands are rounded to BF16
// when staged in LDS; v_mfma_f32_16x16x16_bf16 accumulates in FP32.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace fhs {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (unsigned short)(u >> 16);
}

constexpr int GT = 64;     // C tile edge
constexpr int GK = 16;     // k step (one MFMA depth)
constexpr int GTHREADS = 256;

// C (M x N, row-major ldc) = alpha op(A) op(B) + beta C (+ bias[m]) (ReLU)
// op(A)[m][k] = ta ? A[k*lda + m] : A[m*lda + k]; op(B)[k][n] = tb ? B[n*ldb + k] : B[k*ldb + n]
// Split-K: workgroup z sums k in [z*kchunk, (z+1)*kchunk) and, when part is
// given, stores alpha * (its partial product) to part[z][M][N]; k_gemm_reduce
// then adds the partials in z order (deterministic), beta C, bias and ReLU.
__global__ __launch_bounds__(GTHREADS) void k_gemm_bf16(int M, int N, int K, const float* __restrict__ A,
                                                        int lda, int ta, const float* __restrict__ B, int ldb,
                                                        int tb, float* __restrict__ Cm, int ldc, float alpha,
                                                        float beta, const float* __restrict__ bias, int relu,
                                                        int kchunk, float* __restrict__ part) {
  __shared__ unsigned short As[GT][GK + 4];
  __shared__ unsigned short Bs[GK][GT + 4];
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;   // wave sub-tile origin
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += GK) {
    for (int t = threadIdx.x; t < GT * GK; t += GTHREADS) {
      // A tile: GT rows x GK cols
      int r, c;
      if (ta) { r = t % GT; c = t / GT; } else { c = t % GK; r = t / GK; }
      const int m = m0 + r, k = k0 + c;
      float v = 0.f;
      if (m < M && k < ke) v = ta ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
      As[r][c] = to_bf16(v);
    }
    for (int t = threadIdx.x; t < GT * GK; t += GTHREADS) {
      int r, c;   // r = k, c = n
      if (tb) { r = t % GK; c = t / GK; } else { c = t % GT; r = t / GT; }
      const int k = k0 + r, n = n0 + c;
      float v = 0.f;
      if (n < N && k < ke) v = tb ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n];
      Bs[r][c] = to_bf16(v);
    }
    __syncthreads();
    // lane l: A fragment rows l%16, k 4*(l/16)..+3 ; B fragment k 4*(l/16)..+3, col l%16
    const int fr = lane & 15, fk = 4 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      v4s a;
      a[0] = (short)As[wm + 16 * i + fr][fk + 0];
      a[1] = (short)As[wm + 16 * i + fr][fk + 1];
      a[2] = (short)As[wm + 16 * i + fr][fk + 2];
      a[3] = (short)As[wm + 16 * i + fr][fk + 3];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        v4s b;
        b[0] = (short)Bs[fk + 0][wn + 16 * j + fr];
        b[1] = (short)Bs[fk + 1][wn + 16 * j + fr];
        b[2] = (short)Bs[fk + 2][wn + 16 * j + fr];
        b[3] = (short)Bs[fk + 3][wn + 16 * j + fr];
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D[4*(lane/16) + r][lane%16]
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m < M && n < N) {
          float v = alpha * acc[i][j][r];
          if (part) {
            part[((size_t)blockIdx.z * M + m) * N + n] = v;
            continue;
          }
          if (beta != 0.f) v += beta * Cm[(size_t)m * ldc + n];
          if (bias) v += bias[m];
          if (relu) v = fmaxf(v, 0.f);
          Cm[(size_t)m * ldc + n] = v;
        }
      }
}

__global__ void k_gemm_reduce(int M, int N, int splits, const float* __restrict__ part, float* __restrict__ Cm,
                              int ldc, float beta, const float* __restrict__ bias, int relu) {
  const int64_t total = (int64_t)M * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / N), n = (int)(i % N);
    float v = 0.f;
    for (int z = 0; z < splits; ++z) v += part[(size_t)z * total + i];
    if (beta != 0.f) v += beta * Cm[(size_t)m * ldc + n];
    if (bias) v += bias[m];
    if (relu) v = fmaxf(v, 0.f);
    Cm[(size_t)m * ldc + n] = v;
  }
}

// col[ci*9 + ky*3 + kx][n, y, x] = in[ci][n][y + ky - 1][x + kx - 1]
__global__ void k_im2col3(const float* __restrict__ in, int C, int NB, int H, int W, float* __restrict__ col) {
  const int64_t Q = (int64_t)NB * H * W;
  const int64_t total = (int64_t)C * 9 * Q;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = i % Q;
    const int64_t kk = i / Q;
    const int ci = (int)(kk / 9), r = (int)(kk % 9), ky = r / 3, kx = r % 3;
    const int x = (int)(q % W), y = (int)((q / W) % H), n = (int)(q / ((int64_t)W * H));
    const int yy = y + ky - 1, xx = x + kx - 1;
    float v = 0.f;
    if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = in[(((int64_t)ci * NB + n) * H + yy) * W + xx];
    col[i] = v;
  }
}

// din[ci][n][y][x] = sum_{ky,kx} dcol[ci*9 + ky*3 + kx][n, y - ky + 1, x - kx + 1]
__global__ void k_col2im3(const float* __restrict__ dcol, int C, int NB, int H, int W, float* __restrict__ din) {
  const int64_t Q = (int64_t)NB * H * W;
  const int64_t total = (int64_t)C * Q;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = i % Q;
    const int ci = (int)(i / Q);
    const int x = (int)(q % W), y = (int)((q / W) % H), n = (int)(q / ((int64_t)W * H));
    float s = 0.f;
    for (int ky = 0; ky < 3; ++ky) {
      const int yy = y - ky + 1;
      if (yy < 0 || yy >= H) continue;
      for (int kx = 0; kx < 3; ++kx) {
        const int xx = x - kx + 1;
        if (xx < 0 || xx >= W) continue;
        s += dcol[((int64_t)ci * 9 + ky * 3 + kx) * Q + ((int64_t)n * H + yy) * W + xx];
      }
    }
    din[i] = s;
  }
}

__global__ void k_relu_bwd(const float* __restrict__ y, float* __restrict__ g, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (y[i] <= 0.f) g[i] = 0.f;
}

// 2x2 max pooling over [C*NB] planes of H x W (H, W even)
__global__ void k_maxpool2(const float* __restrict__ in, int64_t planes, int H, int W, float* __restrict__ out) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t total = planes * Ho * Wo;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % Wo), y = (int)((i / Wo) % Ho);
    const int64_t p = i / ((int64_t)Wo * Ho);
    const float* b = in + (p * H + 2 * y) * W + 2 * x;
    out[i] = fmaxf(fmaxf(b[0], b[1]), fmaxf(b[W], b[W + 1]));
  }
}

// gradient goes to the first maximal element of each window
__global__ void k_maxpool2_bwd(const float* __restrict__ in, const float* __restrict__ gout, int64_t planes, int H,
                               int W, float* __restrict__ gin) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t total = planes * Ho * Wo;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % Wo), y = (int)((i / Wo) % Ho);
    const int64_t p = i / ((int64_t)Wo * Ho);
    const int64_t o = (p * H + 2 * y) * W + 2 * x;
    const float v[4] = {in[o], in[o + 1], in[o + W], in[o + W + 1]};
    int a = 0;
    for (int t = 1; t < 4; ++t)
      if (v[t] > v[a]) a = t;
    const float g = gout[i];
    gin[o] = a == 0 ? g : 0.f;
    gin[o + 1] = a == 1 ? g : 0.f;
    gin[o + W] = a == 2 ? g : 0.f;
    gin[o + W + 1] = a == 3 ? g : 0.f;
  }
}

// nearest 2x upsampling of [planes] H x W -> 2H x 2W, and its adjoint
__global__ void k_upsample2(const float* __restrict__ in, int64_t planes, int H, int W, float* __restrict__ out) {
  const int64_t total = planes * 4 * H * W;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % (2 * W)), y = (int)((i / (2 * W)) % (2 * H));
    const int64_t p = i / ((int64_t)4 * H * W);
    out[i] = in[(p * H + y / 2) * W + x / 2];
  }
}

__global__ void k_upsample2_bwd(const float* __restrict__ gout, int64_t planes, int H, int W,
                                float* __restrict__ gin) {
  const int64_t total = planes * H * W;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % W), y = (int)((i / W) % H);
    const int64_t p = i / ((int64_t)H * W);
    const float* g = gout + (p * 2 * H + 2 * y) * (2 * W) + 2 * x;
    gin[i] = g[0] + g[1] + g[2 * W] + g[2 * W + 1];
  }
}

// db[c] = sum_q g[c][q]  (one workgroup per channel, FP64 accumulation)
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ g, int64_t Q, float* __restrict__ db) {
  __shared__ double sh[4];
  double s[1] = {0.0};
  const float* gc = g + (int64_t)blockIdx.x * Q;
  for (int64_t i = threadIdx.x; i < Q; i += blockDim.x) s[0] += gc[i];
  block_sum<1>(s, sh);
  if (threadIdx.x == 0) db[blockIdx.x] = (float)s[0];
}

// loss partials: sum (p - t)^2 ; grad = scale (p - t)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ p, const float* __restrict__ t, int64_t n,
                                             float scale, float* __restrict__ grad, double* __restrict__ part) {
  __shared__ double sh[4];
  double s[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float d = p[i] - t[i];
    s[0] += (double)d * d;
    if (grad) grad[i] = scale * d;
  }
  block_sum<1>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

unsigned egrid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

}  // namespace
}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_nn_gemm(fhs_ctx* ctx, int M, int N, int K, const float* A, int lda, int trans_a, const float* B,
                int ldb, int trans_b, float* C, int ldc, float alpha, float beta, const float* bias, int relu) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(M > 0 && N > 0 && K > 0 && A && B && C, "nn_gemm: bad argument");
    Scope sc(ctx, "nn_gemm");
    // long K over few output tiles (the weight gradient): split K across
    // workgroups so that the chip fills, partials reduced in a fixed order
    const int64_t tiles = (int64_t)cdiv(N, GT) * cdiv(M, GT);
    int splits = 1;
    if (tiles < 1024 && K >= 4 * 1024) {
      splits = (int)std::min<int64_t>(cdiv(K, 1024), std::max<int64_t>(1, 2048 / tiles));
      splits = std::min(splits, 1024);
    }
    const int kchunk = (int)((cdiv(K, splits) + GK - 1) / GK * GK);
    splits = (int)cdiv(K, kchunk);
    float* part = nullptr;
    if (splits > 1)
      part = (float*)ctx->ws.get("nn_gemm_split", (size_t)splits * M * N * sizeof(float));
    k_gemm_bf16<<<dim3(cdiv(N, GT), cdiv(M, GT), splits), GTHREADS, 0, ctx->stream>>>(
        M, N, K, A, lda, trans_a, B, ldb, trans_b, C, ldc, alpha, beta, bias, relu, kchunk, part);
    check_launch();
    if (splits > 1) {
      k_gemm_reduce<<<egrid((int64_t)M * N), 256, 0, ctx->stream>>>(M, N, splits, part, C, ldc, beta, bias, relu);
      check_launch();
    }
  });
}

int fhs_nn_im2col3(fhs_ctx* ctx, const float* in, int C, int NB, int H, int W, float* col) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_im2col3<<<egrid((int64_t)C * 9 * NB * H * W), 256, 0, ctx->stream>>>(in, C, NB, H, W, col);
    check_launch();
  });
}

int fhs_nn_col2im3(fhs_ctx* ctx, const float* dcol, int C, int NB, int H, int W, float* din) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_col2im3<<<egrid((int64_t)C * NB * H * W), 256, 0, ctx->stream>>>(dcol, C, NB, H, W, din);
    check_launch();
  });
}

int fhs_nn_relu_bwd(fhs_ctx* ctx, const float* y, float* g, int64_t n) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_relu_bwd<<<egrid(n), 256, 0, ctx->stream>>>(y, g, n);
    check_launch();
  });
}

int fhs_nn_maxpool2(fhs_ctx* ctx, const float* in, int64_t planes, int H, int W, float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(H % 2 == 0 && W % 2 == 0, "nn_maxpool2: even sizes");
    k_maxpool2<<<egrid(planes * H * W / 4), 256, 0, ctx->stream>>>(in, planes, H, W, out);
    check_launch();
  });
}

int fhs_nn_maxpool2_bwd(fhs_ctx* ctx, const float* in, const float* gout, int64_t planes, int H, int W,
                        float* gin) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_maxpool2_bwd<<<egrid(planes * H * W / 4), 256, 0, ctx->stream>>>(in, gout, planes, H, W, gin);
    check_launch();
  });
}

int fhs_nn_upsample2(fhs_ctx* ctx, const float* in, int64_t planes, int H, int W, float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_upsample2<<<egrid(planes * H * W * 4), 256, 0, ctx->stream>>>(in, planes, H, W, out);
    check_launch();
  });
}

int fhs_nn_upsample2_bwd(fhs_ctx* ctx, const float* gout, int64_t planes, int H, int W, float* gin) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_upsample2_bwd<<<egrid(planes * H * W), 256, 0, ctx->stream>>>(gout, planes, H, W, gin);
    check_launch();
  });
}

int fhs_nn_bias_grad(fhs_ctx* ctx, const float* g, int C, int64_t Q, float* db) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    k_bias_grad<<<C, 256, 0, ctx->stream>>>(g, Q, db);
    check_launch();
  });
}

int fhs_nn_mse(fhs_ctx* ctx, const float* pred, const float* target, int64_t n, float grad_scale, float* grad,
               double* loss) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    const unsigned G = egrid(n) < 1024 ? egrid(n) : 1024;
    double* part = (double*)ctx->ws.get("nn_mse_part", G * sizeof(double));
    k_mse<<<G, 256, 0, ctx->stream>>>(pred, target, n, grad_scale, grad, part);
    check_launch();
    std::vector<double> h(G);
    FHS_HIP(hipMemcpyAsync(h.data(), part, G * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (double v : h) s += v;
    if (loss) *loss = s;
  });
}

int fhs_nn_adam(fhs_ctx* ctx, float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                float beta2, float eps, int step) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(step >= 1, "nn_adam: step >= 1");
    const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    k_adam<<<egrid(n), 256, 0, ctx->stream>>>(w, g, m, v, n, lr, beta1, beta2, eps, c1, c2);
    check_launch();
  });
}

}  // extern "C"
