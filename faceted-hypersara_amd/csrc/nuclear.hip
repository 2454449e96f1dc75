// Weighted nuclear-norm dual step on P x L matrices, batched over facets
// (src/fhs/fhs_update_dual_lowrankness.m:40-45,
//  src/hs/hs_update_dual_lowrankness_serial.m:5-10).
//
//   Z = v0 + w .* xhat_q                    (P_q x L, w = apodization or 1)
//   v0 <- Z - U diag(max(s - beta0 w0, 0)) V'   (economy SVD of Z)
//       = Z V diag(min(1, beta0 w0_i / s_i)) V'     (Gram form, s_i > 0)
//   g0 = sigma * w .* v0
// Three launches:
//   1. Gram Z'Z in FP64 (FP32 inputs widened, so every product is exact).
//      Forming the Gram squares the condition number, and the prox of a
//      near-low-rank cube amplifies Gram errors by ~10^3: FP32 accumulation
//      (2e-4 relative error of the prox, measured) and also BF16 hi/mid/lo
//      splits on the matrix cores, whose 16-product sums are rounded to FP32
//      (Gram error 4e-9, prox error 1.1e-5 > 1e-5), fail; FP64 sums keep the
//      small singular values, those near the threshold, accurate.  One
//      partial 16 x 16 Gram per 1024-pixel chunk; HBM-bound (16 L P bytes
//      per facet).
//   2. Per facet (one wavefront): FP64 sum of the partials, cyclic Jacobi
//      eigen-decomposition warm-started from the previous iteration's
//      eigenvectors (the Gram changes little between iterations, so one or
//      two sweeps suffice), descending sort, M = V diag(f) V'.
//   3. Apply v0 = Z M (and g0) per pixel.
// Channels are padded to 16 (L <= 16).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

namespace fhs {

constexpr int NUC_L = 16;
constexpr int GRAM_THREADS = 256;
constexpr int GRAM_CHUNK = 1024;

struct FacetDev {
  int Q;
  int Ny, Nx;
  const int* oy;        // [Q] facet origin (rows)
  const int* ox;        // [Q]
  const int* py;        // [Q] facet dims
  const int* px;        // [Q]
  const int64_t* voff;  // [Q] offset of facet q's L x P_q block in v0
  const int64_t* woff;  // [Q] offset of facet q's window (P_q), or null
  const float* win;     // windows, or null (w = 1)
};

namespace {

__device__ __forceinline__ float z_val(const FacetDev& F, int q, int L, int P, int py, int c, int p,
                                       const float* __restrict__ v0, const float* __restrict__ xhat,
                                       int64_t N) {
  if (c >= L || p >= P) return 0.f;
  const int iy = F.oy[q] + p % py, ix = F.ox[q] + p / py;
  const float w = F.win ? F.win[F.woff[q] + p] : 1.f;
  const float xv = xhat ? xhat[(int64_t)c * N + iy + (int64_t)F.Ny * ix] : 0.f;
  const float vv = v0 ? v0[F.voff[q] + (int64_t)c * P + p] : 0.f;
  return vv + w * xv;
}

constexpr int GRAM_TP = 256;              // pixels staged per tile (one per thread)
constexpr int GRAM_LD = GRAM_TP + 1;      // LDS row stride (floats)

// Partial Gram of one GRAM_CHUNK-pixel chunk of facet q in FP64.  Per tile,
// thread t forms z[c][t] for the 16 (padded) channels and stages it in LDS
// as FP32 (coalesced along pixels; the next tile is loaded meanwhile).  Thread t then owns the 4 x 4 Gram block
// (t % 16) / 4, (t % 16) % 4 for the pixels t / 16 + 16 n of the tile: per
// pixel it reads 8 values and issues 16 FP64 FMAs (exact products of FP32
// inputs), so LDS traffic is a quarter of one read pair per product.  The 16
// pixel groups' blocks are added in a fixed order at the end.
__global__ __launch_bounds__(GRAM_THREADS) void k_gram(FacetDev F, int L, const float* __restrict__ v0,
                                                       const float* __restrict__ xhat,
                                                       double* __restrict__ partial, int nchunk) {
  __shared__ float zs[NUC_L][GRAM_LD];
  __shared__ double red[GRAM_THREADS / 16][NUC_L * NUC_L];
  const int q = blockIdx.y, chunk = blockIdx.x;
  const int py = F.py[q], P = py * F.px[q];
  const int64_t N = (int64_t)F.Ny * F.Nx;
  const int blk = threadIdx.x % 16, g = threadIdx.x / 16;
  const int i0 = 4 * (blk / 4), j0 = 4 * (blk % 4);
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  const int p0 = chunk * GRAM_CHUNK;
  // the next tile's values are loaded into registers while this one is summed
  float z[NUC_L];
  auto load = [&](int pb) {
    const int p = pb + (int)threadIdx.x;
#pragma unroll
    for (int c = 0; c < NUC_L; ++c) z[c] = z_val(F, q, L, P, py, c, p < P ? p : P, v0, xhat, N);
  };
  load(p0);
  for (int pb = p0; pb < p0 + GRAM_CHUNK; pb += GRAM_TP) {
#pragma unroll
    for (int c = 0; c < NUC_L; ++c) zs[c][threadIdx.x] = z[c];
    __syncthreads();
    if (pb + GRAM_TP < p0 + GRAM_CHUNK) load(pb + GRAM_TP);
#pragma unroll 4
    for (int n = 0; n < GRAM_TP / 16; ++n) {
      const int pl = g + 16 * n;
      double zi[4], zj[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        zi[r] = (double)zs[i0 + r][pl];
        zj[r] = (double)zs[j0 + r][pl];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fma(zi[a], zj[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) red[g][(i0 + a) * NUC_L + j0 + b] = acc[a][b];
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < GRAM_THREADS / 16; ++w) s += red[w][threadIdx.x];
  partial[((int64_t)q * nchunk + chunk) * NUC_L * NUC_L + threadIdx.x] = s;
}

// Mixed-precision partial Gram (opt-in, fhs_lowrank_set_gram_precision): the
// same 1024-pixel chunk partials as k_gram, from BF16 split products on the
// matrix cores.  z (FP32) is split exactly into BF16 parts, z = h + m + l
// (SPLIT = 3) or approximately, z ~ h + l (SPLIT = 2, 16 mantissa bits), and
// Z'Z is accumulated by v_mfma_f32_16x16x32_bf16 from the products h'h,
// h'm, m'h, m'm, h'l, l'h (SPLIT 3) or h'h, h'l, l'h (SPLIT 2) in FP32 over
// a wave's 256 pixels of the chunk; the four waves' 16 x 16 sums are added in
// FP64.
// Each of the 64 lanes holds channel c = lane % 16 at 8 consecutive pixels
// (group lane / 16) of the instruction's 32: the A operand (Z' rows) and the
// B operand (Z columns) are the same register, so the product is Z'Z for
// any assignment of the 32 pixels to the k index.
typedef __bf16 gm_bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short gm_u16x8 __attribute__((ext_vector_type(8)));
typedef float gm_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

template <int SPLIT>
__global__ __launch_bounds__(GRAM_THREADS) void k_gram_mfma(FacetDev F, int L, const float* __restrict__ v0,
                                                            const float* __restrict__ xhat,
                                                            double* __restrict__ partial, int nchunk) {
  constexpr int LD = GRAM_TP + 4;   // row stride (floats): 16-byte rows, rows 4 banks apart
  __shared__ __align__(16) float zs[NUC_L][LD];
  __shared__ double red[GRAM_THREADS / 64][NUC_L * NUC_L];
  const int q = blockIdx.y, chunk = blockIdx.x;
  const int py = F.py[q], P = py * F.px[q];
  const int64_t N = (int64_t)F.Ny * F.Nx;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane & 15, grp = lane >> 4;
  const int p0 = chunk * GRAM_CHUNK;
  // tiles of 256 pixels as in k_gram: thread t loads pixel t's 16 channels
  // (coalesced along pixels), the next tile's loads overlapping this one's
  // products; wave w multiplies pixels [64 w, 64 w + 64) of the tile
  float z[NUC_L];
  auto load = [&](int pb) {
    const int p = pb + (int)threadIdx.x;
#pragma unroll
    for (int k = 0; k < NUC_L; ++k) z[k] = z_val(F, q, L, P, py, k, p < P ? p : P, v0, xhat, N);
  };
  gm_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  load(p0);
  for (int pb = p0; pb < p0 + GRAM_CHUNK; pb += GRAM_TP) {
#pragma unroll
    for (int k = 0; k < NUC_L; ++k) zs[k][threadIdx.x] = z[k];
    __syncthreads();
    if (pb + GRAM_TP < p0 + GRAM_CHUNK) load(pb + GRAM_TP);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const float4* src = reinterpret_cast<const float4*>(&zs[c][64 * wave + 32 * s + 8 * grp]);
      const float4 a0 = src[0], a1 = src[1];
      const float zz[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
      gm_u16x8 h, m, l;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        h[j] = bf16_rne(zz[j]);
        const float r = zz[j] - bf16_val(h[j]);   // exact
        if (SPLIT == 3) {
          m[j] = bf16_rne(r);
          l[j] = bf16_rne(r - bf16_val(m[j]));     // r - m exact; l rounds the last bits
        } else {
          l[j] = bf16_rne(r);
        }
      }
      const gm_bf16x8 H = __builtin_bit_cast(gm_bf16x8, h), Lo = __builtin_bit_cast(gm_bf16x8, l);
      if (SPLIT == 3) {
        const gm_bf16x8 M = __builtin_bit_cast(gm_bf16x8, m);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(M, M, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H, Lo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lo, H, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H, M, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(M, H, acc, 0, 0, 0);
      } else {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H, Lo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Lo, H, acc, 0, 0, 0);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(H, H, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // D[4 grp + r][c] sits in lane (c, grp), register r
#pragma unroll
  for (int r = 0; r < 4; ++r) red[wave][(4 * grp + r) * NUC_L + c] = (double)acc[r];
  __syncthreads();
  const int e = threadIdx.x;
  partial[((int64_t)q * nchunk + chunk) * NUC_L * NUC_L + e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

// FP64 sum of the per-chunk partial Grams: one workgroup of 1024 threads per
// facet; thread t adds the chunks k = t / 256 (mod 4) of entry t % 256 with
// 4 independent accumulators (loads stay in flight), and the four quarters
// are added in a fixed order.
__global__ __launch_bounds__(1024) void k_gram_reduce(const double* __restrict__ partial, int nchunk,
                                                      const int* __restrict__ nchunk_q,
                                                      double* __restrict__ gram) {
  __shared__ double quarter[4][NUC_L * NUC_L];
  const int q = blockIdx.x, e = threadIdx.x % (NUC_L * NUC_L), part = threadIdx.x / (NUC_L * NUC_L);
  const int nc = nchunk_q[q];
  const double* pq = partial + (int64_t)q * nchunk * NUC_L * NUC_L + e;
  double a[4] = {0, 0, 0, 0};
  int k = part;
  for (; k + 12 < nc; k += 16) {
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] += pq[(int64_t)(k + 4 * j) * NUC_L * NUC_L];
  }
  for (; k < nc; k += 4) a[0] += pq[(int64_t)k * NUC_L * NUC_L];
  quarter[part][e] = (a[0] + a[1]) + (a[2] + a[3]);
  __syncthreads();
  if (part == 0)
    gram[(int64_t)q * NUC_L * NUC_L + e] = (quarter[0][e] + quarter[1][e]) + (quarter[2][e] + quarter[3][e]);
}

// One wavefront per facet.  mode 0: dual step (writes M); mode 1: singular
// values only.  Vw: [Q][16][16] FP64 warm-start eigenvectors (in/out).
__global__ __launch_bounds__(64) void k_eig(int L, const double* __restrict__ gram, double* __restrict__ Vw,
                                            const double* __restrict__ thresh, int mode,
                                            float* __restrict__ Mout, double* __restrict__ sv) {
  __shared__ double A[NUC_L][NUC_L + 1];
  __shared__ double V[NUC_L][NUC_L + 1];
  __shared__ double T[NUC_L][NUC_L + 1];
  __shared__ double rc[8], rs[8];
  __shared__ int pp[8], qq[8];
  __shared__ double sval[NUC_L];
  __shared__ int order[NUC_L];
  __shared__ int done;
  const int q = blockIdx.x, t = threadIdx.x;
  // 1. Gram (FP64, reduced by k_gram_reduce)
  for (int e = t; e < NUC_L * NUC_L; e += 64) {
    T[e / NUC_L][e % NUC_L] = gram[(int64_t)q * NUC_L * NUC_L + e];
    V[e / NUC_L][e % NUC_L] = Vw[(int64_t)q * NUC_L * NUC_L + e];
  }
  __syncthreads();
  // symmetrise, then A = V' T V (warm start)
  for (int e = t; e < NUC_L * NUC_L; e += 64) {
    const int i = e / NUC_L, j = e % NUC_L;
    if (i < j) {
      const double m = 0.5 * (T[i][j] + T[j][i]);
      T[i][j] = m;
      T[j][i] = m;
    }
  }
  __syncthreads();
  __shared__ double TV[NUC_L][NUC_L + 1];
  for (int e = t; e < NUC_L * NUC_L; e += 64) {
    const int i = e / NUC_L, j = e % NUC_L;
    double s = 0.0;
    for (int k = 0; k < NUC_L; ++k) s += T[i][k] * V[k][j];
    TV[i][j] = s;
  }
  __syncthreads();
  for (int e = t; e < NUC_L * NUC_L; e += 64) {
    const int i = e / NUC_L, j = e % NUC_L;
    double s = 0.0;
    for (int k = 0; k < NUC_L; ++k) s += V[k][i] * TV[k][j];
    A[i][j] = s;
  }
  __syncthreads();
  for (int e = t; e < NUC_L * NUC_L; e += 64) {
    const int i = e / NUC_L, j = e % NUC_L;
    if (i < j) {
      const double m = 0.5 * (A[i][j] + A[j][i]);
      A[i][j] = m;
      A[j][i] = m;
    }
  }
  __syncthreads();
  // 2. cyclic Jacobi, round-robin pairing (8 disjoint pairs per round)
  __shared__ double offp[64], diap[64];
  for (int sweep = 0; sweep < 20; ++sweep) {
    // converged when the off-diagonal mass is below 1e-24 of the diagonal's
    // (eigenvalues to ~1e-12 relative: far below the FP32 data precision)
    {
      double off = 0.0, dia = 0.0;
      for (int e = t; e < NUC_L * NUC_L; e += 64) {
        const int i = e / NUC_L, j = e % NUC_L;
        const double a = A[i][j] * A[i][j];
        if (i == j)
          dia += a;
        else
          off += a;
      }
      offp[t] = off;
      diap[t] = dia;
    }
    __syncthreads();
    if (t == 0) {
      double off = 0.0, dia = 0.0;
      for (int k = 0; k < 64; ++k) {
        off += offp[k];
        dia += diap[k];
      }
      done = (off <= 1e-24 * dia) || dia == 0.0;
    }
    __syncthreads();
    if (done) break;
    for (int round = 0; round < NUC_L - 1; ++round) {
      if (t < 8) {
        // circle method: position 0 fixed, others rotate
        auto pos = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + round) % (NUC_L - 1); };
        int a = pos(t), b = pos(NUC_L - 1 - t);
        if (a > b) {
          const int tmp = a;
          a = b;
          b = tmp;
        }
        pp[t] = a;
        qq[t] = b;
        const double apq = A[a][b];
        double c = 1.0, s = 0.0;
        if (fabs(apq) > 1e-300) {
          const double tau = (A[b][b] - A[a][a]) / (2.0 * apq);
          const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + tt * tt);
          s = tt * c;
        }
        rc[t] = c;
        rs[t] = s;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int e = t; e < 8 * NUC_L; e += 64) {
        const int k = e / NUC_L, i = e % NUC_L;
        const int a = pp[k], b = qq[k];
        const double c = rc[k], s = rs[k];
        const double x = A[i][a], y = A[i][b];
        A[i][a] = c * x - s * y;
        A[i][b] = s * x + c * y;
        const double vx = V[i][a], vy = V[i][b];
        V[i][a] = c * vx - s * vy;
        V[i][b] = s * vx + c * vy;
      }
      __syncthreads();
      // rows: A <- J' A
      for (int e = t; e < 8 * NUC_L; e += 64) {
        const int k = e / NUC_L, j = e % NUC_L;
        const int a = pp[k], b = qq[k];
        const double c = rc[k], s = rs[k];
        const double x = A[a][j], y = A[b][j];
        A[a][j] = c * x - s * y;
        A[b][j] = s * x + c * y;
      }
      __syncthreads();
    }
  }
  // 3. singular values, descending order
  if (t < NUC_L) sval[t] = sqrt(fmax(A[t][t], 0.0));
  __syncthreads();
  if (t == 0) {
    for (int i = 0; i < NUC_L; ++i) order[i] = i;
    for (int i = 1; i < NUC_L; ++i) {
      const int k = order[i];
      int j = i - 1;
      while (j >= 0 && sval[order[j]] < sval[k]) {
        order[j + 1] = order[j];
        --j;
      }
      order[j + 1] = k;
    }
  }
  __syncthreads();
  if (t < NUC_L) sv[(int64_t)q * NUC_L + t] = t < L ? sval[order[t]] : 0.0;
  for (int e = t; e < NUC_L * NUC_L; e += 64) Vw[(int64_t)q * NUC_L * NUC_L + e] = V[e / NUC_L][e % NUC_L];
  if (mode != 0) return;
  // f for eigenpair order[r] (r-th largest): min(1, thresh_r / s)
  __shared__ double f[NUC_L];
  if (t < NUC_L) {
    const int r = t;
    const int k = order[r];
    const double s = sval[k];
    const double th = r < L ? thresh[(int64_t)q * NUC_L + r] : 0.0;
    f[k] = s > 0.0 ? fmin(1.0, th / s) : 1.0;
  }
  __syncthreads();
  for (int e = t; e < NUC_L * NUC_L; e += 64) {
    const int i = e / NUC_L, j = e % NUC_L;
    double s = 0.0;
    for (int k = 0; k < NUC_L; ++k) s += V[i][k] * f[k] * V[j][k];
    Mout[(int64_t)q * NUC_L * NUC_L + e] = (float)s;
  }
}

// v0 <- Z M; g0 = gscale * w .* v0 (when g0 != null, HS form writes the image)
__global__ __launch_bounds__(256) void k_nuc_apply(FacetDev F, int L, float* __restrict__ v0,
                                                   const float* __restrict__ xhat,
                                                   const float* __restrict__ M, float gscale,
                                                   float* __restrict__ g0) {
  __shared__ float Ms[NUC_L * NUC_L];
  const int q = blockIdx.y;
  for (int e = threadIdx.x; e < NUC_L * NUC_L; e += blockDim.x) Ms[e] = M[(int64_t)q * NUC_L * NUC_L + e];
  __syncthreads();
  const int py = F.py[q], P = py * F.px[q];
  const int64_t N = (int64_t)F.Ny * F.Nx;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float z[NUC_L];
#pragma unroll
  for (int c = 0; c < NUC_L; ++c) z[c] = z_val(F, q, L, P, py, c, p, v0, xhat, N);
  const float w = F.win ? F.win[F.woff[q] + p] : 1.f;
  const int iy = F.oy[q] + p % py, ix = F.ox[q] + p / py;
  for (int c = 0; c < L; ++c) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < NUC_L; ++k) s += z[k] * Ms[k * NUC_L + c];
    v0[F.voff[q] + (int64_t)c * P + p] = s;
    if (g0) g0[(int64_t)c * N + iy + (int64_t)F.Ny * ix] = gscale * w * s;
  }
}

// overlap-add of facet contributions: g0 = gscale * sum_q w_q .* v0_q
// rowf/colf: up to 2 facet rows/cols covering each image row/col (-1 none)
__global__ __launch_bounds__(256) void k_overlap_add(FacetDev F, int L, int Qy,
                                                     const float* __restrict__ v0,
                                                     const int* __restrict__ rowf,
                                                     const int* __restrict__ colf, float gscale,
                                                     float* __restrict__ g0) {
  const int64_t N = (int64_t)F.Ny * F.Nx;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int iy = (int)(i % F.Ny), ix = (int)(i / F.Ny);
    float acc[NUC_L];
#pragma unroll
    for (int c = 0; c < NUC_L; ++c) acc[c] = 0.f;
    for (int a = 0; a < 2; ++a) {
      const int qy = rowf[2 * iy + a];
      if (qy < 0) continue;
      for (int b = 0; b < 2; ++b) {
        const int qx = colf[2 * ix + b];
        if (qx < 0) continue;
        const int q = qx * Qy + qy;
        const int py = F.py[q];
        const int P = py * F.px[q];
        const int p = (iy - F.oy[q]) + py * (ix - F.ox[q]);
        const float w = F.win ? F.win[F.woff[q] + p] : 1.f;
        for (int c = 0; c < L; ++c) acc[c] += w * v0[F.voff[q] + (int64_t)c * P + p];
      }
    }
    for (int c = 0; c < L; ++c) g0[(int64_t)c * N + i] = gscale * acc[c];
  }
}

}  // namespace
}  // namespace fhs

// Facet layout + per-facet persistent state (warm-start eigenvectors).
struct fhs_lowrank {
  int device = 0;
  int Q = 0, Qy = 0, Qx = 0, L = 0, Ny = 0, Nx = 0;
  int nchunk = 0;
  int64_t vsize = 0;                   // total v0 floats
  std::vector<int> h_oy, h_ox, h_py, h_px;
  std::vector<int64_t> h_voff, h_woff;
  fhs::FacetDev F{};
  int* d_geom = nullptr;               // oy, ox, py, px, nchunk_q  (5 Q ints)
  int64_t* d_off = nullptr;            // voff, woff (2 Q)
  float* d_win = nullptr;
  int* d_rowf = nullptr;
  int* d_colf = nullptr;
  double* d_V = nullptr;               // [Q][16][16]
  double* d_thresh = nullptr;          // [Q][16]
  double* d_sv = nullptr;              // [Q][16]
  float* d_M = nullptr;                // [Q][16][16]
  double* d_partial = nullptr;         // [Q][nchunk][256]
  double* d_gram = nullptr;            // [Q][256]
  int gram_mode = 0;                   // 0 FP64, 3 / 2 BF16 split parts on the matrix cores
};

namespace fhs {

static void lowrank_run(fhs_ctx* ctx, fhs_lowrank* R, const float* v0, const float* xhat, int mode) {
  {
    Scope sc(ctx, "nuclear_gram");
    const dim3 grid(R->nchunk, R->Q);
    if (R->gram_mode == 3)
      k_gram_mfma<3><<<grid, GRAM_THREADS, 0, ctx->stream>>>(R->F, R->L, v0, xhat, R->d_partial, R->nchunk);
    else if (R->gram_mode == 2)
      k_gram_mfma<2><<<grid, GRAM_THREADS, 0, ctx->stream>>>(R->F, R->L, v0, xhat, R->d_partial, R->nchunk);
    else
      k_gram<<<grid, GRAM_THREADS, 0, ctx->stream>>>(R->F, R->L, v0, xhat, R->d_partial, R->nchunk);
    check_launch();
  }
  {
    Scope sc(ctx, "nuclear_eig");
    k_gram_reduce<<<R->Q, 1024, 0, ctx->stream>>>(R->d_partial, R->nchunk, R->d_geom + 4 * R->Q, R->d_gram);
    check_launch();
    k_eig<<<R->Q, 64, 0, ctx->stream>>>(R->L, R->d_gram, R->d_V, R->d_thresh, mode, R->d_M, R->d_sv);
    check_launch();
  }
}

// v0 step for all facets; g0 written per pixel (single facet, no overlap) or
// by overlap-add.
void lowrank_dual_step(fhs_ctx* ctx, fhs_lowrank* R, float* v0, const float* xhat, float gscale,
                       float* g0) {
  lowrank_run(ctx, R, v0, xhat, 0);
  int maxP = 0;
  for (int q = 0; q < R->Q; ++q) maxP = std::max(maxP, R->h_py[q] * R->h_px[q]);
  const bool direct = R->Q == 1 && R->h_py[0] == R->Ny && R->h_px[0] == R->Nx;
  {
    Scope sc(ctx, "nuclear_apply");
    k_nuc_apply<<<dim3(cdiv(maxP, 256), R->Q), 256, 0, ctx->stream>>>(R->F, R->L, v0, xhat, R->d_M, gscale,
                                                                      direct ? g0 : nullptr);
    check_launch();
  }
  if (!direct && g0) {
    Scope sc(ctx, "overlap_add");
    const int64_t N = (int64_t)R->Ny * R->Nx;
    k_overlap_add<<<std::min<int64_t>(cdiv(N, 256), 4096), 256, 0, ctx->stream>>>(
        R->F, R->L, R->Qy, v0, R->d_rowf, R->d_colf, gscale, g0);
    check_launch();
  }
}

void lowrank_set_thresh(fhs_ctx* ctx, fhs_lowrank* R, const double* h_thresh) {
  FHS_HIP(hipMemcpyAsync(R->d_thresh, h_thresh, (size_t)R->Q * NUC_L * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
}

int lowrank_facets(const fhs_lowrank* R) { return R->Q; }
int lowrank_channels(const fhs_lowrank* R) { return R->L; }

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_lowrank_create_host(fhs_ctx* ctx, int Ny, int Nx, int L, int Qy, int Qx, const int* oy,
                            const int* ox, const int* py, const int* px, const float* windows,
                            fhs_lowrank** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(L >= 1 && L <= NUC_L, "lowrank_create: 1 <= L <= 16 channels");
    FHS_CHECK(Qy >= 1 && Qx >= 1, "lowrank_create: bad facet grid");
    const int Q = Qy * Qx;
    fhs_lowrank* R = new fhs_lowrank();
    R->device = ctx->device;
    R->Q = Q;
    R->Qy = Qy;
    R->Qx = Qx;
    R->L = L;
    R->Ny = Ny;
    R->Nx = Nx;
    R->h_oy.assign(oy, oy + Q);
    R->h_ox.assign(ox, ox + Q);
    R->h_py.assign(py, py + Q);
    R->h_px.assign(px, px + Q);
    std::vector<int> nchq(Q);
    int64_t voff = 0, woff = 0;
    R->h_voff.resize(Q);
    R->h_woff.resize(Q);
    std::vector<int> rowf(2 * Ny, -1), colf(2 * Nx, -1);
    for (int q = 0; q < Q; ++q) {
      const int qy = q % Qy, qx = q / Qy;   // q = qx*Qy + qy (MATLAB facet order)
      FHS_CHECK(oy[q] >= 0 && ox[q] >= 0 && py[q] >= 1 && px[q] >= 1 && oy[q] + py[q] <= Ny &&
                    ox[q] + px[q] <= Nx,
                "lowrank_create: facet outside the image");
      const int P = py[q] * px[q];
      nchq[q] = (int)cdiv(P, GRAM_CHUNK);
      R->nchunk = std::max(R->nchunk, nchq[q]);
      R->h_voff[q] = voff;
      voff += (int64_t)L * P;
      R->h_woff[q] = woff;
      woff += P;
      for (int y = oy[q]; y < oy[q] + py[q]; ++y) {
        int* r = &rowf[2 * y];
        if (r[0] != qy && r[1] != qy) {
          FHS_CHECK(r[1] < 0, "lowrank_create: a pixel row is covered by more than 2 facets");
          (r[0] < 0 ? r[0] : r[1]) = qy;
        }
      }
      for (int x = ox[q]; x < ox[q] + px[q]; ++x) {
        int* c = &colf[2 * x];
        if (c[0] != qx && c[1] != qx) {
          FHS_CHECK(c[1] < 0, "lowrank_create: a pixel column is covered by more than 2 facets");
          (c[0] < 0 ? c[0] : c[1]) = qx;
        }
      }
    }
    R->vsize = voff;
    FHS_HIP(hipSetDevice(ctx->device));
    std::vector<int> geom;
    geom.insert(geom.end(), oy, oy + Q);
    geom.insert(geom.end(), ox, ox + Q);
    geom.insert(geom.end(), py, py + Q);
    geom.insert(geom.end(), px, px + Q);
    geom.insert(geom.end(), nchq.begin(), nchq.end());
    FHS_HIP(hipMalloc(&R->d_geom, geom.size() * sizeof(int)));
    FHS_HIP(hipMemcpy(R->d_geom, geom.data(), geom.size() * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int64_t> offs(R->h_voff);
    offs.insert(offs.end(), R->h_woff.begin(), R->h_woff.end());
    FHS_HIP(hipMalloc(&R->d_off, offs.size() * sizeof(int64_t)));
    FHS_HIP(hipMemcpy(R->d_off, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if (windows) {
      FHS_HIP(hipMalloc(&R->d_win, woff * sizeof(float)));
      FHS_HIP(hipMemcpy(R->d_win, windows, woff * sizeof(float), hipMemcpyHostToDevice));
    }
    FHS_HIP(hipMalloc(&R->d_rowf, rowf.size() * sizeof(int)));
    FHS_HIP(hipMemcpy(R->d_rowf, rowf.data(), rowf.size() * sizeof(int), hipMemcpyHostToDevice));
    FHS_HIP(hipMalloc(&R->d_colf, colf.size() * sizeof(int)));
    FHS_HIP(hipMemcpy(R->d_colf, colf.data(), colf.size() * sizeof(int), hipMemcpyHostToDevice));
    std::vector<double> eye((size_t)Q * NUC_L * NUC_L, 0.0);
    for (int q = 0; q < Q; ++q)
      for (int i = 0; i < NUC_L; ++i) eye[(size_t)q * NUC_L * NUC_L + i * NUC_L + i] = 1.0;
    FHS_HIP(hipMalloc(&R->d_V, eye.size() * sizeof(double)));
    FHS_HIP(hipMemcpy(R->d_V, eye.data(), eye.size() * sizeof(double), hipMemcpyHostToDevice));
    FHS_HIP(hipMalloc(&R->d_thresh, (size_t)Q * NUC_L * sizeof(double)));
    FHS_HIP(hipMemset(R->d_thresh, 0, (size_t)Q * NUC_L * sizeof(double)));
    FHS_HIP(hipMalloc(&R->d_sv, (size_t)Q * NUC_L * sizeof(double)));
    FHS_HIP(hipMalloc(&R->d_M, (size_t)Q * NUC_L * NUC_L * sizeof(float)));
    FHS_HIP(hipMalloc(&R->d_partial, (size_t)Q * R->nchunk * NUC_L * NUC_L * sizeof(double)));
    FHS_HIP(hipMalloc(&R->d_gram, (size_t)Q * NUC_L * NUC_L * sizeof(double)));
    FacetDev& F = R->F;
    F.Q = Q;
    F.Ny = Ny;
    F.Nx = Nx;
    F.oy = R->d_geom;
    F.ox = R->d_geom + Q;
    F.py = R->d_geom + 2 * Q;
    F.px = R->d_geom + 3 * Q;
    F.voff = R->d_off;
    F.woff = R->d_off + Q;
    F.win = R->d_win;
    *out = R;
  });
}

int fhs_lowrank_destroy(fhs_lowrank* R) {
  if (!R) return FHS_OK;
  (void)hipSetDevice(R->device);
  void* ps[] = {R->d_geom, R->d_off, R->d_win, R->d_rowf, R->d_colf, R->d_V,
                R->d_thresh, R->d_sv, R->d_M, R->d_partial, R->d_gram};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete R;
  return FHS_OK;
}

int fhs_lowrank_set_gram_precision(fhs_lowrank* R, int mode) {
  if (!R || (mode != 0 && mode != 2 && mode != 3)) return FHS_ERR_ARG;
  R->gram_mode = mode;
  return FHS_OK;
}

int fhs_lowrank_info(const fhs_lowrank* R, int64_t* v0_size) {
  if (!R) return FHS_ERR_ARG;
  if (v0_size) *v0_size = R->vsize;
  return FHS_OK;
}

int fhs_update_dual_lowrankness(fhs_ctx* ctx, fhs_lowrank* R, float* v0, const float* xhat,
                                const double* weights0, double beta0, float sigma0, float* g0) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(R && v0 && xhat && weights0, "update_dual_lowrankness: null argument");
    std::vector<double> th((size_t)R->Q * NUC_L, 0.0);
    for (int q = 0; q < R->Q; ++q)
      for (int i = 0; i < R->L; ++i) th[(size_t)q * NUC_L + i] = beta0 * weights0[(size_t)q * R->L + i];
    lowrank_set_thresh(ctx, R, th.data());
    lowrank_dual_step(ctx, R, v0, xhat, sigma0, g0);
    FHS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int fhs_lowrank_singular_values(fhs_ctx* ctx, fhs_lowrank* R, const float* x, double* sv) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(R && x && sv, "lowrank_singular_values: null argument");
    lowrank_run(ctx, R, nullptr, x, 1);
    std::vector<double> h((size_t)R->Q * NUC_L);
    FHS_HIP(hipMemcpyAsync(h.data(), R->d_sv, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < R->Q; ++q)
      for (int i = 0; i < R->L; ++i) sv[(size_t)q * R->L + i] = h[(size_t)q * NUC_L + i];
  });
}

int fhs_nuclear_norm(fhs_ctx* ctx, const float* x, int Ny, int Nx, int L, double* out) {
  if (!ctx) return FHS_ERR_ARG;
  fhs_lowrank* R = nullptr;
  const int oy = 0, ox = 0;
  int st = fhs_lowrank_create_host(ctx, Ny, Nx, L, 1, 1, &oy, &ox, &Ny, &Nx, nullptr, &R);
  if (st != FHS_OK) return st;
  std::vector<double> sv(L);
  st = fhs_lowrank_singular_values(ctx, R, x, sv.data());
  fhs_lowrank_destroy(R);
  if (st == FHS_OK) {
    double s = 0.0;
    for (double v : sv) s += std::fabs(v);
    *out = s;
  }
  return st;
}

}  // extern "C"
