// Oversampled scaled 2-D FFT of the NUFFT: A = fft2(pad(sn .* x)) and its
// adjoint, batched over channels.
//
// The grid is Kty x Ktx (column-major, y contiguous); the image Ny x Nx sits
// in its top-left corner.  Each transform is two passes through HBM:
//   forward : (1) y-pass: one workgroup per image column, scale by sn, zero
//                 pad to Kty and FFT in LDS -> T (Kty x Nx, only Nx columns
//                 exist, the rest are zero);
//             (2) x-pass: one workgroup per B consecutive grid rows, FFT of
//                 length Ktx over the Nx non-zero columns of T -> Fx.
//   adjoint : (1') x-pass inverse keeping the first Nx outputs -> T;
//             (2') y-pass inverse, keep Ny outputs, * sn, real part, * scale.
// The x-pass reads and writes B consecutive complex values per grid column
// (B = 16 -> 128-byte segments, B = 8 -> 64, B = 4 -> 32) and keeps B transforms interleaved in LDS
// (s[i*B + b]) so that consecutive lanes touch consecutive LDS words.
// FFTs are radix-4 Stockham (autosort) with one radix-2 stage when log2 n is
// odd; twiddles come from an FP64-computed table exp(-2 pi i k/n).
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"
#include "nufft.hpp"

namespace fhs {
namespace {

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

template <int N, int B, int NT, bool INV>
__device__ __forceinline__ void fft_lds(float2* s, const float2* tw) {
  const int tid = threadIdx.x;
  constexpr int LOG = ilog2(N);
  int Ns = 1;
  if constexpr (LOG & 1) {
    constexpr int NB = (B * N / 2 + NT - 1) / NT;
    float2 v[NB][2];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int u = tid + q * NT;
      if (u < B * N / 2) {
        const int b = u % B, j = u / B;
        v[q][0] = s[j * B + b];
        v[q][1] = s[(j + N / 2) * B + b];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int u = tid + q * NT;
      if (u < B * N / 2) {
        const int b = u % B, j = u / B;
        s[(2 * j) * B + b] = cadd(v[q][0], v[q][1]);
        s[(2 * j + 1) * B + b] = csub(v[q][0], v[q][1]);
      }
    }
    __syncthreads();
    Ns = 2;
  }
#pragma unroll 1
  for (; Ns < N; Ns *= 4) {
    constexpr int NB = (B * N / 4 + NT - 1) / NT;
    const int tstride = N / (4 * Ns);
    float2 v[NB][4];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int u = tid + q * NT;
      if (u < B * N / 4) {
        const int b = u % B, j = u / B;
        const int k = j & (Ns - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[q][r] = s[(j + r * (N / 4)) * B + b];
        if (k) {
#pragma unroll
          for (int r = 1; r < 4; ++r) {
            float2 w = tw[r * k * tstride];
            if (INV) w.y = -w.y;
            v[q][r] = cmul(v[q][r], w);
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int u = tid + q * NT;
      if (u < B * N / 4) {
        const int b = u % B, j = u / B;
        const int k = j & (Ns - 1);
        const float2 a0 = cadd(v[q][0], v[q][2]);
        const float2 a1 = csub(v[q][0], v[q][2]);
        const float2 a2 = cadd(v[q][1], v[q][3]);
        const float2 d = csub(v[q][1], v[q][3]);
        // forward: (v1 - v3) * (-i); inverse: * (+i)
        const float2 a3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
        const int base = (j / Ns) * Ns * 4 + k;
        s[(base)*B + b] = cadd(a0, a2);
        s[(base + Ns) * B + b] = cadd(a1, a3);
        s[(base + 2 * Ns) * B + b] = csub(a0, a2);
        s[(base + 3 * Ns) * B + b] = csub(a1, a3);
      }
    }
    __syncthreads();
  }
}

// The radix-4 stages read twiddles tw[r k tstride] < 3N/4.  They are staged
// in LDS once per workgroup (read behind the data loads' barrier) instead of
// a cache round trip in every stage.  The y-passes (one column, N x 8 bytes
// of LDS) always stage them; the x-passes up to N = 2048 only, since at
// N = 4096 the table would halve their occupancy.
template <int N>
constexpr int TW_LDS = (N < 4 || N > 2048) ? 1 : 3 * N / 4;
template <int N>
constexpr int TW_LDS_Y = N < 4 ? 1 : 3 * N / 4;
template <int N, int NT, int TWN>
__device__ __forceinline__ const float2* stage_twiddles(const float2* __restrict__ tw, float2* tws) {
  if constexpr (TWN == 1) return tw;
  for (int i = threadIdx.x; i < TWN; i += NT) tws[i] = tw[i];
  return tws;
}

template <int N>
constexpr int nt_contig() { return N / 4 < 64 ? 64 : (N / 4 > 1024 ? 1024 : N / 4); }
// grid rows per x-pass workgroup: 16 up to N = 512, 8 at N = 1024 (two
// workgroups per CU), 4 from N = 2048 (64 KB of rows plus the twiddles, two
// workgroups per CU; 128 KB and one per CU at N = 4096, where 2 rows, 16-byte
// segments, measured slower).  Segments below 128 bytes are merged in one
// XCD's L2 by xpass_row0.
template <int N>
constexpr int b_strided() { return N >= 2048 ? 4 : (N >= 1024 ? 8 : 16); }
template <int N>
constexpr int nt_strided() {
  return b_strided<N>() * N / 4 > 1024 ? 1024 : (b_strided<N>() * N / 4 < 64 ? 64 : b_strided<N>() * N / 4);
}

// x-pass workgroup -> first grid row.  Workgroups are dispatched to the 8
// XCDs round-robin (linear id mod 8; gridDim.x is a multiple of 8 here, so
// blockIdx.x mod 8).  When B rows are fewer than a 128-byte line (G = 16 / B
// workgroups share each line of every grid column), workgroups b, b + 8, ...,
// b + 8 (G - 1) take G adjacent row groups, so the line is fetched into one
// XCD's L2 instead of G.
template <int N>
__device__ __forceinline__ int xpass_row0() {
  constexpr int B = b_strided<N>();
  constexpr int G = 16 / B;
  const int b = blockIdx.x;
  if constexpr (G == 1) {
    return b * B;
  } else {
    if (gridDim.x % (8 * G)) return b * B;
    return ((b / (8 * G)) * (8 * G) + (b % 8) * G + (b / 8) % G) * B;
  }
}

// (1) forward y-pass
template <int N>
__global__ __launch_bounds__(nt_contig<N>()) void k_fwd_y(const float* __restrict__ x,
                                                         const float* __restrict__ sy,
                                                         const float* __restrict__ sx, int Ny, int Nx,
                                                         const float2* __restrict__ tw,
                                                         float2* __restrict__ T) {
  constexpr int NT = nt_contig<N>();
  __shared__ float2 s[N];
  __shared__ float2 tws[TW_LDS_Y<N>];
  const int ix = blockIdx.x, l = blockIdx.y;
  const float2* twp = stage_twiddles<N, NT, TW_LDS_Y<N>>(tw, tws);
  const float* xc = x + ((size_t)l * Nx + ix) * Ny;
  const float scx = sx[ix];
  constexpr int PER = (N + NT - 1) / NT;
  float v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = threadIdx.x + j * NT;
    v[j] = i < Ny ? xc[i] * sy[i] * scx : 0.f;
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < N) s[i] = make_float2(v[j], 0.f);
  }
  __syncthreads();
  fft_lds<N, 1, NT, false>(s, twp);
  float2* out = T + ((size_t)l * Nx + ix) * N;
  for (int i = threadIdx.x; i < N; i += NT) out[i] = s[i];
}

// (2) forward x-pass: N = Ktx
template <int N>
__global__ __launch_bounds__(nt_strided<N>()) void k_fwd_x(const float2* __restrict__ T, int Kty,
                                                          int Nx, const float2* __restrict__ tw,
                                                          float2* __restrict__ Fx) {
  constexpr int B = b_strided<N>();
  constexpr int NT = nt_strided<N>();
  __shared__ float2 s[N * B];
  __shared__ float2 tws[TW_LDS<N>];
  const int ky0 = xpass_row0<N>(), l = blockIdx.y;
  const float2* twp = stage_twiddles<N, NT, TW_LDS<N>>(tw, tws);
  const float2* Tl = T + (size_t)l * Nx * Kty;
  constexpr int PER = N * B / NT;   // every load issued before the first LDS store
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    const int b = t % B, i = t / B;
    v[j] = i < Nx ? Tl[(size_t)i * Kty + ky0 + b] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) s[threadIdx.x + j * NT] = v[j];
  __syncthreads();
  fft_lds<N, B, NT, false>(s, twp);
  float2* F = Fx + (size_t)l * N * Kty;
  for (int t = threadIdx.x; t < N * B; t += NT) {
    const int b = t % B, k = t / B;
    F[(size_t)k * Kty + ky0 + b] = s[t];
  }
}

// (1') adjoint x-pass: N = Ktx, keep the first Nx outputs
template <int N>
__global__ __launch_bounds__(nt_strided<N>()) void k_adj_x(const float2* __restrict__ g, int Kty,
                                                          int Nx, const float2* __restrict__ tw,
                                                          float2* __restrict__ T) {
  constexpr int B = b_strided<N>();
  constexpr int NT = nt_strided<N>();
  __shared__ float2 s[N * B];
  __shared__ float2 tws[TW_LDS<N>];
  const int ky0 = xpass_row0<N>(), l = blockIdx.y;
  const float2* twp = stage_twiddles<N, NT, TW_LDS<N>>(tw, tws);
  const float2* gl = g + (size_t)l * N * Kty;
  constexpr int PER = N * B / NT;   // every load issued before the first LDS store
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    v[j] = gl[(size_t)(t / B) * Kty + ky0 + t % B];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) s[threadIdx.x + j * NT] = v[j];
  __syncthreads();
  fft_lds<N, B, NT, true>(s, twp);
  float2* Tl = T + (size_t)l * Nx * Kty;
  for (int t = threadIdx.x; t < Nx * B; t += NT) {
    const int b = t % B, i = t / B;
    Tl[(size_t)i * Kty + ky0 + b] = s[t];
  }
}

// Split x-passes for N >= X_SPLIT_MIN when the image holds at most half the
// grid columns (Nx <= N / 2, true of every 2x-oversampled grid).  With
// H = N / 2 and W = exp(-2 pi i / N):
//   forward : only x[0 .. H) is non-zero, so
//             X[2m] = FFT_H(x)[m],  X[2m + 1] = FFT_H(W^n x[n])[m];
//   adjoint : only y[0 .. H) is kept, so
//             y[n] = IFFT_H(g[2m])[n] + conj(W^n) IFFT_H(g[2m + 1])[n].
// The two H-point transforms run one after the other in the same H x B
// rows of LDS (64 KB at N = 4096 instead of 128 KB: two workgroups per CU),
// the H-point twiddles being every other entry of the N-point table.
constexpr int X_SPLIT_MIN = 4096;

template <int N>
__global__ __launch_bounds__(nt_strided<N>()) void k_fwd_x_split(const float2* __restrict__ T, int Kty,
                                                                int Nx, const float2* __restrict__ tw,
                                                                float2* __restrict__ Fx) {
  constexpr int H = N / 2;
  constexpr int B = b_strided<N>();
  constexpr int NT = nt_strided<N>();
  static_assert(H * B % NT == 0, "split x-pass: whole rows per thread");
  __shared__ float2 s[H * B];
  __shared__ float2 tws[3 * H / 4];
  const int ky0 = xpass_row0<N>(), l = blockIdx.y;
  for (int i = threadIdx.x; i < 3 * H / 4; i += NT) tws[i] = tw[2 * i];
  const float2* Tl = T + (size_t)l * Nx * Kty;
  constexpr int PER = H * B / NT;
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    const int b = t % B, i = t / B;
    v[j] = i < Nx ? Tl[(size_t)i * Kty + ky0 + b] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) s[threadIdx.x + j * NT] = v[j];
  __syncthreads();
  fft_lds<H, B, NT, false>(s, tws);
  float2* F = Fx + (size_t)l * N * Kty;
  // each thread reads back, then overwrites, only its own LDS slots
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    const int b = t % B, k = t / B;
    F[(size_t)(2 * k) * Kty + ky0 + b] = s[t];
    s[t] = cmul(v[j], tw[k]);
  }
  __syncthreads();
  fft_lds<H, B, NT, false>(s, tws);
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    const int b = t % B, k = t / B;
    F[(size_t)(2 * k + 1) * Kty + ky0 + b] = s[t];
  }
}

template <int N>
__global__ __launch_bounds__(nt_strided<N>()) void k_adj_x_split(const float2* __restrict__ g, int Kty,
                                                                int Nx, const float2* __restrict__ tw,
                                                                float2* __restrict__ T) {
  constexpr int H = N / 2;
  constexpr int B = b_strided<N>();
  constexpr int NT = nt_strided<N>();
  static_assert(H * B % NT == 0, "split x-pass: whole rows per thread");
  __shared__ float2 s[H * B];
  __shared__ float2 tws[3 * H / 4];
  const int ky0 = xpass_row0<N>(), l = blockIdx.y;
  for (int i = threadIdx.x; i < 3 * H / 4; i += NT) tws[i] = tw[2 * i];
  const float2* gl = g + (size_t)l * N * Kty;
  constexpr int PER = H * B / NT;
  float2 v[PER], e[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    v[j] = gl[(size_t)(2 * (t / B)) * Kty + ky0 + t % B];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) s[threadIdx.x + j * NT] = v[j];
  // the odd rows are loaded while the even half is transformed
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    v[j] = gl[(size_t)(2 * (t / B) + 1) * Kty + ky0 + t % B];
  }
  __syncthreads();
  fft_lds<H, B, NT, true>(s, tws);
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    e[j] = s[t];
    s[t] = v[j];
  }
  __syncthreads();
  fft_lds<H, B, NT, true>(s, tws);
  float2* Tl = T + (size_t)l * Nx * Kty;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int t = threadIdx.x + j * NT;
    const int b = t % B, n = t / B;
    if (n < Nx) {
      float2 w = tw[n];
      w.y = -w.y;
      Tl[(size_t)n * Kty + ky0 + b] = cadd(e[j], cmul(s[t], w));
    }
  }
}

// (2') adjoint y-pass: N = Kty, keep Ny outputs, real part, scale
template <int N>
__global__ __launch_bounds__(nt_contig<N>()) void k_adj_y(const float2* __restrict__ T,
                                                         const float* __restrict__ sy,
                                                         const float* __restrict__ sx, int Ny, int Nx,
                                                         const float2* __restrict__ tw,
                                                         const float* __restrict__ scale,
                                                         float* __restrict__ out) {
  constexpr int NT = nt_contig<N>();
  __shared__ float2 s[N];
  __shared__ float2 tws[TW_LDS_Y<N>];
  const int ix = blockIdx.x, l = blockIdx.y;
  const float2* twp = stage_twiddles<N, NT, TW_LDS_Y<N>>(tw, tws);
  const float2* Tc = T + ((size_t)l * Nx + ix) * N;
  constexpr int PER = (N + NT - 1) / NT;
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = threadIdx.x + j * NT;
    v[j] = i < N ? Tc[i] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < N) s[i] = v[j];
  }
  __syncthreads();
  fft_lds<N, 1, NT, true>(s, twp);
  const float sc = (scale ? scale[l] : 1.f) * sx[ix];
  float* o = out + ((size_t)l * Nx + ix) * Ny;
  for (int i = threadIdx.x; i < Ny; i += NT) o[i] = sc * sy[i] * s[i].x;
}

template <template <int> class K, class... Args>
void dispatch_pow2(int n, Args&&... args) {
  switch (n) {
#define FHS_CASE(v) \
  case v:           \
    K<v>::run(args...); \
    break;
    FHS_CASE(16) FHS_CASE(32) FHS_CASE(64) FHS_CASE(128) FHS_CASE(256) FHS_CASE(512)
    FHS_CASE(1024) FHS_CASE(2048) FHS_CASE(4096)
#undef FHS_CASE
    default:
      throw Error(FHS_ERR_UNSUPPORTED, "nufft: grid size must be a power of two in [16, 4096]");
  }
}

template <int N>
struct RunFwdY {
  static void run(hipStream_t st, int Nx, int L, const float* x, const fhs_nufft* p, float2* T) {
    k_fwd_y<N><<<dim3(Nx, L), nt_contig<N>(), 0, st>>>(x, p->sy, p->sx, p->Ny, p->Nx, p->twy, T);
  }
};
template <int N>
struct RunFwdX {
  static void run(hipStream_t st, int Kty, int L, const float2* T, const fhs_nufft* p, float2* F) {
    if constexpr (N >= X_SPLIT_MIN) {
      if (2 * p->Nx <= N) {
        k_fwd_x_split<N><<<dim3(Kty / b_strided<N>(), L), nt_strided<N>(), 0, st>>>(T, Kty, p->Nx, p->twx, F);
        return;
      }
    }
    k_fwd_x<N><<<dim3(Kty / b_strided<N>(), L), nt_strided<N>(), 0, st>>>(T, Kty, p->Nx, p->twx, F);
  }
};
template <int N>
struct RunAdjX {
  static void run(hipStream_t st, int Kty, int L, const float2* g, const fhs_nufft* p, float2* T) {
    if constexpr (N >= X_SPLIT_MIN) {
      if (2 * p->Nx <= N) {
        k_adj_x_split<N><<<dim3(Kty / b_strided<N>(), L), nt_strided<N>(), 0, st>>>(g, Kty, p->Nx, p->twx, T);
        return;
      }
    }
    k_adj_x<N><<<dim3(Kty / b_strided<N>(), L), nt_strided<N>(), 0, st>>>(g, Kty, p->Nx, p->twx, T);
  }
};
template <int N>
struct RunAdjY {
  static void run(hipStream_t st, int Nx, int L, const float2* T, const fhs_nufft* p,
                  const float* scale, float* out) {
    k_adj_y<N><<<dim3(Nx, L), nt_contig<N>(), 0, st>>>(T, p->sy, p->sx, p->Ny, p->Nx, p->twy, scale, out);
  }
};

float2* twiddles(int n) {
  std::vector<float2> h(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  FHS_HIP(hipMalloc(&d, n * sizeof(float2)));
  FHS_HIP(hipMemcpy(d, h.data(), n * sizeof(float2), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

// Used by the fused data-fidelity step.
void nufft_forward(fhs_ctx* ctx, const fhs_nufft* p, const float* x, int L, float2* Fx) {
  float2* T = (float2*)ctx->ws.get("nufft_T", (size_t)L * p->Nx * p->Kty * sizeof(float2));
  {
    Scope sc(ctx, "fft_fwd_y");
    dispatch_pow2<RunFwdY>(p->Kty, ctx->stream, p->Nx, L, x, p, T);
    check_launch();
  }
  {
    Scope sc(ctx, "fft_fwd_x");
    dispatch_pow2<RunFwdX>(p->Ktx, ctx->stream, p->Kty, L, (const float2*)T, p, Fx);
    check_launch();
  }
}

void nufft_adjoint(fhs_ctx* ctx, const fhs_nufft* p, const float2* g, int L, const float* scale_dev,
                   float* out) {
  float2* T = (float2*)ctx->ws.get("nufft_T", (size_t)L * p->Nx * p->Kty * sizeof(float2));
  {
    Scope sc(ctx, "fft_adj_x");
    dispatch_pow2<RunAdjX>(p->Ktx, ctx->stream, p->Kty, L, g, p, T);
    check_launch();
  }
  {
    Scope sc(ctx, "fft_adj_y");
    dispatch_pow2<RunAdjY>(p->Kty, ctx->stream, p->Nx, L, (const float2*)T, p, scale_dev, out);
    check_launch();
  }
}

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_nufft_create_host(fhs_ctx* ctx, int Ny, int Nx, int Kty, int Ktx, const double* sy,
                          const double* sx, fhs_nufft** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    auto pow2 = [](int n) { return n >= 16 && n <= 4096 && (n & (n - 1)) == 0; };
    FHS_CHECK(pow2(Kty) && pow2(Ktx), "nufft_create: Kty, Ktx must be powers of two in [16, 4096]");
    FHS_CHECK(Ny >= 1 && Nx >= 1 && Ny <= Kty && Nx <= Ktx, "nufft_create: image larger than grid");
    FHS_HIP(hipSetDevice(ctx->device));
    fhs_nufft* p = new fhs_nufft();
    p->device = ctx->device;
    p->Ny = Ny;
    p->Nx = Nx;
    p->Kty = Kty;
    p->Ktx = Ktx;
    std::vector<float> hy(sy, sy + Ny), hx(sx, sx + Nx);
    FHS_HIP(hipMalloc(&p->sy, Ny * sizeof(float)));
    FHS_HIP(hipMalloc(&p->sx, Nx * sizeof(float)));
    FHS_HIP(hipMemcpy(p->sy, hy.data(), Ny * sizeof(float), hipMemcpyHostToDevice));
    FHS_HIP(hipMemcpy(p->sx, hx.data(), Nx * sizeof(float), hipMemcpyHostToDevice));
    p->twy = twiddles(Kty);
    p->twx = twiddles(Ktx);
    *out = p;
  });
}

int fhs_nufft_info(const fhs_nufft* p, int* Ny, int* Nx, int* Kty, int* Ktx) {
  if (!p) return FHS_ERR_ARG;
  if (Ny) *Ny = p->Ny;
  if (Nx) *Nx = p->Nx;
  if (Kty) *Kty = p->Kty;
  if (Ktx) *Ktx = p->Ktx;
  return FHS_OK;
}

int fhs_nufft_destroy(fhs_nufft* p) {
  if (!p) return FHS_OK;
  (void)hipSetDevice(p->device);
  (void)hipFree(p->sy);
  (void)hipFree(p->sx);
  (void)hipFree(p->twy);
  (void)hipFree(p->twx);
  delete p;
  return FHS_OK;
}

int fhs_apply_scaled_fourier_transform(fhs_ctx* ctx, const fhs_nufft* p, const float* x, int L,
                                       void* Fx) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(p && x && Fx && L >= 1 && L <= 65535, "apply_scaled_fourier_transform: bad argument");
    nufft_forward(ctx, p, x, L, (float2*)Fx);
  });
}

int fhs_apply_scaled_fourier_adjoint(fhs_ctx* ctx, const fhs_nufft* p, const void* g, int L,
                                     const float* scale, float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(p && g && out && L >= 1 && L <= 65535, "apply_scaled_fourier_adjoint: bad argument");
    float* sd = nullptr;
    if (scale) {
      sd = (float*)ctx->ws.get("nufft_scale", L * sizeof(float));
      FHS_HIP(hipMemcpyAsync(sd, scale, L * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    }
    nufft_adjoint(ctx, p, (const float2*)g, L, sd, out);
  });
}

}  // extern "C"
