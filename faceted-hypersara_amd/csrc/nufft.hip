// Oversampled scaled 2-D FFT of the NUFFT: A = fft2(pad(sn .* x)) and its
// adjoint, batched over channels.
//
// The grid is Kty x Ktx (column-major, y contiguous); the image Ny x Nx sits
// in its top-left corner.  Each transform is two passes through HBM:
//   forward : (1) y-pass: one workgroup per image column, scale by sn, zero
//                 pad to Kty and FFT -> T (Kty x Nx, only Nx columns exist,
//                 the rest are zero);
//             (2) x-pass: one workgroup per B consecutive grid rows, FFT of
//                 length Ktx over the Nx non-zero columns of T -> Fx.
//   adjoint : (1') x-pass inverse keeping the first Nx outputs -> T;
//             (2') y-pass inverse, keep Ny outputs, * sn, real part, * scale.
// The y-passes use a register-resident radix-16 Stockham FFT whose first
// stage reads the column from global memory and whose last stage writes it
// (two LDS round trips at 4096 points).  The x-passes read and write B
// consecutive complex values per grid column (B = 16 -> 128-byte segments,
// B = 8 -> 64, B = 4 -> 32), keep the B transforms interleaved in LDS
// (s[i*B + b]) so that consecutive lanes touch consecutive LDS words, and
// run radix-4 Stockham stages in LDS (one radix-2 stage when log2 n is
// odd): measured faster than the register form for these strided passes.
// Twiddles come from an FP64-computed table exp(-2 pi i k/n).
#include <cmath>
#include <string>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "nufft.hpp"

namespace fhs {
namespace {

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// ---- LDS radix-4 Stockham FFT (the x-passes) --------------------------------
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
template <int N, int NT, int TWN>
__device__ __forceinline__ const float2* stage_twiddles(const float2* __restrict__ tw, float2* tws) {
  if constexpr (TWN == 1) return tw;
  for (int i = threadIdx.x; i < TWN; i += NT) tws[i] = tw[i];
  return tws;
}

// ---- register-resident Stockham FFT ----------------------------------------
// A length-N transform (N a power of two) is done in stages of radix R = 16
// (then one smaller radix for what remains: 2048 = 16 16 8).  A stage of
// sub-transform size NS takes butterfly j (0 <= j < N/R) with k = j mod NS:
//   v[r] = in[j + r N/R] * W_N^(r k N/(NS R)),  r < R,   (no twiddle at NS = 1)
//   v <- DFT_R(v),  out[(j / NS) NS R + k + r NS] = v[r],
// with W_N = exp(-2 pi i / N) (conjugated for the inverse).  Each thread keeps
// its butterfly's R values in registers; the first stage reads its inputs
// from global memory and the last writes its outputs there (first stage:
// n = j + r N/R; last stage: n = j + r N/R as well, since NS R = N), so a
// 4096-point transform makes two LDS round trips (three barriers) instead
// of the six of radix 4.  B transforms of a workgroup are interleaved in LDS
// (slot n B + b), with one pad slot per 16 rows so that the stride-R writes
// of the early stages do not all fall in one bank.
// cos / sin of 2 pi k / 16, k = 0..7
__device__ constexpr float kC16[8] = {1.f, 0.923879532511287f, 0.707106781186548f, 0.382683432365090f,
                                      0.f, -0.382683432365090f, -0.707106781186548f, -0.923879532511287f};
__device__ constexpr float kS16[8] = {0.f, 0.382683432365090f, 0.707106781186548f, 0.923879532511287f,
                                      1.f, 0.923879532511287f, 0.707106781186548f, 0.382683432365090f};

// in-place DFT of R = 1, 2, 4, 8 or 16 registers (radix-2 decimation in time)
template <int R, bool INV>
__device__ __forceinline__ void dft_reg(float2* v) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R > 2) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    dft_reg<R / 2, INV>(e);
    dft_reg<R / 2, INV>(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const int q = k * (16 / R);   // W_R^k = W_16^q
      float2 t;
      if (q == 0) {
        t = o[k];
      } else if (q == 4) {          // -i forward, +i inverse
        t = INV ? make_float2(-o[k].y, o[k].x) : make_float2(o[k].y, -o[k].x);
      } else {
        const float2 w = make_float2(kC16[q], INV ? kS16[q] : -kS16[q]);
        t = cmul(o[k], w);
      }
      v[k] = cadd(e[k], t);
      v[k + R / 2] = csub(e[k], t);
    }
  }
}

__host__ __device__ constexpr int fft_pad(int n) { return n + (n >> 4); }

template <int N, int B>
constexpr int fft_lds_slots() { return fft_pad(N) * B + B; }

template <int N>
constexpr int fft_radix(int ns) { return N / ns >= 16 ? 16 : N / ns; }

// one stage; load(q, r, b, n) / store(q, r, b, n, v) serve the first / last
// stage (q: the thread's butterfly slot, r: its value index)
template <int N, int R, int NS, int B, int NT, bool INV, bool FIRST, bool LAST, class Load, class Store>
__device__ __forceinline__ void fft_stage(float2* s, const float2* __restrict__ tw, int tws, Load& load,
                                          Store& store) {
  constexpr int NBF = B * N / R;
  constexpr int PER = (NBF + NT - 1) / NT;
  constexpr int TS = N / (NS * R);
  float2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int u = threadIdx.x + q * NT;
    if (u < NBF) {
      const int b = u % B, j = u / B;
      const int k = j & (NS - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int n = j + r * (N / R);
        v[q][r] = FIRST ? load(q, r, b, n) : s[fft_pad(n) * B + b];
      }
      if (NS > 1 && k) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[(r * k * TS) * tws];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_reg<R, INV>(v[q]);
    }
  }
  if constexpr (!FIRST && !LAST) __syncthreads();   // in place: every read of the stage before its writes
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int u = threadIdx.x + q * NT;
    if (u < NBF) {
      const int b = u % B, j = u / B;
      const int k = j & (NS - 1);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int n = base + r * NS;
        if constexpr (LAST)
          store(q, r, b, n, v[q][r]);
        else
          s[fft_pad(n) * B + b] = v[q][r];
      }
    }
  }
  if constexpr (!LAST) __syncthreads();
}

template <int N, int NS, int B, int NT, bool INV, bool FIRST, class Load, class Store>
__device__ __forceinline__ void fft_stages(float2* s, const float2* __restrict__ tw, int tws, Load& load,
                                           Store& store) {
  constexpr int R = fft_radix<N>(NS);
  constexpr bool LAST = NS * R == N;
  fft_stage<N, R, NS, B, NT, INV, FIRST, LAST>(s, tw, tws, load, store);
  if constexpr (!LAST) fft_stages<N, NS * R, B, NT, INV, false>(s, tw, tws, load, store);
}

// B transforms of length N: inputs from load, outputs to store.  tw is the
// twiddle table exp(-2 pi i m / (N tws)) read at stride tws (tws = 2 for the
// half-length transforms of the split x-passes).  s: fft_lds_slots<N, B>()
// complex values of LDS.  Must be called by all NT threads (it has barriers).
template <int N, int B, int NT, bool INV, class Load, class Store>
__device__ __forceinline__ void fft_reg(float2* s, const float2* __restrict__ tw, int tws, Load& load,
                                        Store& store) {
  fft_stages<N, 1, B, NT, INV, true>(s, tw, tws, load, store);
}

// threads of a workgroup running B transforms of length n: one radix-16
// butterfly each (at least one wave, at most 1024)
template <int n, int B>
constexpr int fft_threads() { return B * n / 16 < 64 ? 64 : (B * n / 16 > 1024 ? 1024 : B * n / 16); }

// grid rows per x-pass workgroup: 16 up to N = 512, 8 at N = 1024 (two
// workgroups per CU), 4 from N = 2048 (64 KB of rows plus the twiddles, two
// workgroups per CU; 128 KB and one per CU at N = 4096, where 2 rows, 16-byte
// segments, measured slower).  Segments below 128 bytes are merged in one
// XCD's L2 by xpass_row0.
template <int N>
constexpr int b_strided() { return N >= 2048 ? 4 : (N >= 1024 ? 8 : 16); }
template <int N, int B>
constexpr int nt_strided() {
  return B * N / 4 > 1024 ? 1024 : (B * N / 4 < 64 ? 64 : B * N / 4);
}

// x-pass workgroup -> first grid row.  Workgroups are dispatched to the 8
// XCDs round-robin (linear id mod 8; gridDim.x is a multiple of 8 here, so
// blockIdx.x mod 8).  When B rows are fewer than a 128-byte line (G = 16 / B
// workgroups share each line of every grid column), workgroups b, b + 8, ...,
// b + 8 (G - 1) take G adjacent row groups, so the line is fetched into one
// XCD's L2 instead of G.
template <int B>
__device__ __forceinline__ int xpass_row0() {
  constexpr int G = 16 / B;
  const int b = blockIdx.x;
  if constexpr (G == 1) {
    return b * B;
  } else {
    if (gridDim.x % (8 * G)) return b * B;
    return ((b / (8 * G)) * (8 * G) + (b % 8) * G + (b / 8) % G) * B;
  }
}

// threads of a y-pass workgroup of an N-point grid
template <int N>
constexpr int nt_y() { return fft_threads<N, 1>(); }

// (1) forward y-pass: one workgroup per image column, scaled by sn, zero
// padded to N = Kty
template <int N>
__global__ __launch_bounds__(nt_y<N>()) void k_fwd_y(const float* __restrict__ x,
                                                              const float* __restrict__ sy,
                                                              const float* __restrict__ sx, int Ny, int Nx,
                                                              const float2* __restrict__ tw,
                                                              float2* __restrict__ T) {
  constexpr int NT = fft_threads<N, 1>();
  __shared__ float2 s[fft_lds_slots<N, 1>()];
  const int ix = blockIdx.x, l = blockIdx.y;
  const float* __restrict__ xc = x + ((size_t)l * Nx + ix) * Ny;
  const float scx = sx[ix];
  float2* __restrict__ out = T + ((size_t)l * Nx + ix) * N;
  auto load = [&](int, int, int, int n) {
    return n < Ny ? make_float2(xc[n] * sy[n] * scx, 0.f) : make_float2(0.f, 0.f);
  };
  auto store = [&](int, int, int, int n, float2 v) { out[n] = v; };
  fft_reg<N, 1, NT, false>(s, tw, 1, load, store);
}

// (2) forward x-pass: N = Ktx
template <int N, int B>
__global__ __launch_bounds__((nt_strided<N, B>())) void k_fwd_x(const float2* __restrict__ T, int Kty,
                                                          int Nx, const float2* __restrict__ tw,
                                                          float2* __restrict__ Fx) {
  constexpr int NT = nt_strided<N, B>();
  __shared__ float2 s[N * B];
  __shared__ float2 tws[TW_LDS<N>];
  const int ky0 = xpass_row0<B>(), l = blockIdx.y;
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
template <int N, int B>
__global__ __launch_bounds__((nt_strided<N, B>())) void k_adj_x(const float2* __restrict__ g, int Kty,
                                                          int Nx, const float2* __restrict__ tw,
                                                          float2* __restrict__ T) {
  constexpr int NT = nt_strided<N, B>();
  __shared__ float2 s[N * B];
  __shared__ float2 tws[TW_LDS<N>];
  const int ky0 = xpass_row0<B>(), l = blockIdx.y;
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

template <int N, int B>
__global__ __launch_bounds__((nt_strided<N, B>())) void k_fwd_x_split(const float2* __restrict__ T, int Kty,
                                                                int Nx, const float2* __restrict__ tw,
                                                                float2* __restrict__ Fx) {
  constexpr int H = N / 2;
  constexpr int NT = nt_strided<N, B>();
  static_assert(H * B % NT == 0, "split x-pass: whole rows per thread");
  __shared__ float2 s[H * B];
  __shared__ float2 tws[3 * H / 4];
  const int ky0 = xpass_row0<B>(), l = blockIdx.y;
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

template <int N, int B>
__global__ __launch_bounds__((nt_strided<N, B>())) void k_adj_x_split(const float2* __restrict__ g, int Kty,
                                                                int Nx, const float2* __restrict__ tw,
                                                                float2* __restrict__ T) {
  constexpr int H = N / 2;
  constexpr int NT = nt_strided<N, B>();
  static_assert(H * B % NT == 0, "split x-pass: whole rows per thread");
  __shared__ float2 s[H * B];
  __shared__ float2 tws[3 * H / 4];
  const int ky0 = xpass_row0<B>(), l = blockIdx.y;
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
__global__ __launch_bounds__(nt_y<N>()) void k_adj_y(const float2* __restrict__ T,
                                                              const float* __restrict__ sy,
                                                              const float* __restrict__ sx, int Ny, int Nx,
                                                              const float2* __restrict__ tw,
                                                              const float* __restrict__ scale,
                                                              float* __restrict__ out) {
  constexpr int NT = fft_threads<N, 1>();
  __shared__ float2 s[fft_lds_slots<N, 1>()];
  const int ix = blockIdx.x, l = blockIdx.y;
  const float2* __restrict__ Tc = T + ((size_t)l * Nx + ix) * N;
  const float sc = (scale ? scale[l] : 1.f) * sx[ix];
  float* __restrict__ o = out + ((size_t)l * Nx + ix) * Ny;
  auto load = [&](int, int, int, int n) { return Tc[n]; };
  auto store = [&](int, int, int, int n, float2 v) {
    if (n < Ny) o[n] = sc * sy[n] * v.x;
  };
  fft_reg<N, 1, NT, true>(s, tw, 1, load, store);
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
    k_fwd_y<N><<<dim3(Nx, L), nt_y<N>(), 0, st>>>(x, p->sy, p->sx, p->Ny, p->Nx, p->twy, T);
  }
};
// Rows per x-pass workgroup: the default first, then the autotuner's other
// candidates (those whose rows fit in LDS).
template <int N, bool SPLIT>
struct XRows {
  static constexpr int n = N >= 4096 ? (SPLIT ? 3 : 2) : N == 2048 ? 3 : N == 1024 ? 3 : N == 512 ? 2 : 1;
  static constexpr int at(int i) {
    if (N >= 4096) return i == 0 ? 4 : i == 1 ? 2 : 8;
    if (N == 2048) return i == 0 ? 4 : i == 1 ? 2 : 8;
    if (N == 1024) return i == 0 ? 8 : i == 1 ? 4 : 16;
    if (N == 512) return i == 0 ? 16 : 8;
    return 16;
  }
};

template <int N, bool SPLIT, class F>
void with_rows(int i, F&& f) {
  using R = XRows<N, SPLIT>;
  switch (i) {
    case 0: f(std::integral_constant<int, R::at(0)>{}); break;
    case 1: if constexpr (R::n > 1) f(std::integral_constant<int, R::at(1)>{}); break;
    default: if constexpr (R::n > 2) f(std::integral_constant<int, R::at(2)>{}); break;
  }
}

template <int N, bool SPLIT>
const char* const* xrow_labels() {
  static const std::string lab[3] = {std::to_string(XRows<N, SPLIT>::at(0)) + " rows",
                                     std::to_string(XRows<N, SPLIT>::at(1)) + " rows",
                                     std::to_string(XRows<N, SPLIT>::at(2)) + " rows"};
  static const char* const p[3] = {lab[0].c_str(), lab[1].c_str(), lab[2].c_str()};
  return p;
}

inline std::string xpass_key(const char* k, int N, int Kty, int Nx, int L) {
  return std::string(k) + "/N" + std::to_string(N) + "xKty" + std::to_string(Kty) + "xNx" + std::to_string(Nx) +
         "xL" + std::to_string(L);
}

template <int N>
struct RunFwdX {
  static void run(fhs_ctx* ctx, int Kty, int L, const float2* T, const fhs_nufft* p, float2* F) {
    const bool split = N >= X_SPLIT_MIN && 2 * p->Nx <= N;
    auto go = [&](auto split_c) {
      constexpr bool SP = decltype(split_c)::value;
      tune_launch(ctx, xpass_key("fft_fwd_x", N, Kty, p->Nx, L), XRows<N, SP>::n, 0, xrow_labels<N, SP>(),
                  [&](int c) {
                    with_rows<N, SP>(c, [&](auto b_c) {
                      constexpr int B = decltype(b_c)::value;
                      if constexpr (SP)
                        k_fwd_x_split<N, B><<<dim3(Kty / B, L), nt_strided<N, B>(), 0, ctx->stream>>>(T, Kty, p->Nx,
                                                                                                    p->twx, F);
                      else
                        k_fwd_x<N, B><<<dim3(Kty / B, L), nt_strided<N, B>(), 0, ctx->stream>>>(T, Kty, p->Nx,
                                                                                              p->twx, F);
                      check_launch();
                    });
                  });
    };
    if constexpr (N >= X_SPLIT_MIN) {
      if (split) {
        go(std::true_type{});
        return;
      }
    }
    go(std::false_type{});
  }
};
template <int N>
struct RunAdjX {
  static void run(fhs_ctx* ctx, int Kty, int L, const float2* g, const fhs_nufft* p, float2* T) {
    const bool split = N >= X_SPLIT_MIN && 2 * p->Nx <= N;
    auto go = [&](auto split_c) {
      constexpr bool SP = decltype(split_c)::value;
      tune_launch(ctx, xpass_key("fft_adj_x", N, Kty, p->Nx, L), XRows<N, SP>::n, 0, xrow_labels<N, SP>(),
                  [&](int c) {
                    with_rows<N, SP>(c, [&](auto b_c) {
                      constexpr int B = decltype(b_c)::value;
                      if constexpr (SP)
                        k_adj_x_split<N, B><<<dim3(Kty / B, L), nt_strided<N, B>(), 0, ctx->stream>>>(g, Kty, p->Nx,
                                                                                                    p->twx, T);
                      else
                        k_adj_x<N, B><<<dim3(Kty / B, L), nt_strided<N, B>(), 0, ctx->stream>>>(g, Kty, p->Nx,
                                                                                              p->twx, T);
                      check_launch();
                    });
                  });
    };
    if constexpr (N >= X_SPLIT_MIN) {
      if (split) {
        go(std::true_type{});
        return;
      }
    }
    go(std::false_type{});
  }
};
template <int N>
struct RunAdjY {
  static void run(hipStream_t st, int Nx, int L, const float2* T, const fhs_nufft* p,
                  const float* scale, float* out) {
    k_adj_y<N><<<dim3(Nx, L), nt_y<N>(), 0, st>>>(T, p->sy, p->sx, p->Ny, p->Nx, p->twy, scale, out);
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

// ---- grids that are not powers of two: runtime-length mixed radix ---------
// Grid sizes N = 2^a 3^b 5^c (an oversampled 2560-pixel axis is 5120 =
// 16 16 4 5) run the same register Stockham stages with the length known at
// run time: radix 16 while N/NS divides by 16, then the remaining power of
// two, then radix 5 and 3 stages (dft3 / dft5 below).  Each thread holds up
// to ceil(16 / R) butterflies of a radix-R stage; the x-passes use the same
// core on B rows per workgroup (no split form).
__device__ __forceinline__ int fft_radix_rt(int rem) {
  if (rem % 16 == 0) return 16;
  const int p2 = rem & -rem;
  if (p2 > 1) return p2;
  return rem % 5 == 0 ? 5 : 3;
}

template <bool INV>
__device__ __forceinline__ float2 mul_i(float2 z, float s) {   // z * (-i s) forward, z * (+i s) inverse
  return INV ? make_float2(-z.y * s, z.x * s) : make_float2(z.y * s, -z.x * s);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2* x) {
  const float c = -0.5f, s = 0.866025403784438647f;   // cos, sin of 2 pi / 3
  const float2 t1 = cadd(x[1], x[2]);
  const float2 t2 = mul_i<INV>(csub(x[1], x[2]), s);
  const float2 u = make_float2(x[0].x + c * t1.x, x[0].y + c * t1.y);
  x[0] = cadd(x[0], t1);
  x[1] = cadd(u, t2);
  x[2] = csub(u, t2);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* x) {
  const float c1 = 0.309016994374947424f, s1 = 0.951056516295153572f;    // 2 pi / 5
  const float c2 = -0.809016994374947424f, s2 = 0.587785252292473129f;   // 4 pi / 5
  const float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
  const float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
  const float2 u1 = make_float2(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
  const float2 u2 = make_float2(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
  const float2 t1 = mul_i<INV>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), 1.f);
  const float2 t2 = mul_i<INV>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), 1.f);
  x[0] = cadd(x[0], cadd(a1, a2));
  x[1] = cadd(u1, t1);
  x[2] = cadd(u2, t2);
  x[3] = csub(u2, t2);
  x[4] = csub(u1, t1);
}

template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2* v) {
  if constexpr (R == 3)
    dft3<INV>(v);
  else if constexpr (R == 5)
    dft5<INV>(v);
  else
    dft_reg<R, INV>(v);
}

template <int R, bool INV, class Load, class Store>
__device__ __forceinline__ void fft_stage_rt(float2* s, const float2* __restrict__ tw, int N, int NS, int B, int NT,
                                             bool first, bool last, Load& load, Store& store) {
  constexpr int PMAX = (16 + R - 1) / R;   // butterflies per thread (NT >= B N / 16)
  const int NBF = B * N / R;
  const int TS = N / (NS * R);
  float2 v[PMAX][R];
#pragma unroll
  for (int q = 0; q < PMAX; ++q) {
    const int u = threadIdx.x + q * NT;
    if (u < NBF) {
      const int b = u % B, j = u / B;
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int n = j + r * (N / R);
        v[q][r] = first ? load(q, r, b, n) : s[fft_pad(n) * B + b];
      }
      if (k) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * TS];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_any<R, INV>(v[q]);
    }
  }
  if (!first && !last) __syncthreads();
#pragma unroll
  for (int q = 0; q < PMAX; ++q) {
    const int u = threadIdx.x + q * NT;
    if (u < NBF) {
      const int b = u % B, j = u / B;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int n = base + r * NS;
        if (last)
          store(q, r, b, n, v[q][r]);
        else
          s[fft_pad(n) * B + b] = v[q][r];
      }
    }
  }
  if (!last) __syncthreads();
}

template <bool INV, class Load, class Store>
__device__ void fft_rt(float2* s, const float2* __restrict__ tw, int N, int B, int NT, Load& load, Store& store) {
  int ns = 1;
  bool first = true;
  while (ns < N) {
    const int R = fft_radix_rt(N / ns);
    const bool last = ns * R == N;
    switch (R) {
      case 16: fft_stage_rt<16, INV>(s, tw, N, ns, B, NT, first, last, load, store); break;
      case 8: fft_stage_rt<8, INV>(s, tw, N, ns, B, NT, first, last, load, store); break;
      case 4: fft_stage_rt<4, INV>(s, tw, N, ns, B, NT, first, last, load, store); break;
      case 2: fft_stage_rt<2, INV>(s, tw, N, ns, B, NT, first, last, load, store); break;
      case 5: fft_stage_rt<5, INV>(s, tw, N, ns, B, NT, first, last, load, store); break;
      default: fft_stage_rt<3, INV>(s, tw, N, ns, B, NT, first, last, load, store); break;
    }
    ns *= R;
    first = false;
  }
}

// threads and LDS bytes of a runtime-length workgroup of B transforms
inline int fft_rt_threads(int N, int B) {
  const int t = (B * N + 15) / 16;
  return std::max(64, (t + 63) / 64 * 64);
}
inline size_t fft_rt_lds(int N, int B) { return ((size_t)fft_pad(N) * B + B) * sizeof(float2); }
// rows per runtime x-pass workgroup: 4, 2 or 1, dividing the other axis, with
// one radix-16 butterfly per thread (B N <= 16384) and the rows in LDS
inline int fft_rt_rows(int N, int other) {
  for (int B : {4, 2, 1})
    if (other % B == 0 && B * N <= 16384 && fft_rt_lds(N, B) <= 150 * 1024) return B;
  return 1;
}

__global__ __launch_bounds__(1024) void k_fwd_y_rt(const float* __restrict__ x, const float* __restrict__ sy,
                                                   const float* __restrict__ sx, int Ny, int Nx, int N,
                                                   const float2* __restrict__ tw, float2* __restrict__ T) {
  extern __shared__ __align__(16) float2 s_rt[];
  const int ix = blockIdx.x, l = blockIdx.y;
  const float* __restrict__ xc = x + ((size_t)l * Nx + ix) * Ny;
  const float scx = sx[ix];
  float2* __restrict__ out = T + ((size_t)l * Nx + ix) * N;
  auto load = [&](int, int, int, int n) {
    return n < Ny ? make_float2(xc[n] * sy[n] * scx, 0.f) : make_float2(0.f, 0.f);
  };
  auto store = [&](int, int, int, int n, float2 v) { out[n] = v; };
  fft_rt<false>(s_rt, tw, N, 1, blockDim.x, load, store);
}

__global__ __launch_bounds__(1024) void k_adj_y_rt(const float2* __restrict__ T, const float* __restrict__ sy,
                                                   const float* __restrict__ sx, int Ny, int Nx, int N,
                                                   const float2* __restrict__ tw, const float* __restrict__ scale,
                                                   float* __restrict__ out) {
  extern __shared__ __align__(16) float2 s_rt[];
  const int ix = blockIdx.x, l = blockIdx.y;
  const float2* __restrict__ Tc = T + ((size_t)l * Nx + ix) * N;
  const float sc = (scale ? scale[l] : 1.f) * sx[ix];
  float* __restrict__ o = out + ((size_t)l * Nx + ix) * Ny;
  auto load = [&](int, int, int, int n) { return Tc[n]; };
  auto store = [&](int, int, int, int n, float2 v) {
    if (n < Ny) o[n] = sc * sy[n] * v.x;
  };
  fft_rt<true>(s_rt, tw, N, 1, blockDim.x, load, store);
}

__global__ __launch_bounds__(1024) void k_fwd_x_rt(const float2* __restrict__ T, int Kty, int Nx, int N, int B,
                                                   const float2* __restrict__ tw, float2* __restrict__ Fx) {
  extern __shared__ __align__(16) float2 s_rt[];
  const int ky0 = blockIdx.x * B, l = blockIdx.y;
  const float2* __restrict__ Tl = T + (size_t)l * Nx * Kty;
  float2* __restrict__ F = Fx + (size_t)l * N * Kty;
  auto load = [&](int, int, int b, int n) {
    return n < Nx ? Tl[(size_t)n * Kty + ky0 + b] : make_float2(0.f, 0.f);
  };
  auto store = [&](int, int, int b, int n, float2 v) { F[(size_t)n * Kty + ky0 + b] = v; };
  fft_rt<false>(s_rt, tw, N, B, blockDim.x, load, store);
}

__global__ __launch_bounds__(1024) void k_adj_x_rt(const float2* __restrict__ g, int Kty, int Nx, int N, int B,
                                                   const float2* __restrict__ tw, float2* __restrict__ T) {
  extern __shared__ __align__(16) float2 s_rt[];
  const int ky0 = blockIdx.x * B, l = blockIdx.y;
  const float2* __restrict__ gl = g + (size_t)l * N * Kty;
  float2* __restrict__ Tl = T + (size_t)l * Nx * Kty;
  auto load = [&](int, int, int b, int n) { return gl[(size_t)n * Kty + ky0 + b]; };
  auto store = [&](int, int, int b, int n, float2 v) {
    if (n < Nx) Tl[(size_t)n * Kty + ky0 + b] = v;
  };
  fft_rt<true>(s_rt, tw, N, B, blockDim.x, load, store);
}

inline bool is_pow2_grid(int n) { return n >= 16 && n <= 4096 && (n & (n - 1)) == 0; }

void fwd_y_any(fhs_ctx* ctx, const fhs_nufft* p, int L, const float* x, float2* T) {
  if (is_pow2_grid(p->Kty)) {
    dispatch_pow2<RunFwdY>(p->Kty, ctx->stream, p->Nx, L, x, p, T);
    return;
  }
  const int nt = fft_rt_threads(p->Kty, 1);
  const size_t lds = fft_rt_lds(p->Kty, 1);
  allow_dynamic_lds((const void*)k_fwd_y_rt, lds);
  k_fwd_y_rt<<<dim3(p->Nx, L), nt, lds, ctx->stream>>>(x, p->sy, p->sx, p->Ny, p->Nx, p->Kty, p->twy, T);
}

void adj_y_any(fhs_ctx* ctx, const fhs_nufft* p, int L, const float2* T, const float* scale, float* out) {
  if (is_pow2_grid(p->Kty)) {
    dispatch_pow2<RunAdjY>(p->Kty, ctx->stream, p->Nx, L, T, p, scale, out);
    return;
  }
  const int nt = fft_rt_threads(p->Kty, 1);
  const size_t lds = fft_rt_lds(p->Kty, 1);
  allow_dynamic_lds((const void*)k_adj_y_rt, lds);
  k_adj_y_rt<<<dim3(p->Nx, L), nt, lds, ctx->stream>>>(T, p->sy, p->sx, p->Ny, p->Nx, p->Kty, p->twy, scale, out);
}

void fwd_x_any(fhs_ctx* ctx, const fhs_nufft* p, int L, const float2* T, float2* F) {
  if (is_pow2_grid(p->Ktx)) {
    dispatch_pow2<RunFwdX>(p->Ktx, ctx, p->Kty, L, T, p, F);
    return;
  }
  const int B = fft_rt_rows(p->Ktx, p->Kty);
  const int nt = fft_rt_threads(p->Ktx, B);
  const size_t lds = fft_rt_lds(p->Ktx, B);
  allow_dynamic_lds((const void*)k_fwd_x_rt, lds);
  k_fwd_x_rt<<<dim3(p->Kty / B, L), nt, lds, ctx->stream>>>(T, p->Kty, p->Nx, p->Ktx, B, p->twx, F);
}

void adj_x_any(fhs_ctx* ctx, const fhs_nufft* p, int L, const float2* g, float2* T) {
  if (is_pow2_grid(p->Ktx)) {
    dispatch_pow2<RunAdjX>(p->Ktx, ctx, p->Kty, L, g, p, T);
    return;
  }
  const int B = fft_rt_rows(p->Ktx, p->Kty);
  const int nt = fft_rt_threads(p->Ktx, B);
  const size_t lds = fft_rt_lds(p->Ktx, B);
  allow_dynamic_lds((const void*)k_adj_x_rt, lds);
  k_adj_x_rt<<<dim3(p->Kty / B, L), nt, lds, ctx->stream>>>(g, p->Kty, p->Nx, p->Ktx, B, p->twx, T);
}

}  // namespace

// Used by the fused data-fidelity step.
void nufft_forward(fhs_ctx* ctx, const fhs_nufft* p, const float* x, int L, float2* Fx) {
  float2* T = (float2*)ctx->ws.get("nufft_T", (size_t)L * p->Nx * p->Kty * sizeof(float2));
  {
    Scope sc(ctx, "fft_fwd_y");
    fwd_y_any(ctx, p, L, x, T);
    check_launch();
  }
  {
    Scope sc(ctx, "fft_fwd_x");
    fwd_x_any(ctx, p, L, (const float2*)T, Fx);
    check_launch();
  }
}

void nufft_adjoint(fhs_ctx* ctx, const fhs_nufft* p, const float2* g, int L, const float* scale_dev,
                   float* out) {
  float2* T = (float2*)ctx->ws.get("nufft_T", (size_t)L * p->Nx * p->Kty * sizeof(float2));
  {
    Scope sc(ctx, "fft_adj_x");
    adj_x_any(ctx, p, L, g, T);
    check_launch();
  }
  {
    Scope sc(ctx, "fft_adj_y");
    adj_y_any(ctx, p, L, (const float2*)T, scale_dev, out);
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
    auto smooth = [](int n) {   // 2^a 3^b 5^c in [16, 8192]
      if (n < 16 || n > 8192) return false;
      for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
      return n == 1;
    };
    FHS_CHECK(smooth(Kty) && smooth(Ktx), "nufft_create: Kty, Ktx must be of the form 2^a 3^b 5^c in [16, 8192]");
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
