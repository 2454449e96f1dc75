// SARA dictionary on the GPU: Psit (analysis) and Psi (synthesis, exact
// adjoint) of db1..db8 + Dirac with 'zpd' extension, batched over bases and
// channels, and the weighted l21 dual step.
//
// Coefficient vector of one channel (length s, MATLAB wavedec2 order per
// basis, bases in order, Dirac last), each basis scaled by 1/sqrt(P):
//   [ A_J H_J V_J D_J H_{J-1} ... D_1 ]_db1 ... [ ... ]_dbK [ x(:) ]_dirac
// One zpd analysis step along an axis with reconstruction filter f (length
// lf): out[k] = sum_j f[j] in[2k + 2 - lf + j], n_out = floor((n + lf - 1)/2).
// a = lo_y lo_x, h = hi_y lo_x, v = lo_y hi_x, d = hi_y hi_x (y = rows).
//
// Kernels work on output tiles (analysis 32 x 16, synthesis 32 x 32): the
// input window (plus filter halo) is staged in LDS, filtered along x then y
// (separable, 2 lf FMAs per output).  Analysis runs one workgroup per (tile,
// basis, channel) at every level; its 32 x 16 tiles keep the LDS at 24 KB,
// so five workgroups fit on a CU instead of three with 32 x 32 tiles (one
// workgroup per tile running all the bases from one image window measured
// slower: the bases then run one after the other).  Synthesis mirrors the analysis,
// and its finest level sums all bases and the Dirac part in registers so the
// image is written once.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

namespace fhs {

constexpr int WMAX_BASES = 8;
constexpr int WMAX_LEVELS = 6;
constexpr int WMAX_LF = 16;
constexpr int WT = 32;           // output tile edge
constexpr int WTHREADS = 256;

struct SaraGeom {
  int Ny, Nx, nlevel, nb, dirac;
  float norm;                            // 1/sqrt(P)
  int lf[WMAX_BASES];
  int ny[WMAX_BASES][WMAX_LEVELS + 1];   // level shapes, [0] = image
  int nx[WMAX_BASES][WMAX_LEVELS + 1];
  int64_t boff[WMAX_BASES + 1];          // basis offsets in a channel vector; boff[nb] = dirac
  int64_t aoff[WMAX_BASES];              // A_J offset inside basis (= 0)
  int64_t doff[WMAX_BASES][WMAX_LEVELS + 1];  // offset of H_k (V, D follow) inside basis
  int64_t scr[WMAX_BASES][WMAX_LEVELS + 1];   // scratch offset of approx level k (per channel)
  int64_t scr_per_chan;
  int64_t s;                             // coefficients per channel
  float lo[WMAX_BASES][WMAX_LF];
  float hi[WMAX_BASES][WMAX_LF];
};

}  // namespace fhs

struct fhs_sara {
  int device = 0;
  fhs::SaraGeom g{};
};

namespace fhs {
namespace {

// ---- analysis of one T x T output tile from an LDS window -----------------
// win: window origin at input (wy0, wx0), height WH (column-major: [x][y]);
// the basis' own window starts at offset o (longest-filter window at level 1).
struct AnOut {
  float* v1;          // channel vector (coefficients accumulate here)
  float* approx;      // scratch for A_k (nullptr at the last level)
  int64_t a_off;      // offset of A_k in approx (or of A_J in v1 at the last level)
  int64_t h_off;      // offset of H_k in v1 (V, D follow)
  int ny, nx;         // output level shape
  bool last;
};

// analysis tiles: AY x AX outputs (rows x columns)
constexpr int AY = 32, AX = 16;
constexpr int AN_PER = AY * AX / WTHREADS;   // outputs per thread

// old values of the coefficients a tile accumulates into (+=), loaded at
// kernel start so that their latency overlaps the window load
struct AnPre {
  float a[AN_PER], h[AN_PER], v[AN_PER], d[AN_PER];
};

__device__ __forceinline__ void an_prefetch(int ky0, int kx0, const AnOut& out, AnPre& pre) {
#pragma unroll
  for (int q = 0; q < AN_PER; ++q) {
    const int t = threadIdx.x + q * WTHREADS;
    const int gy = ky0 + t % AY, gx = kx0 + t / AY;
    pre.a[q] = pre.h[q] = pre.v[q] = pre.d[q] = 0.f;
    if (gy < out.ny && gx < out.nx) {
      const int64_t p = gy + (int64_t)out.ny * gx;
      const int64_t sz = (int64_t)out.ny * out.nx;
      if (out.last) pre.a[q] = out.v1[out.a_off + p];
      pre.h[q] = out.v1[out.h_off + p];
      pre.v[q] = out.v1[out.h_off + sz + p];
      pre.d[q] = out.v1[out.h_off + 2 * sz + p];
    }
  }
}

template <int LF>
__device__ void an_tile_t(const float* win, int WH, const float* lo, const float* hi, int ky0, int kx0, float* Lx,
                          float* Hx, const AnOut& out, const AnPre& pre) {
  constexpr int H = 2 * AY + LF - 2;   // rows of the window
  float rlo[LF], rhi[LF];
#pragma unroll
  for (int j = 0; j < LF; ++j) {
    rlo[j] = lo[j];
    rhi[j] = hi[j];
  }
  // x pass: Lx[kx][r], Hx[kx][r] for r < H, kx < AX
  for (int t = threadIdx.x; t < H * AX; t += WTHREADS) {
    const int r = t % H, kx = t / H;
    float sl = 0.f, sh = 0.f;
    const float* col = win + (size_t)(2 * kx) * WH + r;
#pragma unroll
    for (int j = 0; j < LF; ++j) {
      const float w = col[(size_t)j * WH];
      sl += rlo[j] * w;
      sh += rhi[j] * w;
    }
    Lx[kx * H + r] = sl;
    Hx[kx * H + r] = sh;
  }
  __syncthreads();
  // y pass
#pragma unroll
  for (int q = 0; q < AN_PER; ++q) {
    const int t = threadIdx.x + q * WTHREADS;
    const int ky = t % AY, kx = t / AY;
    const int gy = ky0 + ky, gx = kx0 + kx;
    if (gy < out.ny && gx < out.nx) {
      const float* lc = Lx + kx * H + 2 * ky;
      const float* hc = Hx + kx * H + 2 * ky;
      float a = 0.f, h = 0.f, v = 0.f, d = 0.f;
#pragma unroll
      for (int j = 0; j < LF; ++j) {
        a += rlo[j] * lc[j];
        h += rhi[j] * lc[j];
        v += rlo[j] * hc[j];
        d += rhi[j] * hc[j];
      }
      const int64_t p = gy + (int64_t)out.ny * gx;
      const int64_t sz = (int64_t)out.ny * out.nx;
      if (out.last)
        out.v1[out.a_off + p] = pre.a[q] + a;
      else
        out.approx[out.a_off + p] = a;
      out.v1[out.h_off + p] = pre.h[q] + h;
      out.v1[out.h_off + sz + p] = pre.v[q] + v;
      out.v1[out.h_off + 2 * sz + p] = pre.d[q] + d;
    }
  }
  __syncthreads();
}

// filter lengths of db1..db8 are 2, 4, ..., 16: one unrolled instance each
__device__ void an_tile(const float* win, int WH, int lf, const float* lo, const float* hi, int ky0, int kx0,
                        float* Lx, float* Hx, const AnOut& out, const AnPre& pre) {
  switch (lf) {
    case 2: an_tile_t<2>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    case 4: an_tile_t<4>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    case 6: an_tile_t<6>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    case 8: an_tile_t<8>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    case 10: an_tile_t<10>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    case 12: an_tile_t<12>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    case 14: an_tile_t<14>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
    default: an_tile_t<16>(win, WH, lo, hi, ky0, kx0, Lx, Hx, out, pre); break;
  }
}

constexpr int WIN_ROWS = 2 * AY + WMAX_LF - 2;   // 78
constexpr int WIN_COLS = 2 * AX + WMAX_LF - 2;   // 46
constexpr int WIN_PER_THREAD = (WIN_ROWS * WIN_COLS + WTHREADS - 1) / WTHREADS;

// load window [wy0, wy0+WH) x [wx0, wx0+WW) of a column-major ny x nx image;
// every load of a thread is issued before the first LDS store, so the
// window arrives in one memory latency instead of one per element
__device__ void load_window(float* win, const float* img, int ny, int nx, int wy0, int wx0, int WH,
                            int WW, float scale) {
  float v[WIN_PER_THREAD];
  const int n = WH * WW;
#pragma unroll
  for (int i = 0; i < WIN_PER_THREAD; ++i) {
    const int t = threadIdx.x + i * WTHREADS;
    v[i] = 0.f;
    if (t < n) {
      const int r = t % WH, c = t / WH;
      const int gy = wy0 + r, gx = wx0 + c;
      if (gy >= 0 && gy < ny && gx >= 0 && gx < nx) v[i] = img[gy + (int64_t)ny * gx];
    }
  }
#pragma unroll
  for (int i = 0; i < WIN_PER_THREAD; ++i) {
    const int t = threadIdx.x + i * WTHREADS;
    if (t < n) win[t] = v[i] * scale;
  }
  __syncthreads();
}

__global__ __launch_bounds__(WTHREADS) void k_an_lk(SaraGeom G, int k, const float* __restrict__ x,
                                                    float* __restrict__ v1, float* __restrict__ scr) {
  __shared__ float win[WIN_ROWS * WIN_COLS];
  __shared__ float Lx[AX * WIN_ROWS];
  __shared__ float Hx[AX * WIN_ROWS];
  __shared__ float flo[WMAX_LF], fhi[WMAX_LF];
  const int b = blockIdx.z % G.nb, l = blockIdx.z / G.nb;
  const int ky0 = blockIdx.x * AY, kx0 = blockIdx.y * AX;
  if (ky0 >= G.ny[b][k] || kx0 >= G.nx[b][k]) return;
  const int lf = G.lf[b];
  if (threadIdx.x < lf) {
    flo[threadIdx.x] = G.lo[b][threadIdx.x];
    fhi[threadIdx.x] = G.hi[b][threadIdx.x];
  }
  const int WH = 2 * AY + lf - 2, WW = 2 * AX + lf - 2;   // window rows, columns
  float* sc = scr + (int64_t)l * G.scr_per_chan;
  AnOut out;
  out.v1 = v1 + (int64_t)l * G.s;
  out.last = k == G.nlevel;
  out.approx = sc;
  out.a_off = out.last ? G.boff[b] + G.aoff[b] : G.scr[b][k];
  out.h_off = G.boff[b] + G.doff[b][k];
  out.ny = G.ny[b][k];
  out.nx = G.nx[b][k];
  AnPre pre;
  an_prefetch(ky0, kx0, out, pre);
  if (k == 1)   // level 1 reads the image, scaled by 1/sqrt(P)
    load_window(win, x + (int64_t)l * G.Ny * G.Nx, G.Ny, G.Nx, 2 * ky0 + 2 - lf, 2 * kx0 + 2 - lf, WH, WW,
                G.norm);
  else
    load_window(win, sc + G.scr[b][k - 1], G.ny[b][k - 1], G.nx[b][k - 1], 2 * ky0 + 2 - lf,
                2 * kx0 + 2 - lf, WH, WW, 1.f);
  an_tile(win, WH, lf, flo, fhi, ky0, kx0, Lx, Hx, out, pre);
}

// ---- synthesis -------------------------------------------------------------
// Output tile [iy0, iy0+WT) x [ix0, ix0+WT) of the finer level (shape py x px)
// from the four subbands of level k (shape ny x nx).  Coefficient rows that
// reach the tile: k in [ceil((i0-1)/2), floor((i0+WT-1+lf-2)/2)].
constexpr int CMAX = WT / 2 + WMAX_LF / 2 + 1;   // 25

struct SynIn {
  const float* a;   // approx of level k (column-major ny x nx)
  const float* h;   // H_k, V_k = h + sz, D_k = h + 2 sz
  int ny, nx;
};

// A synthesis tile in three steps, so that a caller can overlap the global
// loads of the next tile with the arithmetic of the current one:
//   syn_load_t   issues the loads of the 4 subband tiles (CN x CN) into registers,
//   syn_store_t  writes them to LDS (cs[s][cx][cy]; a barrier must follow),
//   syn_compute_t accumulates the tile (WT*WT/WTHREADS values per thread, placed
//                 as syn_acc_coord says) into acc.
// Output i takes the LF/2 coefficients k = floor(i/2) + m (m < LF/2) with
// filter tap j = (i & 1) + LF - 2 - 2m; the staged tile is zero beyond the
// subband, so no clamping is needed.
constexpr int CS_MAX = (4 * CMAX * CMAX + WTHREADS - 1) / WTHREADS;   // registers per thread

template <int LF>
__device__ __forceinline__ void syn_load_t(const SynIn& in, int iy0, int ix0, float* cv) {
  constexpr int CN = WT / 2 + LF / 2 + 1;
  constexpr int NCS = 4 * CN * CN;
  const int cy0 = iy0 / 2, cx0 = ix0 / 2;
  const int64_t sz = (int64_t)in.ny * in.nx;
#pragma unroll
  for (int i = 0; i < CS_MAX; ++i) {
    const int t = threadIdx.x + i * WTHREADS;
    cv[i] = 0.f;
    if (t < NCS) {
      const int cyl = t % CN, cxl = (t / CN) % CN, s = t / (CN * CN);
      const int cy = cy0 + cyl, cx = cx0 + cxl;
      if (cy < in.ny && cx < in.nx) {
        const int64_t p = cy + (int64_t)in.ny * cx;
        cv[i] = s == 0 ? in.a[p] : in.h[(s - 1) * sz + p];
      }
    }
  }
}

template <int LF>
__device__ __forceinline__ void syn_store_t(const float* cv, float* cs) {
  constexpr int CN = WT / 2 + LF / 2 + 1;
  constexpr int NCS = 4 * CN * CN;
#pragma unroll
  for (int i = 0; i < CS_MAX; ++i) {
    const int t = threadIdx.x + i * WTHREADS;
    if (t < NCS) cs[t] = cv[i];
  }
}

// Output coordinates of a thread's accumulator entry q of syn_compute_t: the
// thread makes the column pair ixl = 2c, 2c + 1 of row iyl for the pair
// items t = threadIdx.x + (q / 2) WTHREADS (c = t / WT, iyl = t % WT).
__device__ __forceinline__ void syn_acc_coord(int q, int& iyl, int& ixl) {
  const int t = threadIdx.x + (q >> 1) * WTHREADS;
  iyl = t % WT;
  ixl = 2 * (t / WT) + (q & 1);
}

// expects cs staged and the caller's filter writes visible (after a barrier).
// The two outputs of a parity pair (2u, 2u + 1) take the same LF/2
// coefficients with the even and the odd taps, so each item makes both and
// reads every staged value once for two outputs (half the LDS reads).
template <int LF>
__device__ __forceinline__ void syn_compute_t(const float* lo, const float* hi, int iy0, int ix0, int py, int px,
                                              const float* cs, float* Ly, float* Hy, float* acc) {
  constexpr int CN = WT / 2 + LF / 2 + 1;
  float rlo[LF], rhi[LF];
#pragma unroll
  for (int j = 0; j < LF; ++j) {
    rlo[j] = lo[j];
    rhi[j] = hi[j];
  }
  // y pass: Ly[cx][iy] = sum_k lo a[k][cx] + hi h[k][cx];  Hy: v, d
  // (iy0 is even: output 2u + par takes k = u + m with tap par + LF - 2 - 2m)
  for (int t = threadIdx.x; t < (WT / 2) * CN; t += WTHREADS) {
    const int u = t % (WT / 2), cxl = t / (WT / 2);
    const float* a0 = cs + (0 * CN + cxl) * CN + u;
    const float* h0 = cs + (1 * CN + cxl) * CN + u;
    const float* v0 = cs + (2 * CN + cxl) * CN + u;
    const float* d0 = cs + (3 * CN + cxl) * CN + u;
    float sl0 = 0.f, sh0 = 0.f, sl1 = 0.f, sh1 = 0.f;
#pragma unroll
    for (int m = 0; m < LF / 2; ++m) {
      const int je = LF - 2 - 2 * m, jo = je + 1;
      const float a = a0[m], h = h0[m], v = v0[m], d = d0[m];
      sl0 += rlo[je] * a + rhi[je] * h;
      sh0 += rlo[je] * v + rhi[je] * d;
      sl1 += rlo[jo] * a + rhi[jo] * h;
      sh1 += rlo[jo] * v + rhi[jo] * d;
    }
    Ly[cxl * WT + 2 * u] = sl0;
    Hy[cxl * WT + 2 * u] = sh0;
    Ly[cxl * WT + 2 * u + 1] = sl1;
    Hy[cxl * WT + 2 * u + 1] = sh1;
  }
  __syncthreads();
  // x pass: column pair (2c, 2c + 1) of row iyl (ix0 is even)
#pragma unroll
  for (int q2 = 0; q2 < WT * WT / 2 / WTHREADS; ++q2) {
    const int t = threadIdx.x + q2 * WTHREADS;
    const int iyl = t % WT, c = t / WT;
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int m = 0; m < LF / 2; ++m) {
      const int je = LF - 2 - 2 * m, jo = je + 1;
      const float ly = Ly[(c + m) * WT + iyl], hy = Hy[(c + m) * WT + iyl];
      s0 += rlo[je] * ly + rhi[je] * hy;
      s1 += rlo[jo] * ly + rhi[jo] * hy;
    }
    if (iy0 + iyl < py) {
      if (ix0 + 2 * c < px) acc[2 * q2] += s0;
      if (ix0 + 2 * c + 1 < px) acc[2 * q2 + 1] += s1;
    }
  }
  __syncthreads();
}

template <int LF>
__device__ void syn_tile_t(const SynIn& in, const float* lo, const float* hi, int iy0, int ix0, int py, int px,
                           float* cs, float* Ly, float* Hy, float* acc) {
  float cv[CS_MAX];
  syn_load_t<LF>(in, iy0, ix0, cv);
  syn_store_t<LF>(cv, cs);
  __syncthreads();
  syn_compute_t<LF>(lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc);
}

__device__ void syn_tile(const SynIn& in, int lf, const float* lo, const float* hi, int iy0, int ix0, int py,
                         int px, float* cs, float* Ly, float* Hy, float* acc) {
  switch (lf) {
    case 2: syn_tile_t<2>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    case 4: syn_tile_t<4>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    case 6: syn_tile_t<6>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    case 8: syn_tile_t<8>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    case 10: syn_tile_t<10>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    case 12: syn_tile_t<12>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    case 14: syn_tile_t<14>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
    default: syn_tile_t<16>(in, lo, hi, iy0, ix0, py, px, cs, Ly, Hy, acc); break;
  }
}

// runtime filter length version (the finest level, whose basis loop would
// otherwise hold all eight unrolled instances): accumulates the tile into acc
__device__ void syn_tile_dyn(const SynIn& in, int lf, const float* lo, const float* hi, int iy0, int ix0,
                         int py, int px, float* cs, float* Ly, float* Hy, float* acc) {
  const int cy0 = (iy0 - 1) >= 0 ? (iy0) / 2 : 0;    // ceil((iy0-1)/2) for iy0 >= 0
  const int cx0 = (ix0 - 1) >= 0 ? (ix0) / 2 : 0;
  const int CN = WT / 2 + lf / 2 + 1;
  const int64_t sz = (int64_t)in.ny * in.nx;
  // load 4 subband tiles CN x CN: cs[s][cx][cy] (all loads issued first)
  constexpr int CS_PER_THREAD = (4 * CMAX * CMAX + WTHREADS - 1) / WTHREADS;   // 10
  float cv[CS_PER_THREAD];
  const int ncs = 4 * CN * CN;
#pragma unroll
  for (int i = 0; i < CS_PER_THREAD; ++i) {
    const int t = threadIdx.x + i * WTHREADS;
    cv[i] = 0.f;
    if (t < ncs) {
      const int cyl = t % CN, cxl = (t / CN) % CN, s = t / (CN * CN);
      const int cy = cy0 + cyl, cx = cx0 + cxl;
      if (cy < in.ny && cx < in.nx) {
        const int64_t p = cy + (int64_t)in.ny * cx;
        cv[i] = s == 0 ? in.a[p] : in.h[(s - 1) * sz + p];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < CS_PER_THREAD; ++i) {
    const int t = threadIdx.x + i * WTHREADS;
    if (t < ncs) cs[t] = cv[i];
  }
  __syncthreads();
  // y pass: Ly[cx][iy] = sum_k lo[.] a[k][cx] + hi[.] h[k][cx];  Hy: v, d
  for (int t = threadIdx.x; t < WT * CN; t += WTHREADS) {
    const int iyl = t % WT, cxl = t / WT;
    const int i = iy0 + iyl;
    float sl = 0.f, sh = 0.f;
    // k with j = i - 2k - 2 + lf in [0, lf)
    int kmin = (i - 1 + 1) / 2;            // ceil((i-1)/2) for i >= 0
    int kmax = (i + lf - 2) / 2;
    kmin = max(kmin, cy0);
    kmax = min(kmax, cy0 + CN - 1);
    for (int kk = kmin; kk <= kmax; ++kk) {
      const int j = i - 2 * kk - 2 + lf;
      const int cyl = kk - cy0;
      const float fa = lo[j], fh = hi[j];
      sl += fa * cs[(0 * CN + cxl) * CN + cyl] + fh * cs[(1 * CN + cxl) * CN + cyl];
      sh += fa * cs[(2 * CN + cxl) * CN + cyl] + fh * cs[(3 * CN + cxl) * CN + cyl];
    }
    Ly[cxl * WT + iyl] = sl;
    Hy[cxl * WT + iyl] = sh;
  }
  __syncthreads();
  // x pass
  int q = 0;
  for (int t = threadIdx.x; t < WT * WT; t += WTHREADS, ++q) {
    const int iyl = t % WT, ixl = t / WT;
    const int i = ix0 + ixl;
    float s = 0.f;
    int kmin = (i) / 2;
    int kmax = (i + lf - 2) / 2;
    kmin = max(kmin, cx0);
    kmax = min(kmax, cx0 + CN - 1);
    for (int kk = kmin; kk <= kmax; ++kk) {
      const int j = i - 2 * kk - 2 + lf;
      const int cxl = kk - cx0;
      s += lo[j] * Ly[cxl * WT + iyl] + hi[j] * Hy[cxl * WT + iyl];
    }
    if (iy0 + iyl < py && i < px) acc[q] += s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(WTHREADS) void k_syn_lk(SaraGeom G, int k, const float* __restrict__ v1,
                                                     float* __restrict__ scr) {
  __shared__ float cs[4 * CMAX * CMAX];
  __shared__ float Ly[CMAX * WT];
  __shared__ float Hy[CMAX * WT];
  __shared__ float flo[WMAX_LF], fhi[WMAX_LF];
  const int b = blockIdx.z % G.nb, l = blockIdx.z / G.nb;
  const int iy0 = blockIdx.x * WT, ix0 = blockIdx.y * WT;
  const int py = G.ny[b][k - 1], px = G.nx[b][k - 1];
  if (iy0 >= py || ix0 >= px) return;
  const int lf = G.lf[b];
  if (threadIdx.x < lf) {
    flo[threadIdx.x] = G.lo[b][threadIdx.x];
    fhi[threadIdx.x] = G.hi[b][threadIdx.x];
  }
  const float* v1l = v1 + (int64_t)l * G.s;
  float* sc = scr + (int64_t)l * G.scr_per_chan;
  SynIn in;
  in.a = k == G.nlevel ? v1l + G.boff[b] + G.aoff[b] : sc + G.scr[b][k];
  in.h = v1l + G.boff[b] + G.doff[b][k];
  in.ny = G.ny[b][k];
  in.nx = G.nx[b][k];
  float acc[WT * WT / WTHREADS];
#pragma unroll
  for (int q = 0; q < WT * WT / WTHREADS; ++q) acc[q] = 0.f;
  syn_tile(in, lf, flo, fhi, iy0, ix0, py, px, cs, Ly, Hy, acc);
  float* o = sc + G.scr[b][k - 1];
#pragma unroll
  for (int q = 0; q < WT * WT / WTHREADS; ++q) {
    int iyl, ixl;
    syn_acc_coord(q, iyl, ixl);
    const int iy = iy0 + iyl, ix = ix0 + ixl;
    if (iy < py && ix < px) o[iy + (int64_t)py * ix] = acc[q];
  }
}

// finest level: all bases + Dirac, out = scale * Psi(v1)
__global__ __launch_bounds__(WTHREADS) void k_syn_l1(SaraGeom G, const float* __restrict__ v1,
                                                     const float* __restrict__ scr, float scale,
                                                     float* __restrict__ out) {
  __shared__ float cs[4 * CMAX * CMAX];
  __shared__ float Ly[CMAX * WT];
  __shared__ float Hy[CMAX * WT];
  __shared__ float flo[WMAX_LF], fhi[WMAX_LF];
  const int l = blockIdx.z;
  const int iy0 = blockIdx.x * WT, ix0 = blockIdx.y * WT;
  const float* v1l = v1 + (int64_t)l * G.s;
  const float* sc = scr + (int64_t)l * G.scr_per_chan;
  float acc[WT * WT / WTHREADS];
#pragma unroll
  for (int q = 0; q < WT * WT / WTHREADS; ++q) acc[q] = 0.f;
  for (int b = 0; b < G.nb; ++b) {
    const int lf = G.lf[b];
    if (threadIdx.x < lf) {
      flo[threadIdx.x] = G.lo[b][threadIdx.x];
      fhi[threadIdx.x] = G.hi[b][threadIdx.x];
    }
    SynIn in;
    in.a = G.nlevel == 1 ? v1l + G.boff[b] + G.aoff[b] : sc + G.scr[b][1];
    in.h = v1l + G.boff[b] + G.doff[b][1];
    in.ny = G.ny[b][1];
    in.nx = G.nx[b][1];
    syn_tile_dyn(in, lf, flo, fhi, iy0, ix0, G.Ny, G.Nx, cs, Ly, Hy, acc);
  }
  const float* dir = v1l + G.boff[G.nb];
  float* o = out + (int64_t)l * G.Ny * G.Nx;
  int q = 0;
  for (int t = threadIdx.x; t < WT * WT; t += WTHREADS, ++q) {
    const int iy = iy0 + t % WT, ix = ix0 + t / WT;
    if (iy < G.Ny && ix < G.Nx) {
      const int64_t p = iy + (int64_t)G.Ny * ix;
      float s = acc[q] * G.norm;
      if (G.dirac) s += dir[p] * G.norm;
      o[p] = scale * s;
    }
  }
}

// The standard SARA dictionary (db1..db8 in order: filter length 2(b+1)),
// finest level: the basis loop is unrolled at compile time so every basis
// runs its fixed-length tile code, and the subband loads of basis B+1 are
// issued before the arithmetic of basis B (software pipelining).
__device__ __forceinline__ SynIn syn_l1_in(const SaraGeom& G, const float* v1l, const float* sc, int b) {
  SynIn in;
  in.a = G.nlevel == 1 ? v1l + G.boff[b] + G.aoff[b] : sc + G.scr[b][1];
  in.h = v1l + G.boff[b] + G.doff[b][1];
  in.ny = G.ny[b][1];
  in.nx = G.nx[b][1];
  return in;
}

template <int B>
__device__ __forceinline__ void syn_l1_std_bases(const SaraGeom& G, const float* v1l, const float* sc, int iy0,
                                                 int ix0, float* flo, float* fhi, float* cs, float* Ly, float* Hy,
                                                 float* cv, float* acc) {
  constexpr int LF = 2 * (B + 1);
  if (threadIdx.x < LF) {
    flo[threadIdx.x] = G.lo[B][threadIdx.x];
    fhi[threadIdx.x] = G.hi[B][threadIdx.x];
  }
  syn_store_t<LF>(cv, cs);   // basis B's subbands, loaded one step ahead
  __syncthreads();
  if constexpr (B + 1 < WMAX_BASES) syn_load_t<2 * (B + 2)>(syn_l1_in(G, v1l, sc, B + 1), iy0, ix0, cv);
  syn_compute_t<LF>(flo, fhi, iy0, ix0, G.Ny, G.Nx, cs, Ly, Hy, acc);
  if constexpr (B + 1 < WMAX_BASES)
    syn_l1_std_bases<B + 1>(G, v1l, sc, iy0, ix0, flo, fhi, cs, Ly, Hy, cv, acc);
}

__global__ __launch_bounds__(WTHREADS) void k_syn_l1_std(SaraGeom G, const float* __restrict__ v1,
                                                         const float* __restrict__ scr, float scale,
                                                         float* __restrict__ out) {
  __shared__ float cs[4 * CMAX * CMAX];
  __shared__ float Ly[CMAX * WT];
  __shared__ float Hy[CMAX * WT];
  __shared__ float flo[WMAX_LF], fhi[WMAX_LF];
  const int l = blockIdx.z;
  const int iy0 = blockIdx.x * WT, ix0 = blockIdx.y * WT;
  const float* v1l = v1 + (int64_t)l * G.s;
  const float* sc = scr + (int64_t)l * G.scr_per_chan;
  float acc[WT * WT / WTHREADS] = {0.f, 0.f, 0.f, 0.f};
  float cv[CS_MAX];
  syn_load_t<2>(syn_l1_in(G, v1l, sc, 0), iy0, ix0, cv);
  syn_l1_std_bases<0>(G, v1l, sc, iy0, ix0, flo, fhi, cs, Ly, Hy, cv, acc);
  const float* dir = v1l + G.boff[G.nb];
  float* o = out + (int64_t)l * G.Ny * G.Nx;
#pragma unroll
  for (int q = 0; q < WT * WT / WTHREADS; ++q) {
    int iyl, ixl;
    syn_acc_coord(q, iyl, ixl);
    const int iy = iy0 + iyl, ix = ix0 + ixl;
    if (iy < G.Ny && ix < G.Nx) {
      const int64_t p = iy + (int64_t)G.Ny * ix;
      float s = acc[q] * G.norm;
      if (G.dirac) s += dir[p] * G.norm;
      o[p] = scale * s;
    }
  }
}

// ---- l21 dual step, Dirac analysis fused ------------------------------------
// r_l = v1_l (+ xhat_l(p)/sqrt(P) on the Dirac block); l2 = ||r||_2 over l;
// v1 <- r - max(l2 - beta w, 0)/(l2 + eps) r.   mode 1: only sum l2 (prior).
__global__ __launch_bounds__(256) void k_l21(SaraGeom G, float* __restrict__ v1,
                                             const float* __restrict__ xhat, int L,
                                             const float* __restrict__ w, double beta, int mode,
                                             double* __restrict__ partial) {
  __shared__ double sh[256 / 64];
  double tot[1] = {0.0};
  const int64_t s = G.s;
  const int64_t dstart = G.boff[G.nb];
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < s;
       c += (int64_t)gridDim.x * blockDim.x) {
    const bool dir = G.dirac && c >= dstart;
    constexpr int LREG = 16;   // channels kept in registers between the norm and the update
    float rv[LREG];
    double ss = 0.0;
#pragma unroll
    for (int l = 0; l < LREG; ++l) {
      if (l < L) {
        float r = v1[(int64_t)l * s + c];
        if (dir && xhat) r += xhat[(int64_t)l * G.Ny * G.Nx + (c - dstart)] * G.norm;
        rv[l] = r;
        ss += (double)r * r;
      }
    }
    for (int l = LREG; l < L; ++l) {
      float r = v1[(int64_t)l * s + c];
      if (dir && xhat) r += xhat[(int64_t)l * G.Ny * G.Nx + (c - dstart)] * G.norm;
      ss += (double)r * r;
    }
    const double l2 = sqrt(ss);
    if (mode == 1) {
      tot[0] += l2;
      continue;
    }
    if (mode == 2) {   // reweighting: w = upsilon / (upsilon + l2), beta = upsilon
      partial[c] = beta / (beta + l2);
      continue;
    }
    const double t = beta * (w ? (double)w[c] : 1.0);
    const float f = (float)(fmax(l2 - t, 0.0) / (l2 + 2.220446049250313e-16));
#pragma unroll
    for (int l = 0; l < LREG; ++l)
      if (l < L) v1[(int64_t)l * s + c] = rv[l] - f * rv[l];
    for (int l = LREG; l < L; ++l) {
      float r = v1[(int64_t)l * s + c];
      if (dir && xhat) r += xhat[(int64_t)l * G.Ny * G.Nx + (c - dstart)] * G.norm;
      v1[(int64_t)l * s + c] = r - f * r;
    }
  }
  if (mode == 1) {
    block_sum<1>(tot, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot[0];
  }
}

// The same dual step (mode 0) four consecutive coefficients per thread with
// 16-byte loads and stores: every channel's vector load of a step is issued
// before the norms need them.  For s, the Dirac offset and Ny*Nx multiples of
// 4 (a group of four is then wholly Dirac or wholly wavelet).
__global__ __launch_bounds__(256) void k_l21_v4(SaraGeom G, float* __restrict__ v1,
                                                const float* __restrict__ xhat, int L,
                                                const float* __restrict__ w, double beta) {
  constexpr int LREG = 16;
  const int64_t s4 = G.s / 4;
  const int64_t dstart = G.boff[G.nb];
  const int64_t npix = (int64_t)G.Ny * G.Nx;
  float4* __restrict__ v = reinterpret_cast<float4*>(v1);
  for (int64_t c4 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c4 < s4;
       c4 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = 4 * c4;
    const bool dir = G.dirac && c >= dstart && xhat;
    float4 rv[LREG];
    double ss[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int l = 0; l < LREG; ++l) {
      if (l < L) {
        float4 r = v[(int64_t)l * s4 + c4];
        if (dir) {
          const float4 xv = *reinterpret_cast<const float4*>(xhat + (int64_t)l * npix + (c - dstart));
          r.x += xv.x * G.norm;
          r.y += xv.y * G.norm;
          r.z += xv.z * G.norm;
          r.w += xv.w * G.norm;
        }
        rv[l] = r;
        ss[0] += (double)r.x * r.x;
        ss[1] += (double)r.y * r.y;
        ss[2] += (double)r.z * r.z;
        ss[3] += (double)r.w * r.w;
      }
    }
    float4 wv = make_float4(1.f, 1.f, 1.f, 1.f);
    if (w) wv = *reinterpret_cast<const float4*>(w + c);
    const float wa[4] = {wv.x, wv.y, wv.z, wv.w};
    float f[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double l2 = sqrt(ss[j]);
      f[j] = (float)(fmax(l2 - beta * (double)wa[j], 0.0) / (l2 + 2.220446049250313e-16));
    }
#pragma unroll
    for (int l = 0; l < LREG; ++l)
      if (l < L) {
        const float4 r = rv[l];
        v[(int64_t)l * s4 + c4] = make_float4(r.x - f[0] * r.x, r.y - f[1] * r.y, r.z - f[2] * r.z, r.w - f[3] * r.w);
      }
  }
}

__global__ void k_f64_to_f32(const double* __restrict__ a, int64_t n, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_sum_l(const double* __restrict__ partial, int n, double* out) {
  __shared__ double sh[4];
  double s[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) s[0] += partial[i];
  block_sum<1>(s, sh);
  if (threadIdx.x == 0) *out = s[0];
}

// v1 := Psit(x) on the Dirac block when no l21 step follows (plain analysis)
__global__ void k_dirac_copy(SaraGeom G, const float* __restrict__ x, int L, float* __restrict__ v1,
                             int accumulate) {
  const int64_t N = (int64_t)G.Ny * G.Nx;
  const int64_t dstart = G.boff[G.nb];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * L;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(i / N);
    const int64_t p = i - (int64_t)l * N;
    float* d = v1 + (int64_t)l * G.s + dstart + p;
    const float v = x[i] * G.norm;
    *d = accumulate ? *d + v : v;
  }
}

}  // namespace

// coefficients (+)= Psit(x) for the wavelet bases (Dirac block untouched)
void sara_analysis_wavelets(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, float* v1) {
  const SaraGeom& G = S->g;
  float* scr = (float*)ctx->ws.get("sara_scratch", std::max<int64_t>(1, G.scr_per_chan * L) * sizeof(float));
  int my = 0, mx = 0;
  for (int b = 0; b < G.nb; ++b) {
    my = std::max(my, G.ny[b][1]);
    mx = std::max(mx, G.nx[b][1]);
  }
  Scope sc(ctx, "wavelet_analysis");
  (void)my;
  (void)mx;
  for (int k = 1; k <= G.nlevel; ++k) {
    int ny = 0, nx = 0;
    for (int b = 0; b < G.nb; ++b) {
      ny = std::max(ny, G.ny[b][k]);
      nx = std::max(nx, G.nx[b][k]);
    }
    k_an_lk<<<dim3(cdiv(ny, AY), cdiv(nx, AX), G.nb * L), WTHREADS, 0, ctx->stream>>>(G, k, x, v1, scr);
    check_launch();
  }
}

// out = scale * Psi(v1) (wavelets + Dirac)
void sara_synthesis(fhs_ctx* ctx, const fhs_sara* S, const float* v1, int L, float scale, float* out) {
  const SaraGeom& G = S->g;
  float* scr = (float*)ctx->ws.get("sara_scratch", std::max<int64_t>(1, G.scr_per_chan * L) * sizeof(float));
  Scope sc(ctx, "wavelet_synthesis");
  for (int k = G.nlevel; k >= 2 && G.nb > 0; --k) {
    int py = 0, px = 0;
    for (int b = 0; b < G.nb; ++b) {
      py = std::max(py, G.ny[b][k - 1]);
      px = std::max(px, G.nx[b][k - 1]);
    }
    k_syn_lk<<<dim3(cdiv(py, WT), cdiv(px, WT), G.nb * L), WTHREADS, 0, ctx->stream>>>(G, k, v1, scr);
    check_launch();
  }
  bool std_dict = G.nb == WMAX_BASES;
  for (int b = 0; b < G.nb && std_dict; ++b) std_dict = G.lf[b] == 2 * (b + 1);
  if (std_dict)
    k_syn_l1_std<<<dim3(cdiv(G.Ny, WT), cdiv(G.Nx, WT), L), WTHREADS, 0, ctx->stream>>>(G, v1, scr, scale, out);
  else
    k_syn_l1<<<dim3(cdiv(G.Ny, WT), cdiv(G.Nx, WT), L), WTHREADS, 0, ctx->stream>>>(G, v1, scr, scale, out);
  check_launch();
}

// v1 <- l21 dual step of (v1 + Psit(xhat)); g1 = sigma1 * Psi(v1)
void sara_dual_step(fhs_ctx* ctx, const fhs_sara* S, float* v1, const float* xhat, int L,
                    const float* w, double beta, float sigma1, float* g1) {
  const SaraGeom& G = S->g;
  if (G.nb) sara_analysis_wavelets(ctx, S, xhat, L, v1);
  {
    Scope sc(ctx, "l21_threshold");
    const bool v4 = L <= 16 && G.s % 4 == 0 && G.boff[G.nb] % 4 == 0 && ((int64_t)G.Ny * G.Nx) % 4 == 0 &&
                    (((uintptr_t)v1 | (uintptr_t)xhat | (uintptr_t)w) % 16) == 0;
    if (v4)
      k_l21_v4<<<std::min<int64_t>(cdiv(G.s / 4, 256), 2048), 256, 0, ctx->stream>>>(G, v1, xhat, L, w, beta);
    else
      k_l21<<<cdiv(G.s, 256) < 2048 ? cdiv(G.s, 256) : 2048, 256, 0, ctx->stream>>>(G, v1, xhat, L, w, beta, 0,
                                                                                     nullptr);
    check_launch();
  }
  sara_synthesis(ctx, S, v1, L, sigma1, g1);
}

int64_t sara_ncoefs(const fhs_sara* S) { return S->g.s; }

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_sara_create_host(fhs_ctx* ctx, int Ny, int Nx, int nlevel, int nbases,
                         const int* filter_len, const double* lo_r, int with_dirac, fhs_sara** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(Ny >= 1 && Nx >= 1, "sara_create: bad image size");
    FHS_CHECK(nlevel >= 1 && nlevel <= WMAX_LEVELS, "sara_create: nlevel must be in [1, 6]");
    FHS_CHECK(nbases >= 0 && nbases <= WMAX_BASES, "sara_create: at most 8 wavelet bases");
    FHS_CHECK(nbases + (with_dirac ? 1 : 0) >= 1, "sara_create: empty dictionary");
    fhs_sara* S = new fhs_sara();
    S->device = ctx->device;
    SaraGeom& G = S->g;
    G.Ny = Ny;
    G.Nx = Nx;
    G.nlevel = nlevel;
    G.nb = nbases;
    G.dirac = with_dirac ? 1 : 0;
    G.norm = (float)(1.0 / std::sqrt((double)(nbases + G.dirac)));
    int64_t off = 0, soff = 0;
    const double* f = lo_r;
    for (int b = 0; b < nbases; ++b) {
      const int lf = filter_len[b];
      FHS_CHECK(lf >= 2 && lf <= WMAX_LF && lf % 2 == 0, "sara_create: filter length must be even in [2, 16]");
      G.lf[b] = lf;
      for (int j = 0; j < lf; ++j) {
        G.lo[b][j] = (float)f[j];
        G.hi[b][j] = (float)((j % 2 ? -1.0 : 1.0) * f[lf - 1 - j]);
      }
      f += lf;
      G.ny[b][0] = Ny;
      G.nx[b][0] = Nx;
      for (int k = 1; k <= nlevel; ++k) {
        G.ny[b][k] = (G.ny[b][k - 1] + lf - 1) / 2;
        G.nx[b][k] = (G.nx[b][k - 1] + lf - 1) / 2;
      }
      // scratch: approx levels 1..nlevel-1 (analysis), also reused by synthesis
      for (int k = 1; k < nlevel; ++k) {
        G.scr[b][k] = soff;
        soff += (int64_t)G.ny[b][k] * G.nx[b][k];
      }
      G.boff[b] = off;
      G.aoff[b] = 0;
      int64_t p = (int64_t)G.ny[b][nlevel] * G.nx[b][nlevel];
      for (int k = nlevel; k >= 1; --k) {
        G.doff[b][k] = p;
        p += 3 * (int64_t)G.ny[b][k] * G.nx[b][k];
      }
      off += p;
    }
    G.boff[nbases] = off;
    G.s = off + (G.dirac ? (int64_t)Ny * Nx : 0);
    G.scr_per_chan = soff;
    *out = S;
  });
}

int fhs_sara_destroy(fhs_sara* S) {
  delete S;
  return FHS_OK;
}

int fhs_sara_info(const fhs_sara* S, int64_t* n_coefs) {
  if (!S) return FHS_ERR_ARG;
  if (n_coefs) *n_coefs = S->g.s;
  return FHS_OK;
}

int fhs_sara_analysis(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, float* coefs,
                      int accumulate) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(S && x && coefs && L >= 1 && L * std::max(1, S->g.nb) <= 65535, "sara_analysis: bad argument");
    if (!accumulate)
      FHS_HIP(hipMemsetAsync(coefs, 0, (size_t)S->g.s * L * sizeof(float), ctx->stream));
    if (S->g.nb) sara_analysis_wavelets(ctx, S, x, L, coefs);
    if (S->g.dirac) {
      k_dirac_copy<<<1024, 256, 0, ctx->stream>>>(S->g, x, L, coefs, 1);
      check_launch();
    }
  });
}

int fhs_sara_synthesis(fhs_ctx* ctx, const fhs_sara* S, const float* coefs, int L, float scale,
                       float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(S && coefs && out && L >= 1 && L * std::max(1, S->g.nb) <= 65535, "sara_synthesis: bad argument");
    sara_synthesis(ctx, S, coefs, L, scale, out);
  });
}

int fhs_update_dual_sparsity(fhs_ctx* ctx, const fhs_sara* S, float* v1, const float* xhat, int L,
                             const float* weights1, double beta1, float sigma1, float* g1) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(S && v1 && xhat && g1 && L >= 1 && L * std::max(1, S->g.nb) <= 65535, "update_dual_sparsity: bad argument");
    sara_dual_step(ctx, S, v1, xhat, L, weights1, beta1, sigma1, g1);
  });
}

int fhs_sara_update_weights(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, double upsilon,
                            float* weights1) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(S && x && weights1 && L >= 1 && L * std::max(1, S->g.nb) <= 65535,
              "sara_update_weights: bad argument");
    const SaraGeom& G = S->g;
    float* c = (float*)ctx->ws.get("sara_prior_coefs", (size_t)G.s * L * sizeof(float));
    FHS_HIP(hipMemsetAsync(c, 0, (size_t)G.s * L * sizeof(float), ctx->stream));
    if (G.nb) sara_analysis_wavelets(ctx, S, x, L, c);
    double* w = (double*)ctx->ws.get("sara_weights_f64", (size_t)G.s * sizeof(double));
    const unsigned nbk = cdiv(G.s, 256) < 2048 ? cdiv(G.s, 256) : 2048;
    k_l21<<<nbk, 256, 0, ctx->stream>>>(G, c, G.dirac ? x : nullptr, L, nullptr, upsilon, 2, w);
    check_launch();
    k_f64_to_f32<<<nbk, 256, 0, ctx->stream>>>(w, G.s, weights1);
    check_launch();
  });
}

int fhs_compute_sara_prior(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, double* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(S && x && out && L >= 1 && L * std::max(1, S->g.nb) <= 65535, "compute_sara_prior: bad argument");
    const SaraGeom& G = S->g;
    float* c = (float*)ctx->ws.get("sara_prior_coefs", (size_t)G.s * L * sizeof(float));
    FHS_HIP(hipMemsetAsync(c, 0, (size_t)G.s * L * sizeof(float), ctx->stream));
    if (G.nb) sara_analysis_wavelets(ctx, S, x, L, c);
    const unsigned nbk = cdiv(G.s, 256) < 2048 ? cdiv(G.s, 256) : 2048;
    double* part = (double*)ctx->ws.get("sara_prior_partial", (nbk + 1) * sizeof(double));
    k_l21<<<nbk, 256, 0, ctx->stream>>>(G, c, G.dirac ? x : nullptr, L, nullptr, 0.0, 1, part);
    check_launch();
    k_sum_l<<<1, 256, 0, ctx->stream>>>(part, nbk, part + nbk);
    check_launch();
    FHS_HIP(hipMemcpyAsync(out, part + nbk, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
