// Device-resident interpolation matrix G of one channel.
//
// Forward (degridding, y = G Fx) reads G transposed: value j of row m sits at
// vals[j*M + m], so the 64 lanes of a wavefront (64 consecutive rows) load
// 64 consecutive complex values per tap (one 512-byte coalesced load).
// Two row formats share that value layout:
//   ELL    : explicit column per tap (cols[j*M + m]), any sparsity pattern;
//   WINDOW : one anchor per row; tap j = jx*Ky + jy reads grid point
//            ((ky0 + jy) mod Kty, (kx0 + jx) mod Ktx) - the dense min-max
//            window, no column indices (8 B per nonzero instead of 12).
//   SEP    : the min-max interpolator of a 2-D NUFFT is the Kronecker
//            product of two 1-D interpolators, G[m, (jy, jx)] =
//            cy[m][jy] cx[m][jx]: a row stores Ky + Kx complex values
//            (112 B for 7 x 7 instead of 392 B) plus its anchor.  The CSC of
//            the adjoint then stores only (row << 6 | tap) per nonzero (4 B).
// Adjoint (gridding, Fx = G' y) is computed per grid point from a CSC copy
// (colptr / rows / conj-ready values): each output is owned by one thread,
// no atomics, deterministic summation order.
#pragma once

#include "common.hpp"

namespace fhs {

enum GLayout { G_ELL = 0, G_WINDOW = 1, G_SEP = 2 };
constexpr int SEP_KMAX = 8;   // max taps per axis of the separable layout

struct GDev {
  int layout;
  int M, K;
  int Kty, Ktx, Ky, Kx;
  int ncols;
  const int* cols;      // ELL: [K][M]
  const int* anchor;    // WINDOW: [M]
  const float2* vals;   // [K][M]
  const int* colptr;    // [ncols + 1]
  const int* csc_rows;  // [nnz]
  const float2* csc_vals;
  const int* ucol;      // adjoint work units: unit u = columns [ucol[u], ucol[u+1])
  int nunits;
  const unsigned* csc_pk;  // SEP: (row << 6) | (jx*Ky + jy) per nonzero
};

// value and row of CSC entry e
__device__ __forceinline__ void g_csc_entry(const GDev& g, int e, float2& v, int& r) {
  if (g.layout == G_SEP) {
    const unsigned pk = g.csc_pk[e];
    r = (int)(pk >> 6);
    const int tap = (int)(pk & 63u), jx = tap / g.Ky, jy = tap - jx * g.Ky;
    v = cmul(g.vals[(size_t)jy * g.M + r], g.vals[(size_t)(g.Ky + jx) * g.M + r]);
  } else {
    v = g.csc_vals[e];
    r = g.csc_rows[e];
  }
}

// Adjoint work units: contiguous column ranges holding about ADJ_ENTRIES
// nonzeros (and at most ADJ_COLS columns), so every workgroup streams the
// same amount of CSC data whatever the uv density.
constexpr int ADJ_THREADS = 256;
constexpr int ADJ_ENTRIES = 2048;
constexpr int ADJ_COLS = 2048;
constexpr int ADJ_LDS = 4096;

// sum over taps j of row m: G[m, c_j] * f(c_j)
template <class F>
__device__ __forceinline__ float2 g_row_apply(const GDev& g, int m, F&& f) {
  float2 acc = make_float2(0.f, 0.f);
  if (g.layout == G_ELL) {
    for (int j = 0; j < g.K; ++j) {
      const int c = g.cols[(size_t)j * g.M + m];
      const float2 v = g.vals[(size_t)j * g.M + m];
      const float2 x = f(c);
      acc.x += v.x * x.x - v.y * x.y;
      acc.y += v.x * x.y + v.y * x.x;
    }
  } else if (g.layout == G_SEP) {
    const int a = g.anchor[m];
    const int ky0 = a % g.Kty, kx0 = a / g.Kty;
    float2 cy[SEP_KMAX];
#pragma unroll
    for (int jy = 0; jy < SEP_KMAX; ++jy)
      cy[jy] = jy < g.Ky ? g.vals[(size_t)jy * g.M + m] : make_float2(0.f, 0.f);
    for (int jx = 0; jx < g.Kx; ++jx) {
      int kx = kx0 + jx;
      if (kx >= g.Ktx) kx -= g.Ktx;
      const int colbase = kx * g.Kty;
      float2 s = make_float2(0.f, 0.f);
#pragma unroll
      for (int jy = 0; jy < SEP_KMAX; ++jy) {
        if (jy < g.Ky) {
          int ky = ky0 + jy;
          if (ky >= g.Kty) ky -= g.Kty;
          const float2 x = f(colbase + ky);
          s.x += cy[jy].x * x.x - cy[jy].y * x.y;
          s.y += cy[jy].x * x.y + cy[jy].y * x.x;
        }
      }
      const float2 c = g.vals[(size_t)(g.Ky + jx) * g.M + m];
      acc.x += c.x * s.x - c.y * s.y;
      acc.y += c.x * s.y + c.y * s.x;
    }
  } else {
    const int a = g.anchor[m];
    const int ky0 = a % g.Kty, kx0 = a / g.Kty;
    int j = 0;
    for (int jx = 0; jx < g.Kx; ++jx) {
      int kx = kx0 + jx;
      if (kx >= g.Ktx) kx -= g.Ktx;
      const int colbase = kx * g.Kty;
      for (int jy = 0; jy < g.Ky; ++jy, ++j) {
        int ky = ky0 + jy;
        if (ky >= g.Kty) ky -= g.Kty;
        const float2 v = g.vals[(size_t)j * g.M + m];
        const float2 x = f(colbase + ky);
        acc.x += v.x * x.x - v.y * x.y;
        acc.y += v.x * x.y + v.y * x.x;
      }
    }
  }
  return acc;
}

__device__ __forceinline__ float2 g_row_dot(const GDev& g, int m, const float2* __restrict__ Fx) {
  return g_row_apply(g, m, [&](int c) { return Fx[c]; });
}

// sum over rows r of column k: conj(G[r,k]) * w(r)
template <class W>
__device__ __forceinline__ float2 g_col_dot(const GDev& g, int k, W&& w) {
  float2 acc = make_float2(0.f, 0.f);
  const int e0 = g.colptr[k], e1 = g.colptr[k + 1];
  for (int e = e0; e < e1; ++e) {
    float2 v;
    int r;
    g_csc_entry(g, e, v, r);
    const float2 y = w(r);
    acc.x += v.x * y.x + v.y * y.y;
    acc.y += v.x * y.y - v.y * y.x;
  }
  return acc;
}

// One adjoint work unit: the unit's CSC entries are loaded coalesced (lane
// t takes entries t, t+256, ...), their products conj(G) w(row) staged in
// LDS, then each thread sums whole columns from LDS and writes them: every
// output column belongs to exactly one unit (no atomics).  extra(c) is added
// to column c (the G w term of the Hermitian DR form).
template <class W, class X>
__device__ __forceinline__ void g_adjoint_unit(const GDev& g, int u, W&& w, X&& extra,
                                               float2* __restrict__ out, float2* prod) {
  const int c0 = g.ucol[u], c1 = g.ucol[u + 1];
  const int e0 = g.colptr[c0], n = g.colptr[c1] - e0;
  if (n <= ADJ_LDS) {
    for (int e = threadIdx.x; e < n; e += ADJ_THREADS) {
      float2 v;
      int r;
      g_csc_entry(g, e0 + e, v, r);
      const float2 y = w(r);
      prod[e] = make_float2(v.x * y.x + v.y * y.y, v.x * y.y - v.y * y.x);
    }
    __syncthreads();
    for (int c = c0 + threadIdx.x; c < c1; c += ADJ_THREADS) {
      const int a = g.colptr[c] - e0, b = g.colptr[c + 1] - e0;
      float2 s = make_float2(0.f, 0.f);
      for (int e = a; e < b; ++e) s = cadd(s, prod[e]);
      out[c] = cadd(s, extra(c));
    }
  } else {   // a column longer than the LDS stage: direct gather
    for (int c = c0 + threadIdx.x; c < c1; c += ADJ_THREADS) out[c] = cadd(g_col_dot(g, c, w), extra(c));
  }
}

}  // namespace fhs

struct fhs_gmat {
  int device = 0;
  fhs::GDev d{};
  int64_t M = 0, ncols = 0, nnz = 0;
  void* mem[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};
