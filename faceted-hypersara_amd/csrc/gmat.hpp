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
//            cy[m][jy] cx[m][jx]: a row stores Ky + Kx complex values plus
//            its anchor, row-major in one 128-byte line (cy in slots 0-7, cx
//            in slots 8-15, zero padded) instead of 392 B for 7 x 7 taps, so
//            any gather order costs one line per row.  The CSC of the adjoint
//            (small grids) stores only (row << 6 | tap) per nonzero (4 B).
// Adjoint (gridding, Fx = G' y) is computed per grid point, each output
// owned by one thread (no atomics, deterministic summation order): from a
// CSC copy (colptr / rows / values) in general, by tile gridding for the
// separable layout on grids of whole 32 x 32 tiles.
#pragma once

#include <hip/hip_fp16.h>

#include "common.hpp"

namespace fhs {

enum GLayout { G_ELL = 0, G_WINDOW = 1, G_SEP = 2 };
constexpr int SEP_KMAX = 16;  // max taps per axis of the separable layout
// A separable row holds 2 kmax complex slots (cy in 0..kmax-1, cx in
// kmax..2 kmax-1, zero padded), kmax = 8 (128-byte rows, windows of up to
// 8 taps: the min-max interpolator) or 16 (256-byte rows: w-projected rows).
inline int sep_kmax_for(int Ky, int Kx) { return (Ky <= 8 && Kx <= 8) ? 8 : 16; }

struct GDev {
  int layout;
  int M, K;
  int Kty, Ktx, Ky, Kx;
  int ncols;
  const int* cols;      // ELL: [K][M]
  const int* anchor;    // WINDOW: [M]
  const float2* vals;   // ELL / WINDOW: [K][M]; SEP: [M][2 kmax] (__half2 slots if half)
  int kmax;             // SEP: complex slots per axis in a row (8 or 16)
  int half;             // SEP on a tiled grid: rows stored as FP16 (re, im) pairs, 4 B a slot
  const int* colptr;    // [ncols + 1]
  const int* csc_rows;  // [nnz]
  const float2* csc_vals;
  const int* ucol;      // adjoint work units: unit u = columns [ucol[u], ucol[u+1])
  int nunits;
  const unsigned* csc_pk;  // SEP: (row << 6) | (jx*Ky + jy) per nonzero
  // SEP on a tiled grid: adjoint by tile gridding (no CSC; nunits = items)
  int tiled, tiles_y, tiles_x;
  const int* tile_rows;    // rows listed under every tile their window touches
  const int* tile_cell;    // local anchor cell of each listed row
  // work items {tile, begin, end, slot}, largest first: slot -1 writes the
  // tile, slot >= 0 writes partial[slot] (a tile split over several items,
  // summed in slot order by the reduce pass), slot -2 zeroes the empty tiles
  // empty_tiles[begin, end)
  const int4* items;
  const int* empty_tiles;
  const int4* red;         // {tile, first slot, slot count} of the split tiles
  int nred;
  float2* partial;         // [slots][TILE_Y * TILE_X]
  // forward by tile: rows grouped by the tile of their anchor (ascending row
  // order inside a tile), work items {tile, begin, end, 0} of <= FWD_ITEM rows
  const int* prim_rows;
  const int4* fitems;
  int nfitems;
  // holographic form (square, rows = columns = grid points): the full
  // Hermitian stencil G'G of each active mode i, stencil[o * nact + i] for the
  // (2Ky-1) x (2Kx-1) neighbourhood offsets o = (dy + Ky-1) + (2Ky-1)(dx + Kx-1),
  // and act[p] = i for the active grid points (-1 elsewhere)
  const float2* stencil;
  const int* act;
  int nact;
};

// one FP16 slot (re, im), as stored in a 32-bit word, in FP32
__device__ __forceinline__ float2 h2f(unsigned u) {
  __half2 h;
  __builtin_memcpy(&h, &u, 4);
  return __half22float2(h);
}

// the 2 KM complex slots of separable row m (KM 16-byte loads, KM / 2 for
// FP16 rows); KM = g.kmax
template <int KM>
__device__ __forceinline__ void sep_load(const GDev& g, int m, float2* cy, float2* cx) {
  if (g.half) {
    const uint4* q = reinterpret_cast<const uint4*>(reinterpret_cast<const __half2*>(g.vals) + (size_t)m * (2 * KM));
#pragma unroll
    for (int i = 0; i < KM / 4; ++i) {
      const uint4 a = q[i], b = q[KM / 4 + i];
      cy[4 * i] = h2f(a.x);
      cy[4 * i + 1] = h2f(a.y);
      cy[4 * i + 2] = h2f(a.z);
      cy[4 * i + 3] = h2f(a.w);
      cx[4 * i] = h2f(b.x);
      cx[4 * i + 1] = h2f(b.y);
      cx[4 * i + 2] = h2f(b.z);
      cx[4 * i + 3] = h2f(b.w);
    }
    return;
  }
  const float4* p = reinterpret_cast<const float4*>(g.vals + (size_t)m * (2 * KM));
#pragma unroll
  for (int i = 0; i < KM / 2; ++i) {
    const float4 a = p[i], b = p[KM / 2 + i];
    cy[2 * i] = make_float2(a.x, a.y);
    cy[2 * i + 1] = make_float2(a.z, a.w);
    cx[2 * i] = make_float2(b.x, b.y);
    cx[2 * i + 1] = make_float2(b.z, b.w);
  }
}

// value and row of CSC entry e
__device__ __forceinline__ void g_csc_entry(const GDev& g, int e, float2& v, int& r) {
  if (g.layout == G_SEP) {
    const unsigned pk = g.csc_pk[e];
    r = (int)(pk >> 6);
    const int tap = (int)(pk & 63u), jx = tap / g.Ky, jy = tap - jx * g.Ky;
    v = cmul(g.vals[(size_t)r * 2 * g.kmax + jy], g.vals[(size_t)r * 2 * g.kmax + g.kmax + jx]);
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
  } else if (g.layout == G_SEP && g.kmax > 8) {
    const int a = g.anchor[m];
    const int ky0 = a % g.Kty, kx0 = a / g.Kty;
    const float2* row = g.vals + (size_t)m * 2 * g.kmax;
    for (int jx = 0; jx < g.Kx; ++jx) {
      int kx = kx0 + jx;
      if (kx >= g.Ktx) kx -= g.Ktx;
      const int colbase = kx * g.Kty;
      float2 s = make_float2(0.f, 0.f);
      for (int jy = 0; jy < g.Ky; ++jy) {
        int ky = ky0 + jy;
        if (ky >= g.Kty) ky -= g.Kty;
        const float2 x = f(colbase + ky), c = row[jy];
        s.x += c.x * x.x - c.y * x.y;
        s.y += c.x * x.y + c.y * x.x;
      }
      const float2 c = row[g.kmax + jx];
      acc.x += c.x * s.x - c.y * s.y;
      acc.y += c.x * s.y + c.y * s.x;
    }
  } else if (g.layout == G_SEP) {
    const int a = g.anchor[m];
    const int ky0 = a % g.Kty, kx0 = a / g.Kty;
    float2 cy[8], cx[8];
    sep_load<8>(g, m, cy, cx);
#pragma unroll
    for (int jx = 0; jx < 8; ++jx) {   // fully unrolled: cy / cx stay in registers
      if (jx < g.Kx) {
        int kx = kx0 + jx;
        if (kx >= g.Ktx) kx -= g.Ktx;
        const int colbase = kx * g.Kty;
        float2 s = make_float2(0.f, 0.f);
#pragma unroll
        for (int jy = 0; jy < 8; ++jy) {
          if (jy < g.Ky) {
            int ky = ky0 + jy;
            if (ky >= g.Kty) ky -= g.Kty;
            const float2 x = f(colbase + ky);
            s.x += cy[jy].x * x.x - cy[jy].y * x.y;
            s.y += cy[jy].x * x.y + cy[jy].y * x.x;
          }
        }
        const float2 c = cx[jx];
        acc.x += c.x * s.x - c.y * s.y;
        acc.y += c.x * s.y + c.y * s.x;
      }
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

// Tile-gridding adjoint of the separable layout.
//
// The grid is cut into TILE_Y x TILE_X tiles.  At creation every row is
// listed under each tile its Ky x Kx window touches (1, 2 or 4 tiles), with
// its anchor cell relative to the tile ("local cell": (TILE_Y + Ky - 1) x
// (TILE_X + Kx - 1) cells, x-major, the list sorted by it).  One workgroup
// owns one tile and stages its listed rows in LDS a chunk at a time.  Two
// forms, chosen by the row width:
//  - 8-slot rows (7 x 7 windows): a cell-indexed gather.  The chunk is
//    staged as Q[k][jx] = conj(cx[jx]) w(row), C[k][jy] = conj(cy[jy]);
//    where each local cell starts in the chunk is found (the first row of
//    each occupied cell marks it, a block-wide suffix minimum fills the
//    rest), and every thread sums, for the grid points it owns, the rows of
//    the Ky consecutive cells of each of the Kx cell columns that reach it.
//  - 16-slot rows (the w-projected 15 x 15 windows): on the matrix cores.
//    Restricted to one tile the adjoint is a sum of rank-1 updates,
//      out[y][x] = sum_rows conj(cy[y - dy]) * (w(row) conj(cx[x - dx])),
//    (dy, dx) the row's window origin relative to the tile: a GEMM whose
//    inner dimension runs over the tile's rows.  Each of the 4 waves owns
//    one 16 x 16 quadrant, accumulated by v_mfma_f32_16x16x4f32 in FP32, two
//    complex rows per instruction pair (K = 4 real terms, re/im split over
//    the A/B lanes), over the rows whose window meets its quadrant.  Rows
//    are staged 32 at a time by all 256 threads (8 per row, coalesced), the
//    next chunk loaded during this chunk's updates; 32-row chunks keep the
//    LDS at 17 KB, so eight workgroups fit on a CU (the register limit).  A
//    7 x 7 window fills under a fifth of such an update, which is why the
//    narrow rows keep the gather.
// Every output point is written once, by one thread, after a fixed-order
// sum: no atomics, deterministic.  The uv density of a tile varies by
// orders of magnitude (empty corners, a dense centre), so the work is cut
// into items of at most TILE_ITEM listed rows: a dense tile is spread over
// several workgroups whose partial tiles a second pass adds in a fixed
// order, and empty tiles are zeroed in groups.
constexpr int TILE_Y = 32, TILE_X = 32;
constexpr int TILE_PTS = TILE_Y * TILE_X / ADJ_THREADS;
constexpr int TILE_ITEM = 1024;                      // listed rows per work item
constexpr int TILE_EMPTY_GROUP = 8;                  // empty tiles zeroed per work item
constexpr int FWD_ITEM = 1024;                       // rows per forward work item
template <int KM>
constexpr int fwd_win() { return (TILE_Y + KM - 1) * (TILE_X + KM - 1); }   // forward window (complex)
constexpr int TILE_GCH = ADJ_THREADS;                                        // rows staged per chunk (gather)
template <int KM>
constexpr int tile_cells_max() { return (TILE_Y + KM - 1) * (TILE_X + KM - 1); }

// LDS of a tile workgroup (dynamic shared memory).  Matrix-core form
// (16-slot rows): the staged C and B slots of a chunk, then each staged
// row's window origin.  Gather form (8-slot rows): the staged rows with an
// odd stride of Kx + Ky + 1 complex values (conj(cx) w, then conj(cy)), so
// lanes reading consecutive rows hit different banks; their local cells;
// the cell start offsets.
__host__ __device__ inline size_t tile_lds_bytes(int Ky, int Kx);
__host__ __device__ inline size_t tile_lds_bytes_gather(int Ky, int Kx) {
  return (size_t)TILE_GCH * (Ky + Kx + 1) * 8 + (size_t)(TILE_GCH + ADJ_THREADS / 64) * 4 +
         ((size_t)(TILE_Y + Ky - 1) * (TILE_X + Kx - 1) + 1) * 4;
}

// LDS of an adjoint workgroup that may run the CSC unit path (bytes)
constexpr size_t ADJ_CSC_LDS_BYTES = ADJ_LDS * sizeof(float2);

inline bool tile_grid_ok(int Kty, int Ktx, int Ky, int Kx) {
  return Kty % TILE_Y == 0 && Ktx % TILE_X == 0 && Kty >= TILE_Y + Ky - 1 && Ktx >= TILE_X + Kx - 1;
}

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// Matrix-core tile layout in LDS (16-slot rows).  Staged row i's slots sit
// at [i * TM_STRIDE + TM_GAP, + 16) of the C and B arrays, with TM_GAP zero
// slots before the first row and after every row: a lane reading slot
// jy = qy0 + lane - dy of a row whose window meets its quadrant lands in
// [-15, 30] around the row's first slot, inside its values or the zero gaps,
// so the reads need no bounds checks.
constexpr int TM_GAP = 16, TM_STRIDE = 32;
template <int CH>
__host__ __device__ constexpr size_t tile_mfma_lds_bytes() {
  return 2 * (size_t)(CH * TM_STRIDE + TM_GAP) * 8 + (size_t)CH * 4 + (size_t)4 * (CH + 4) * 4;
}
constexpr int TILE_MCH = 32;                          // rows staged per chunk (matrix cores)

__host__ __device__ inline size_t tile_lds_bytes(int Ky, int Kx) {
  return (Ky > 8 || Kx > 8) ? tile_mfma_lds_bytes<TILE_MCH>() : tile_lds_bytes_gather(Ky, Kx);
}

template <int KM, class W>
__device__ __forceinline__ void g_adjoint_tile_mfma(const GDev& g, int item, W&& w, float2* __restrict__ out,
                                                    unsigned char* lds) {
  static_assert(ADJ_THREADS == 256, "one wave per 16 x 16 quadrant of a 32 x 32 tile");
  static_assert(KM <= 16 && TM_GAP >= 15 && TM_STRIDE - 16 >= 15, "zero gaps cover a 16-lane read");
  constexpr int CH = TILE_MCH;
  constexpr int NSLOT = CH * TM_STRIDE + TM_GAP;
  float2* C = reinterpret_cast<float2*>(lds);
  float2* B = C + NSLOT;
  int* meta = reinterpret_cast<int*>(B + NSLOT);       // [CH] window origin of staged row i
  int* wl = meta + CH;                                 // [4][CH + 4] this wave's rows, as slot offsets
  const int4 it = g.items[item];
  if (it.w == -2) {   // a group of empty tiles: zeros
    for (int e = it.y; e < it.z; ++e) {
      const int t = g.empty_tiles[e];
      const int ty = t % g.tiles_y, tx = t / g.tiles_y;
#pragma unroll
      for (int q = 0; q < TILE_PTS; ++q) {
        const int p = threadIdx.x + q * ADJ_THREADS;
        out[(size_t)(tx * TILE_X + p / TILE_Y) * g.Kty + ty * TILE_Y + p % TILE_Y] = make_float2(0.f, 0.f);
      }
    }
    return;
  }
  const int t = it.x;
  const int ty = t % g.tiles_y, tx = t / g.tiles_y;
  const int ncy = TILE_Y + g.Ky - 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15;
  const bool odd = (lane >> 4) & 1;
  const int half = lane >> 5;                          // A/B lanes 0-31: first row of a pair, 32-63: second
  const int qy0 = (wv & 1) * 16, qx0 = (wv >> 1) * 16;
  int* mine = wl + wv * (CH + 4);
  // the gaps (and unused slots) are zero for the whole item
  for (int k = threadIdx.x; k < NSLOT; k += ADJ_THREADS) {
    C[k] = make_float2(0.f, 0.f);
    B[k] = make_float2(0.f, 0.f);
  }
  mfma_f4 cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f}, cre2 = cre, cim2 = cim;
  // staging: TPR = ADJ_THREADS / CH threads per row, thread (row sk, part
  // sp) loads 16 / TPR contiguous 16-byte words of the row's 2 KM slots
  // (the first half of the parts cy, the second cx), so a chunk's rows
  // arrive coalesced by all four waves; the next chunk's loads are issued
  // before this chunk's matrix-core loop
  constexpr int TPR = ADJ_THREADS / CH, FPT = 16 / TPR, SPT = 2 * FPT;   // float4 / slots per thread
  static_assert(ADJ_THREADS % CH == 0 && 16 % TPR == 0 && KM == 16 && TPR >= 2, "row staging split");
  const int sk = threadIdx.x / TPR, sp = threadIdx.x % TPR;
  float4 pre[FPT];
  uint4 preh = make_uint4(0u, 0u, 0u, 0u);   // FP16 rows: the raw word, converted when staged
  float2 pv = make_float2(0.f, 0.f);
  int pm = 0;
  bool pok = false;
  auto fetch = [&](int c) {
    pok = sk < min(CH, it.z - c);
    if (pok) {
      const int r = g.tile_rows[c + sk];
      const int cell = g.tile_cell[c + sk];
      if (sp >= TPR / 2) pv = w(r);
      if (g.half) {   // the same 4 slots in one 16-byte word of FP16 pairs
        preh = reinterpret_cast<const uint4*>(reinterpret_cast<const __half2*>(g.vals) + (size_t)r * (2 * KM))[sp];
      } else {
        const float4* src = reinterpret_cast<const float4*>(g.vals + (size_t)r * (2 * KM)) + FPT * sp;
#pragma unroll
        for (int j = 0; j < FPT; ++j) pre[j] = src[j];
      }
      pm = (cell % ncy - (g.Ky - 1) + 32) | ((cell / ncy - (g.Kx - 1) + 32) << 8);
    }
  };
  fetch(it.y);
  for (int c0 = it.y; c0 < it.z; c0 += CH) {
    const int n = min(CH, it.z - c0);
    __syncthreads();   // the previous chunk's reads are done (and the zeroing)
    if (pok) {
      if (g.half) {
        static_assert(FPT == 2, "FP16 rows: 4 slots per thread");
        const float2 a = h2f(preh.x), b = h2f(preh.y), c = h2f(preh.z), d = h2f(preh.w);
        pre[0] = make_float4(a.x, a.y, b.x, b.y);
        pre[1] = make_float4(c.x, c.y, d.x, d.y);
      }
      const int s0 = (sp % (TPR / 2)) * SPT;   // first slot of this thread's part
      if (sp < TPR / 2) {
        float2* qc = C + sk * TM_STRIDE + TM_GAP + s0;
#pragma unroll
        for (int j = 0; j < FPT; ++j) {   // conj(cy)
          qc[2 * j] = s0 + 2 * j < g.Ky ? make_float2(pre[j].x, -pre[j].y) : make_float2(0.f, 0.f);
          qc[2 * j + 1] = s0 + 2 * j + 1 < g.Ky ? make_float2(pre[j].z, -pre[j].w) : make_float2(0.f, 0.f);
        }
        if (sp == 0) meta[sk] = pm;
      } else {
        float2* qb = B + sk * TM_STRIDE + TM_GAP + s0;
        const float2 v = pv;
#pragma unroll
        for (int j = 0; j < FPT; ++j) {   // conj(cx) v
          qb[2 * j] = s0 + 2 * j < g.Kx ? make_float2(pre[j].x * v.x + pre[j].y * v.y, pre[j].x * v.y - pre[j].y * v.x)
                                        : make_float2(0.f, 0.f);
          qb[2 * j + 1] = s0 + 2 * j + 1 < g.Kx
                              ? make_float2(pre[j].z * v.x + pre[j].w * v.y, pre[j].z * v.y - pre[j].w * v.x)
                              : make_float2(0.f, 0.f);
        }
      }
    }
    __syncthreads();
    // this wave's rows (window meets its quadrant), listed as the slot
    // offsets of lane 0's reads: C at row + qy0 - dy, B at row + qx0 - dx
    int nh = 0;
    for (int b0 = 0; b0 < n; b0 += 64) {
      const int k = b0 + lane;
      bool hit = false;
      int e = 0;
      if (k < n) {
        const int m = meta[k];
        const int dy = (m & 0xff) - 32, dx = ((m >> 8) & 0xff) - 32;
        hit = dy < qy0 + 16 && dy + g.Ky > qy0 && dx < qx0 + 16 && dx + g.Kx > qx0;
        const int base = k * TM_STRIDE + TM_GAP;
        e = (base + qy0 - dy) | ((base + qx0 - dx) << 16);
      }
      const unsigned long long mask = __ballot(hit);
      const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
      if (hit) mine[nh + pos] = e;
      nh += __popcll(mask);
    }
    // rows missing from the last group of four read the leading zero gap
    if (lane < 3) mine[nh + lane] = 0;
    __syncthreads();
    if (c0 + CH < it.z) fetch(c0 + CH);
    // rank-1 updates, two rows per instruction pair and two pairs per step
    // (their LDS reads issued together): A lane (i, k) holds re / im (k
    // even / odd) of C at its row's offset + i, B lane (j, k) the matching
    // re / -im or im / re of B, so cre += Re(C B^T), cim += Im(C B^T)
    for (int p = 0; p < nh; p += 4) {
      const int e0 = mine[p + half], e1 = mine[p + 2 + half];
      const float2 c0 = C[(e0 & 0xffff) + li], b0 = B[(e0 >> 16) + li];
      const float2 c1 = C[(e1 & 0xffff) + li], b1 = B[(e1 >> 16) + li];
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(odd ? c0.y : c0.x, odd ? -b0.y : b0.x, cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(odd ? c0.y : c0.x, odd ? b0.x : b0.y, cim, 0, 0, 0);
      cre2 = __builtin_amdgcn_mfma_f32_16x16x4f32(odd ? c1.y : c1.x, odd ? -b1.y : b1.x, cre2, 0, 0, 0);
      cim2 = __builtin_amdgcn_mfma_f32_16x16x4f32(odd ? c1.y : c1.x, odd ? b1.x : b1.y, cim2, 0, 0, 0);
    }
  }
  cre += cre2;   // the two pairs' chains (four independent accumulators in the loop)
  cim += cim2;
  // lane (j, k) holds rows 4k .. 4k + 3 of column j of the quadrant
  const int lx = qx0 + li, ly0 = qy0 + 4 * (lane >> 4);
  if (it.w >= 0) {   // one of several items of this tile: partial sums
    float2* part = g.partial + (size_t)it.w * (TILE_Y * TILE_X) + lx * TILE_Y + ly0;
#pragma unroll
    for (int r = 0; r < 4; ++r) part[r] = make_float2(cre[r], cim[r]);
    return;
  }
  float2* o = out + (size_t)(tx * TILE_X + lx) * g.Kty + ty * TILE_Y + ly0;
#pragma unroll
  for (int r = 0; r < 4; ++r) o[r] = make_float2(cre[r], cim[r]);
}

// ---- the cell-indexed gather (8-slot rows) ----

struct TileView {
  float2* qc;
  int* cell;
  int* start;
  int stride;
};

__device__ __forceinline__ TileView tile_view(const GDev& g, unsigned char* lds) {
  TileView v;
  v.stride = g.Ky + g.Kx + 1;
  v.qc = reinterpret_cast<float2*>(lds);
  v.cell = reinterpret_cast<int*>(lds + (size_t)TILE_GCH * v.stride * 8);
  v.start = v.cell + TILE_GCH + ADJ_THREADS / 64;   // cell[TILE_GCH], scan scratch, start
  return v;
}

// a[c] = min(a[c], a[c+1], ..., a[n-1]) for c < n, by the ADJ_THREADS
// threads of the workgroup (each owns a contiguous segment); wsum: LDS
// scratch of ADJ_THREADS / 64 ints.  Ends with a barrier.
template <int KM>
__device__ __forceinline__ void tile_suffix_min(int* a, int n, int* wsum) {
  constexpr int PER = (tile_cells_max<KM>() + 1 + ADJ_THREADS - 1) / ADJ_THREADS;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int v[PER];
  int m = 0x7fffffff;
#pragma unroll
  for (int j = PER - 1; j >= 0; --j) {
    const int idx = t * PER + j;
    if (idx < n) m = min(m, a[idx]);
    v[j] = m;
  }
  // inclusive suffix minimum of the segment minima over the lanes of the wave
  int x = m;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_down(x, o, 64);
    if (lane + o < 64) x = min(x, y);
  }
  if (lane == 0) wsum[wv] = x;
  int ex = __shfl_down(x, 1, 64);
  if (lane == 63) ex = 0x7fffffff;
  __syncthreads();
  for (int u = wv + 1; u < ADJ_THREADS / 64; ++u) ex = min(ex, wsum[u]);
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int idx = t * PER + j;
    if (idx < n) a[idx] = min(v[j], ex);
  }
  __syncthreads();
}

template <int KM, class W>
__device__ __forceinline__ void g_adjoint_tile_gather(const GDev& g, int item, W&& w, float2* __restrict__ out,
                                                      unsigned char* lds) {
  const TileView s = tile_view(g, lds);
  const int4 it = g.items[item];
  if (it.w == -2) {   // a group of empty tiles: zeros
    for (int e = it.y; e < it.z; ++e) {
      const int t = g.empty_tiles[e];
      const int ty = t % g.tiles_y, tx = t / g.tiles_y;
#pragma unroll
      for (int q = 0; q < TILE_PTS; ++q) {
        const int p = threadIdx.x + q * ADJ_THREADS;
        out[(size_t)(tx * TILE_X + p / TILE_Y) * g.Kty + ty * TILE_Y + p % TILE_Y] = make_float2(0.f, 0.f);
      }
    }
    return;
  }
  const int t = it.x;
  const int ty = t % g.tiles_y, tx = t / g.tiles_y;
  const int ncy = TILE_Y + g.Ky - 1, ncx = TILE_X + g.Kx - 1, ncells = ncy * ncx;
  const int p0 = it.y, p1 = it.z;
  float2 acc[TILE_PTS];
#pragma unroll
  for (int i = 0; i < TILE_PTS; ++i) acc[i] = make_float2(0.f, 0.f);
  // the next chunk's row index and cell are loaded one chunk ahead, so the
  // row-indexed loads of a chunk do not wait behind the index load
  const int i = threadIdx.x;
  int r_next = 0, cell_next = 0;
  if (p0 + i < p1) {
    r_next = g.tile_rows[p0 + i];
    cell_next = g.tile_cell[p0 + i];
  }
  for (int c0 = p0; c0 < p1; c0 += TILE_GCH) {
    const int n = min(TILE_GCH, p1 - c0);
    const int r = r_next, rcell = cell_next;
    if (c0 + TILE_GCH + i < p1) {
      r_next = g.tile_rows[c0 + TILE_GCH + i];
      cell_next = g.tile_cell[c0 + TILE_GCH + i];
    }
    if (i < n) {
      const float2 v = w(r);
      float2 cy[KM], cx[KM];
      sep_load<KM>(g, r, cy, cx);
      float2* q = s.qc + (size_t)i * s.stride;
#pragma unroll
      for (int j = 0; j < KM; ++j) {
        if (j < g.Kx) {
          const float2 c = cx[j];
          q[j] = make_float2(c.x * v.x + c.y * v.y, c.x * v.y - c.y * v.x);   // conj(c) v
        }
        if (j < g.Ky) q[g.Kx + j] = make_float2(cy[j].x, -cy[j].y);
      }
      s.cell[i] = rcell;
    }
    for (int c = threadIdx.x; c <= ncells; c += ADJ_THREADS) s.start[c] = n;
    __syncthreads();
    // start[c] = first staged row with cell >= c: the first row of every
    // occupied cell marks it, then a block-wide suffix minimum fills the rest
    if (i < n && (i == 0 || s.cell[i - 1] != s.cell[i])) s.start[s.cell[i]] = i;
    __syncthreads();
    tile_suffix_min<KM>(s.start, ncells + 1, s.cell + TILE_GCH);
    // occupied cell columns of this chunk (cells are sorted, x-major)
    const int cxmin = s.cell[0] / ncy, cxmax = s.cell[n - 1] / ncy;
#pragma unroll
    for (int q = 0; q < TILE_PTS; ++q) {
      const int p = threadIdx.x + q * ADJ_THREADS;
      const int ly = p % TILE_Y, lx = p / TILE_Y;
      float2 a = acc[q];
      const int jx0 = max(0, lx + g.Kx - 1 - cxmax), jx1 = min(g.Kx - 1, lx + g.Kx - 1 - cxmin);
      for (int jx = jx0; jx <= jx1; ++jx) {
        const int base = (lx + g.Kx - 1 - jx) * ncy;
        const int k0 = s.start[base + ly], k1 = s.start[base + ly + g.Ky];
        for (int k = k0; k < k1; ++k) {
          const int jy = ly + g.Ky - 1 - (s.cell[k] - base);
          const float2* qq = s.qc + (size_t)k * s.stride;
          const float2 cy = qq[g.Kx + jy], qx = qq[jx];
          a.x += cy.x * qx.x - cy.y * qx.y;
          a.y += cy.x * qx.y + cy.y * qx.x;
        }
      }
      acc[q] = a;
    }
    __syncthreads();
  }
  if (it.w >= 0) {   // one of several items of this tile: partial sums
    float2* part = g.partial + (size_t)it.w * (TILE_Y * TILE_X);
#pragma unroll
    for (int q = 0; q < TILE_PTS; ++q) part[threadIdx.x + q * ADJ_THREADS] = acc[q];
    return;
  }
#pragma unroll
  for (int q = 0; q < TILE_PTS; ++q) {
    const int p = threadIdx.x + q * ADJ_THREADS;
    const int ly = p % TILE_Y, lx = p / TILE_Y;
    out[(size_t)(tx * TILE_X + lx) * g.Kty + ty * TILE_Y + ly] = acc[q];
  }
}


template <int KM, class W>
__device__ __forceinline__ void g_adjoint_tile(const GDev& g, int item, W&& w, float2* __restrict__ out,
                                               unsigned char* lds) {
  if constexpr (KM > 8)
    g_adjoint_tile_mfma<KM>(g, item, w, out, lds);
  else
    g_adjoint_tile_gather<KM>(g, item, w, out, lds);
}

// sum of the partial tiles of split tile r, in slot order
__device__ __forceinline__ void g_adjoint_tile_reduce(const GDev& g, int r, float2* __restrict__ out) {
  const int4 rd = g.red[r];
  const int t = rd.x, ty = t % g.tiles_y, tx = t / g.tiles_y;
#pragma unroll
  for (int q = 0; q < TILE_PTS; ++q) {
    const int p = threadIdx.x + q * ADJ_THREADS;
    float2 a = make_float2(0.f, 0.f);
    for (int j = 0; j < rd.z; ++j) a = cadd(a, g.partial[(size_t)(rd.y + j) * (TILE_Y * TILE_X) + p]);
    out[(size_t)(tx * TILE_X + p / TILE_Y) * g.Kty + ty * TILE_Y + p % TILE_Y] = a;
  }
}
// Forward (degridding) of one work item: the (TILE_Y + Ky - 1) x (TILE_X +
// Kx - 1) grid window of the tile is staged in LDS once (all loads issued
// before the stores), then every row of the item reads its Ky x Kx taps from
// LDS instead of gathering them through the caches.  emit(r, G Fx) stores
// row r's value.
template <int KM, class P, class E>
__device__ __forceinline__ void g_forward_tile(const GDev& g, int item, const float2* __restrict__ Fx,
                                               float2* win, P&& pre, E&& emit) {
  const int4 it = g.fitems[item];
  const int t = it.x;
  const int ty0 = (t % g.tiles_y) * TILE_Y, tx0 = (t / g.tiles_y) * TILE_X;
  const int wy = TILE_Y + g.Ky - 1, wx = TILE_X + g.Kx - 1, nw = wy * wx;
  constexpr int PER = (fwd_win<KM>() + ADJ_THREADS - 1) / ADJ_THREADS;
  float2 v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + i * ADJ_THREADS;
    v[i] = make_float2(0.f, 0.f);
    if (e < nw) {
      int ky = ty0 + e % wy, kx = tx0 + e / wy;
      if (ky >= g.Kty) ky -= g.Kty;
      if (kx >= g.Ktx) kx -= g.Ktx;
      v[i] = Fx[(size_t)kx * g.Kty + ky];
    }
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + i * ADJ_THREADS;
    if (e < nw) win[e] = v[i];
  }
  __syncthreads();
  int k = it.y + threadIdx.x;
  int rn = k < it.z ? g.prim_rows[k] : 0;
  for (; k < it.z; k += ADJ_THREADS) {
    const int r = rn;   // the next row index is loaded one step ahead
    if (k + ADJ_THREADS < it.z) rn = g.prim_rows[k + ADJ_THREADS];
    const auto pf = pre(r);   // the caller's per-row loads, issued before the arithmetic
    const int a = g.anchor[r];
    const int ly = a % g.Kty - ty0, lx = a / g.Kty - tx0;
    float2 cy[KM], cx[KM];
    sep_load<KM>(g, r, cy, cx);
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int jx = 0; jx < KM; ++jx) {
      if (jx < g.Kx) {
        const float2* col = win + (lx + jx) * wy + ly;
        float2 s = make_float2(0.f, 0.f);
#pragma unroll
        for (int jy = 0; jy < KM; ++jy) {
          if (jy < g.Ky) {
            const float2 x = col[jy];
            s.x += cy[jy].x * x.x - cy[jy].y * x.y;
            s.y += cy[jy].x * x.y + cy[jy].y * x.x;
          }
        }
        const float2 c = cx[jx];
        acc.x += c.x * s.x - c.y * s.y;
        acc.y += c.x * s.y + c.y * s.x;
      }
    }
    emit(r, acc, pf);
  }
}

// host construction helpers (gmat.hip)
std::vector<int> adjoint_units(const std::vector<int>& colptr);
void sep_csc(int64_t M, const int32_t* anchor, int Kty, int Ktx, int Ky, int Kx, std::vector<int>& colptr,
             std::vector<unsigned>& pk);

}  // namespace fhs

struct fhs_gmat {
  int device = 0;
  fhs::GDev d{};
  int64_t M = 0, ncols = 0, nnz = 0;
  int64_t n_modes = 0;   // holographic form: active Fourier modes (0 for a measurement G)
  void* mem[14] = {};
};
