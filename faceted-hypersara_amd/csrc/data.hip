// Data-fidelity dual step of all channels and blocks
// (src/update_dual_data_fidelity.m:126-147, non-DR branch; DR branch :85-125).
//
// Per iteration:
//   Fx   = A(x_l)                           (batched scaled FFT)
//   gx   = G_l Fx ;  r2 = 2 gx - gx_old ;  gx_old <- gx
//        (by linearity G(2 Fx - Fx_old) = 2 G Fx - G Fx_old: the state keeps
//         G Fx_old, M complex per channel, instead of Fx_old, No per channel)
//   per block: proj = ellipsoid projection (forward-backward, warm start),
//              v2 += pU (r2 - proj), norm_res = ||gx - y||      (FP64)
//   g2_l = G_l' v2_l (atomic-free gather per grid point)
//   Ftx_l = sigma22_l real(At(g2_l))
// DR branch: G is square lower-triangular (holographic H = G + G'), the
// dual uses the closed-form l2 prox, no ellipsoid.
#include <algorithm>

#include "gmat.hpp"
#include "kernels.hpp"
#include "nufft.hpp"

struct fhs_data {
  int device = 0;
  int L = 0;
  int nblocks = 0;
  int64_t Mtot = 0;
  int maxM = 0;
  int adj_grid = 0;                    // max adjoint units + grid-tail units over CSC channels
  int tile_grid = 0;                   // max tile work items over tile-gridded channels
  int red_grid = 0;                    // max split tiles over tile-gridded channels
  size_t tile_lds = 0;                 // LDS bytes of a tile workgroup (max over channels)
  int kmax = 0;                        // slots per axis of the tiled channels' separable rows
  int fwd_grid = 0;                    // max forward work items over tile-gridded channels
  bool any_rows = false;               // some channel without tile gridding (row kernels)
  int64_t No = 0;
  const fhs_nufft* A = nullptr;
  std::vector<int64_t> row_off;        // [L+1]
  std::vector<int64_t> blk_start;      // [nblocks+1]
  std::vector<int> blk_chan;           // [nblocks]
  bool square = false;                 // all G square (DR holographic form)
  fhs::GDev* gdev = nullptr;           // device [L]
  std::vector<fhs::GDev> gd;           // host copy of gdev
  std::vector<const float2*> rows32;   // each channel's G rows as created (FP32)
  std::vector<void*> rows16;           // FP16 copies of the tiled separable rows (row precision 16)
  int64_t* d_row_off = nullptr;        // device [L+1]
  int64_t* d_blk_start = nullptr;      // device [nblocks+1]
  double* d_eps = nullptr;             // device [nblocks]
  float* d_sigma22 = nullptr;          // device [L]
  double* d_norm_res = nullptr;        // device [nblocks] (squared residual norms)
  float2* Fx = nullptr;                // [L][No]
  float2* g2 = nullptr;                // [L][No]
  float2* r2 = nullptr;                // [Mtot]
  int fb_split = 1;                    // workgroups per block in the FB projection
  unsigned* fb_sync = nullptr;         // device [2 nblocks + 1]: arrival counts, generations, error flag
  double* fb_part = nullptr;           // device [nblocks][2][FB_SPLIT_MAX][FB_PART], then [nblocks][FB_SPLIT_MAX]
  bool stencil = false;                // DR form applied through the holographic stencils
  int l2_split = 1;                    // workgroups per block in the DR dual step
  double* l2_part = nullptr;           // device [nblocks][L2_SPLIT_MAX][2] (square G only)
};

namespace fhs {
namespace {

constexpr int FB_THREADS = 512;
constexpr int NXCD = 8;   // MI355X: 8 XCDs, each with its own L2

// Workgroups of a launch are dispatched to the XCDs round-robin (linear id
// mod 8).  The per-channel kernels below are launched as a 1-D grid of
// NXCD * ceil(L / NXCD) * units workgroups and remapped so that XCD x only
// runs channels x, x + 8, ...: one channel's v2 / coefficients / grid
// region is then shared by one L2 instead of being pulled into all eight.
// Units are interleaved over those channels (unit-major), so the first
// units of every channel (the heaviest adjoint tiles) start first.
__device__ __forceinline__ bool xcd_map(int L, int units, int& chan, int& unit) {
  const int b = blockIdx.x;
  const int xcd = b % NXCD, s = b / NXCD;
  const int cpx = (L + NXCD - 1) / NXCD;   // channels per XCD
  chan = xcd + NXCD * (s % cpx);
  unit = s / cpx;
  return chan < L && unit < units;
}

inline unsigned xcd_grid(int L, int units) {
  return (unsigned)(NXCD * ((L + NXCD - 1) / NXCD) * units);
}

__global__ __launch_bounds__(256) void k_data_forward(const GDev* __restrict__ gd,
                                                      const int64_t* __restrict__ row_off,
                                                      const float2* __restrict__ Fx, int64_t No,
                                                      float2* __restrict__ gx_old,
                                                      float2* __restrict__ r2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (g.tiled) return;   // k_data_forward_tiled
  const int m = u * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const float2 gx = g_row_dot(g, m, Fx + (size_t)l * No);
  const int64_t i = row_off[l] + m;
  const float2 old = gx_old[i];
  r2[i] = make_float2(2.f * gx.x - old.x, 2.f * gx.y - old.y);
  gx_old[i] = gx;
}

// Holographic (DR) forward: Lambda (G + G') applied to (2 Fx - Fx_old), and
// the residual operator on Fx, G square M x M (STATE); or y = Lambda (G + G') Fx
// into gx_old (the DR measurement operator, r2 unused).
template <bool STATE>
__global__ __launch_bounds__(256) void k_data_forward_dr(const GDev* __restrict__ gd,
                                                         const int64_t* __restrict__ row_off,
                                                         const float2* __restrict__ Fx, int64_t No,
                                                         const float* __restrict__ lam,
                                                         float2* __restrict__ gx_old,
                                                         float2* __restrict__ r2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  const int m = u * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const float2* F = Fx + (size_t)l * No;
  float2 gx = cadd(g_row_dot(g, m, F), g_col_dot(g, m, [&](int r) { return F[r]; }));
  const int64_t i = row_off[l] + m;
  const float lm = lam ? lam[i] : 1.f;
  gx = cscale(gx, lm);
  if (STATE) {
    const float2 old = gx_old[i];
    r2[i] = make_float2(2.f * gx.x - old.x, 2.f * gx.y - old.y);
  }
  gx_old[i] = gx;
}

// Holographic form applied as a stencil (G square over the grid, with its
// full Hermitian stencil G'G = H + H'): per grid point p of channel l,
//   s_p = sum over the neighbourhood offsets o of stencil[o][act[p]] in(p + o)
// (grid wrap).  DR_STATE: in = Fx, gx = Lambda_p s_p, r2 = 2 gx - gx_old,
// gx_old <- gx (rows off W are left alone: they stay zero); DR_APPLY:
// y_p = Lambda_p s_p, 0 off W; DR_ADJ: in = Lambda .* v2, g2_p = s_p, 0 off W.
enum DrMode { DR_STATE = 0, DR_APPLY = 1, DR_ADJ = 2 };

template <int MODE>
__global__ __launch_bounds__(256) void k_dr_stencil(const GDev* __restrict__ gd,
                                                    const int64_t* __restrict__ row_off,
                                                    const float2* __restrict__ in, int64_t No,
                                                    const float* __restrict__ lam, float2* __restrict__ out,
                                                    float2* __restrict__ r2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  const int p = u * blockDim.x + threadIdx.x;
  if (p >= g.M) return;
  const int64_t base = row_off[l];
  const int i = g.act[p];
  if (i < 0) {
    if (MODE != DR_STATE) out[(MODE == DR_ADJ ? (int64_t)l * No : base) + p] = make_float2(0.f, 0.f);
    return;
  }
  const float2* __restrict__ src = MODE == DR_ADJ ? in + base : in + (size_t)l * No;
  const float* __restrict__ lm = lam ? lam + base : nullptr;
  const int py = p % g.Kty, px = p / g.Kty;
  const int ny = 2 * g.Ky - 1, nx = 2 * g.Kx - 1;
  const float2* __restrict__ st = g.stencil + i;
  float ax = 0.f, ay = 0.f;
  for (int ox = 0; ox < nx; ++ox) {
    int qx = px + ox - (g.Kx - 1);
    qx += qx < 0 ? g.Ktx : (qx >= g.Ktx ? -g.Ktx : 0);
    const int col = qx * g.Kty;
    for (int oy = 0; oy < ny; ++oy) {
      int qy = py + oy - (g.Ky - 1);
      qy += qy < 0 ? g.Kty : (qy >= g.Kty ? -g.Kty : 0);
      const float2 h = st[(size_t)(oy + ny * ox) * g.nact];
      float2 v = src[col + qy];
      if (MODE == DR_ADJ && lm) v = cscale(v, lm[col + qy]);
      ax += h.x * v.x - h.y * v.y;
      ay += h.x * v.y + h.y * v.x;
    }
  }
  float2 s = make_float2(ax, ay);
  if (MODE == DR_ADJ) {
    out[(int64_t)l * No + p] = s;
    return;
  }
  if (lm) s = cscale(s, lm[p]);
  if (MODE == DR_APPLY) {
    out[base + p] = s;
  } else {
    const float2 old = out[base + p];
    r2[base + p] = make_float2(2.f * s.x - old.x, 2.f * s.y - old.y);
    out[base + p] = s;
  }
}

// DR dual step v2 <- sc(v2 + r2 - y, eps) (closed-form l2 prox, sc(z, r) =
// z (1 - min(r/||z||, 1))) and the residual ||gx - y||, over many
// workgroups: each block's rows are cut into S equal
// chunks; pass 1 sums ||v2 + r2 - y||^2 and ||gx - y||^2 per chunk, pass 2
// adds a block's chunk sums in chunk order (every workgroup of the block the
// same way) and scales its own chunk.
constexpr int L2_SPLIT_MAX = 256;
constexpr int L2_THREADS = 256;

__global__ __launch_bounds__(L2_THREADS) void k_l2_partial(const int64_t* __restrict__ blk_start, int S,
                                                           const float2* __restrict__ v2,
                                                           const float2* __restrict__ y,
                                                           const float2* __restrict__ r2,
                                                           const float2* __restrict__ gx,
                                                           double* __restrict__ part) {
  __shared__ double sh[2 * L2_THREADS / 64];
  const int b = blockIdx.x / S, c = blockIdx.x % S;
  const int64_t a0 = blk_start[b], a1 = blk_start[b + 1];
  const int64_t chunk = (a1 - a0 + S - 1) / S;
  const int64_t lo = a0 + c * chunk, hi = lo + chunk < a1 ? lo + chunk : a1;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += L2_THREADS) {
    const float2 vv = v2[i], rr = r2[i], yy = y[i], gg = gx[i];
    const double zx = (double)vv.x + rr.x - yy.x, zy = (double)vv.y + rr.y - yy.y;
    acc[0] += zx * zx + zy * zy;
    const double ex = (double)gg.x - yy.x, ey = (double)gg.y - yy.y;
    acc[1] += ex * ex + ey * ey;
  }
  block_sum<2>(acc, sh);
  if (threadIdx.x == 0) {
    part[((size_t)b * L2_SPLIT_MAX + c) * 2] = acc[0];
    part[((size_t)b * L2_SPLIT_MAX + c) * 2 + 1] = acc[1];
  }
}

__global__ __launch_bounds__(L2_THREADS) void k_l2_apply(const int64_t* __restrict__ blk_start, int S,
                                                         const double* __restrict__ eps_arr,
                                                         float2* __restrict__ v2, const float2* __restrict__ y,
                                                         const float2* __restrict__ r2,
                                                         const double* __restrict__ part,
                                                         double* __restrict__ norm_res) {
  __shared__ float fsh;
  const int b = blockIdx.x / S, c = blockIdx.x % S;
  if (threadIdx.x == 0) {
    double z = 0.0, e = 0.0;
    for (int k = 0; k < S; ++k) {
      z += part[((size_t)b * L2_SPLIT_MAX + k) * 2];
      e += part[((size_t)b * L2_SPLIT_MAX + k) * 2 + 1];
    }
    const double nz = sqrt(z);
    fsh = nz > 0.0 ? (float)(1.0 - fmin(eps_arr[b] / nz, 1.0)) : 0.f;
    if (c == 0) norm_res[b] = e;
  }
  __syncthreads();
  const float f = fsh;
  const int64_t a0 = blk_start[b], a1 = blk_start[b + 1];
  const int64_t chunk = (a1 - a0 + S - 1) / S;
  const int64_t lo = a0 + c * chunk, hi = lo + chunk < a1 ? lo + chunk : a1;
  for (int64_t i = lo + threadIdx.x; i < hi; i += L2_THREADS) {
    const float2 vv = v2[i], rr = r2[i], yy = y[i];
    v2[i] = make_float2((vv.x + rr.x - yy.x) * f, (vv.y + rr.y - yy.y) * f);
  }
}

// ---- forward-backward projection onto the data ellipsoid ------------------
// Per data block: z <- P_B(z - mu pU (z - w)), w = v2/pU + r2, mu = 1/max(pU),
// warm start z = proj, stop after min_it once ||dz|| <= tol ||z|| or at
// max_it; then the dual update v2 = pU (w - z) and the residual ||gx - y||.
//
// A block is split over S workgroups (rows in S contiguous parts), so the
// whole GPU works on the projection instead of one CU per block.  Each
// thread keeps FB_ROWS of its rows (w, z, y, pU) in registers for the whole
// call: the FB iterations then read no memory for them, only rows beyond
// FB_ROWS * FB_THREADS per workgroup are streamed (w kept in the v2 buffer,
// z in proj).  The block-wide sums every iteration needs are combined
// across the S workgroups by fb_combine.  The launch never has more
// workgroups than can be resident at once (fb_split), so every wait ends.
constexpr int FB_ROWS = 24;                 // rows per thread held in registers
constexpr int FB_SPLIT_MAX = 16;            // workgroups per block
constexpr int FB_PART = 4;                  // doubles per workgroup partial
constexpr unsigned FB_SPIN_LIMIT = 1u << 22;
// dynamic LDS per workgroup: y and pU of the resident rows
constexpr size_t FB_LDS = (size_t)FB_ROWS * FB_THREADS * (sizeof(float2) + sizeof(float));

struct FbSync {
  unsigned* cnt;    // [nblocks] arrivals in the current round
  unsigned* gen;    // [nblocks] completed rounds
  double* part;     // [nblocks][2][FB_SPLIT_MAX][FB_PART] partials, by round parity
  unsigned* err;    // set when a wait gave up
  double* res;      // [nblocks][FB_SPLIT_MAX] residual sums of the parts
};

// Combine NV per-thread values over the S workgroups of block blk (sum, or
// max with MAXOP).  Thread 0 publishes the workgroup's total in the slot of
// this round's parity and counts its arrival with an agent-scope atomic; the
// last arrival resets the count and advances the block's generation, the
// others wait for that.  Every workgroup then combines the S partials in
// slot order, so all of them hold bit-identical totals.  A workgroup reuses
// a slot two rounds later, after every workgroup has read it.
// All shared words are accessed with agent-scope atomics, which the GPU
// performs at its device-coherent level (past the per-XCD L2s), and each
// step waits for the previous one's completion (s_waitcnt 0) before it is
// issued; so no L2 writeback or invalidation is needed, which would also
// flush the dirty lines of kernels running beside this one.
// The wait is bounded: after FB_SPIN_LIMIT polls it sets *err and goes on
// (the results are then wrong; the host reports it).
template <bool MAXOP, int NV>
__device__ void fb_combine(const FbSync& fs, int blk, int part, int S, unsigned& gen, int& round,
                           double (&v)[NV], double* sh, double* bc) {
  if (MAXOP) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double m = v[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
      v[k] = m;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < NV; ++k) sh[k * (FB_THREADS / 64) + (threadIdx.x >> 6)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double m = sh[k * (FB_THREADS / 64)];
      for (int w = 1; w < FB_THREADS / 64; ++w) m = fmax(m, sh[k * (FB_THREADS / 64) + w]);
      v[k] = m;
    }
    __syncthreads();
  } else {
    block_sum<NV>(v, sh);
  }
  if (S > 1) {
    if (threadIdx.x == 0) {
      double* slot = fs.part + ((size_t)blk * 2 + (round & 1)) * (FB_SPLIT_MAX * FB_PART);
#pragma unroll
      for (int k = 0; k < NV; ++k)
        __hip_atomic_store(slot + part * FB_PART + k, v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0);   // the partials are stored before the arrival counts
      const unsigned old = __hip_atomic_fetch_add(fs.cnt + blk, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old == (unsigned)S - 1) {
        __hip_atomic_store(fs.cnt + blk, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);   // the reset lands before the generation advances
        __hip_atomic_store(fs.gen + blk, gen + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        unsigned spins = 0;
        while (__hip_atomic_load(fs.gen + blk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins >= FB_SPIN_LIMIT) {
            __hip_atomic_store(fs.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      ++gen;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        double t = __hip_atomic_load(slot + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int w = 1; w < S; ++w) {
          const double u = __hip_atomic_load(slot + w * FB_PART + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          t = MAXOP ? fmax(t, u) : t + u;
        }
        bc[k] = t;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = bc[k];
    __syncthreads();
  }
  ++round;
}

__global__ __launch_bounds__(FB_THREADS) void k_fb_dual(
    const int64_t* __restrict__ blk_start, const double* __restrict__ eps_arr,
    float2* __restrict__ v2, const float2* __restrict__ y, const float* __restrict__ pU,
    const float2* __restrict__ r2, float2* __restrict__ z, const float2* __restrict__ gx,
    int max_it, int min_it, double tol, int S, FbSync fs) {
  __shared__ double sh[4 * FB_THREADS / 64];
  __shared__ double bc[4];
  // y [FB_ROWS][FB_THREADS] then pU [FB_ROWS][FB_THREADS] of the resident
  // rows; every thread reads back only its own slots, so no barrier is needed
  extern __shared__ float2 fb_y[];
  const int blk = blockIdx.x / S, part = blockIdx.x % S;
  const int64_t a0 = blk_start[blk], b0 = blk_start[blk + 1];
  int64_t cs = (b0 - a0 + S - 1) / S;
  cs = (cs + 63) & ~(int64_t)63;
  const int64_t a = a0 + part * cs < b0 ? a0 + part * cs : b0;
  const int64_t b = a + cs < b0 ? a + cs : b0;
  const int64_t tail = a + (int64_t)FB_ROWS * FB_THREADS;   // first streamed row
  const int nloc = (int)(b - a);
  const double eps = eps_arr[blk];
  unsigned gen = 0;
  int round = 0;
  if (S > 1 && threadIdx.x == 0) gen = __hip_atomic_load(fs.gen + blk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  auto live = [&](int k) { return (int)threadIdx.x + k * FB_THREADS < nloc; };
  // Resident row k of this thread is a + threadIdx.x + k FB_THREADS.  Loads
  // use indices clamped into the part (or the block's first row for an empty
  // part), so they are unconditional and batch; rows past the part only
  // feed masked-out terms and are never stored.
  const int64_t base = nloc > 0 ? a : a0;
  const unsigned last = nloc > 0 ? (unsigned)nloc - 1u : 0u;
  // unsigned min: the index is provably below FB_ROWS * FB_THREADS, so the
  // loads take a scalar base and a 32-bit offset instead of a 64-bit address per row
  auto ridx = [&](int k) { return min(threadIdx.x + (unsigned)(k * FB_THREADS), last); };
  float2 rw[FB_ROWS], rz[FB_ROWS];
  float2* ry = fb_y + threadIdx.x;                                    // ry[k * FB_THREADS]
  float* rp = (float*)(fb_y + FB_ROWS * FB_THREADS) + threadIdx.x;   // rp[k * FB_THREADS]
  const float* pUa = pU ? pU + base : nullptr;
  const float2 *v2a = v2 + base, *r2a = r2 + base, *za = z + base, *ya = y + base;
  // two waves of loads (pU, v2, r2 -> w, then z, y) bound the live registers
#pragma unroll
  for (int k = 0; k < FB_ROWS; ++k) {
    const int i = ridx(k);
    const float p = pUa ? pUa[i] : 1.f;
    rp[k * FB_THREADS] = p;
    const float2 vv = v2a[i], r = r2a[i];
    rw[k] = make_float2(vv.x / p + r.x, vv.y / p + r.y);   // w = v2 / pU + r2
  }
#pragma unroll
  for (int k = 0; k < FB_ROWS; ++k) {
    const int i = ridx(k);
    rz[k] = za[i];
    ry[k * FB_THREADS] = ya[i];
  }
  // max(pU) over the block -> step (a clamped duplicate does not change a max)
  double pmv[1] = {0.0};
  if (nloc > 0) {
#pragma unroll
    for (int k = 0; k < FB_ROWS; ++k) pmv[0] = fmax(pmv[0], (double)rp[k * FB_THREADS]);
  }
  for (int64_t i = tail + threadIdx.x; i < b; i += FB_THREADS) pmv[0] = fmax(pmv[0], pU ? (double)pU[i] : 1.0);
  fb_combine<true, 1>(fs, blk, part, S, gen, round, pmv, sh, bc);
  const float pmax = (float)pmv[0];
  const float pm = pmax > 0.f ? pmax : 1.f;
  const float inv_pm = 1.f / pm;
  // a = pU / max pU as one multiply; exactly 1 where pU is maximal
  auto step = [&](float p) { return p == pm ? 1.f : p * inv_pm; };
  // gradient step z - mu pU (z - w), mu = 1/max pU, as (1-a) z + a w with
  // a = pU / max pU (a == 1 exactly where pU is maximal, so a constant pU
  // reproduces w bit for bit and the iteration stops at its fixed point).
  // Pass 1: first gradient-step distance to y.
  double s1[1] = {0.0};
#pragma unroll
  for (int k = 0; k < FB_ROWS; ++k) {
    const float2 w = rw[k], yy = ry[k * FB_THREADS];
    const float aa = step(rp[k * FB_THREADS]);
    const float2 g = make_float2((1.f - aa) * rz[k].x + aa * w.x, (1.f - aa) * rz[k].y + aa * w.y);
    const double dx = (double)g.x - yy.x, dy = (double)g.y - yy.y;
    s1[0] += live(k) ? dx * dx + dy * dy : 0.0;
  }
  for (int64_t i = tail + threadIdx.x; i < b; i += FB_THREADS) {
    const float p = pU ? pU[i] : 1.f;
    const float2 vv = v2[i], r = r2[i], zz = z[i], yy = y[i];
    const float2 w = make_float2(vv.x / p + r.x, vv.y / p + r.y);
    v2[i] = w;
    const float aa = step(p);
    const float2 g = make_float2((1.f - aa) * zz.x + aa * w.x, (1.f - aa) * zz.y + aa * w.y);
    const double dx = (double)g.x - yy.x, dy = (double)g.y - yy.y;
    s1[0] += dx * dx + dy * dy;
  }
  fb_combine<false, 1>(fs, blk, part, S, gen, round, s1, sh, bc);
  double s_cur = s1[0];
  for (int it = 1; it <= max_it; ++it) {
    const double nrm = sqrt(s_cur);
    const float scale = nrm > 0.0 ? (float)fmin(eps / nrm, 1.0) : 1.f;
    double acc[3] = {0.0, 0.0, 0.0};  // ||dz||^2, ||z||^2, next ||g - y||^2
    auto upd = [&](float p, float2 w, float2& zz, float2 yy, bool ok) {
      const float aa = step(p);
      const float2 g = make_float2((1.f - aa) * zz.x + aa * w.x, (1.f - aa) * zz.y + aa * w.y);
      const float2 zn = make_float2(yy.x + (g.x - yy.x) * scale, yy.y + (g.y - yy.y) * scale);
      const double ddx = (double)zn.x - zz.x, ddy = (double)zn.y - zz.y;
      const double zx = zn.x, zy = zn.y;
      const float2 g2 = make_float2((1.f - aa) * zn.x + aa * w.x, (1.f - aa) * zn.y + aa * w.y);
      const double ex = (double)g2.x - yy.x, ey = (double)g2.y - yy.y;
      acc[0] += ok ? ddx * ddx + ddy * ddy : 0.0;
      acc[1] += ok ? zx * zx + zy * zy : 0.0;
      acc[2] += ok ? ex * ex + ey * ey : 0.0;
      zz = zn;
    };
#pragma unroll
    for (int k = 0; k < FB_ROWS; ++k) upd(rp[k * FB_THREADS], rw[k], rz[k], ry[k * FB_THREADS], live(k));
    for (int64_t i = tail + threadIdx.x; i < b; i += FB_THREADS) {
      float2 zz = z[i];
      upd(pU ? pU[i] : 1.f, v2[i], zz, y[i], true);
      z[i] = zz;
    }
    fb_combine<false, 3>(fs, blk, part, S, gen, round, acc, sh, bc);
    const double crit = acc[1] > 0.0 ? sqrt(acc[0] / acc[1]) : 0.0;
    if (it >= min_it && crit <= tol) break;
    s_cur = acc[2];
  }
  // dual update v2 = pU (w - z), z kept as the next warm start, residual ||gx - y||
  double res[1] = {0.0};
  const float2* gxa = gx + base;
  float2 *v2out = v2 + base, *zout = z + base;
#pragma unroll
  for (int k = 0; k < FB_ROWS; ++k) {
    const float2 gg = gxa[ridx(k)], yy = ry[k * FB_THREADS];
    const double ex = (double)gg.x - yy.x, ey = (double)gg.y - yy.y;
    res[0] += live(k) ? ex * ex + ey * ey : 0.0;
  }
#pragma unroll
  for (int k = 0; k < FB_ROWS; ++k) {
    if (live(k)) {
      const unsigned i = threadIdx.x + (unsigned)(k * FB_THREADS);
      const float p = rp[k * FB_THREADS];
      v2out[i] = make_float2(p * (rw[k].x - rz[k].x), p * (rw[k].y - rz[k].y));
      zout[i] = rz[k];
    }
  }
  for (int64_t i = tail + threadIdx.x; i < b; i += FB_THREADS) {
    const float p = pU ? pU[i] : 1.f;
    const float2 w = v2[i], zz = z[i], yy = y[i], gg = gx[i];
    v2[i] = make_float2(p * (w.x - zz.x), p * (w.y - zz.y));
    const double ex = (double)gg.x - yy.x, ey = (double)gg.y - yy.y;
    res[0] += ex * ex + ey * ey;
  }
  // the parts' residual sums are added by k_sum_blocks (no exchange needed)
  block_sum<1>(res, sh);
  if (threadIdx.x == 0) fs.res[(size_t)blk * FB_SPLIT_MAX + part] = res[0];
}

FbSync fb_sync(const fhs_data* D) {
  FbSync f;
  f.cnt = D->fb_sync;
  f.gen = D->fb_sync + D->nblocks;
  f.err = D->fb_sync + 2 * D->nblocks;
  f.part = D->fb_part;
  f.res = D->fb_part + (size_t)D->nblocks * 2 * FB_SPLIT_MAX * FB_PART;
  return f;
}

// Workgroups per block for the FB projection: enough that every row of the
// largest block is held in registers, at most FB_SPLIT_MAX, and never more
// workgroups in total than the GPU holds at once (the combine waits for all
// workgroups of a block, so they must be able to be resident together).
int fb_split_for(const fhs_data* D, int device) {
  int64_t maxb = 0;
  for (int b = 0; b < D->nblocks; ++b) maxb = std::max<int64_t>(maxb, D->blk_start[b + 1] - D->blk_start[b]);
  const int64_t per = (int64_t)FB_ROWS * FB_THREADS;
  int S = (int)std::min<int64_t>(FB_SPLIT_MAX, std::max<int64_t>(1, (maxb + per - 1) / per));
  int occ = 0, cus = 0;
  FHS_HIP(hipFuncSetAttribute((const void*)k_fb_dual, hipFuncAttributeMaxDynamicSharedMemorySize, (int)FB_LDS));
  FHS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fb_dual, FB_THREADS, FB_LDS));
  FHS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const int64_t cap = (int64_t)occ * cus;
  while (S > 1 && (int64_t)D->nblocks * S > cap) --S;
  return S;
}

// g2_l = G_l' (lam .* v2_l) (+ G_l (lam .* v2_l) in the DR form) for the
// channels whose G keeps a CSC copy: one workgroup per adjoint work unit;
// grid points past a channel's columns are zero.
__global__ __launch_bounds__(ADJ_THREADS) void k_data_adjoint(const GDev* __restrict__ gd,
                                                              const int64_t* __restrict__ row_off,
                                                              const float2* __restrict__ v2,
                                                              const float* __restrict__ lam, int dr,
                                                              int64_t No, float2* __restrict__ g2, int L,
                                                              int units) {
  __shared__ float2 prod[ADJ_LDS];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (g.tiled) return;
  float2* out = g2 + (size_t)l * No;
  if (u >= g.nunits) {
    // zero the grid tail beyond the matrix columns (only when ncols < No)
    const int64_t extra_units = (No - g.ncols + ADJ_COLS - 1) / ADJ_COLS;
    const int64_t t = u - g.nunits;
    if (t < extra_units) {
      const int64_t c0 = g.ncols + t * ADJ_COLS;
      for (int64_t c = c0 + threadIdx.x; c < c0 + ADJ_COLS && c < No; c += ADJ_THREADS)
        out[c] = make_float2(0.f, 0.f);
    }
    return;
  }
  const float2* v = v2 + row_off[l];
  const float* lm = lam ? lam + row_off[l] : nullptr;
  auto none = [](int) { return make_float2(0.f, 0.f); };
  auto wl = [&](int r) { return cscale(v[r], lm[r]); };
  auto wp = [&](int r) { return v[r]; };
  if (lm) {
    if (dr)
      g_adjoint_unit(g, u, wl, [&](int c) { return g_row_apply(g, c, wl); }, out, prod);
    else
      g_adjoint_unit(g, u, wl, none, out, prod);
  } else {
    if (dr)
      g_adjoint_unit(g, u, wp, [&](int c) { return g_row_apply(g, c, wp); }, out, prod);
    else
      g_adjoint_unit(g, u, wp, none, out, prod);
  }
}

// g2_l = G_l' (lam .* v2_l) for the channels with a tile-gridded separable G
// (one workgroup per grid tile; never the DR form, whose H is square).
template <bool LAM, int KM>
__global__ __launch_bounds__(ADJ_THREADS) void k_data_adjoint_tiled(const GDev* __restrict__ gd,
                                                                    const int64_t* __restrict__ row_off,
                                                                    const float2* __restrict__ v2,
                                                                    const float* __restrict__ lam, int64_t No,
                                                                    float2* __restrict__ g2, int L, int units) {
  extern __shared__ __align__(16) unsigned char t[];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nunits) return;
  const float2* __restrict__ v = v2 + row_off[l];
  const float* __restrict__ lm = LAM ? lam + row_off[l] : nullptr;
  if (LAM)
    g_adjoint_tile<KM>(g, u, [=](int r) { return cscale(v[r], lm[r]); }, g2 + (size_t)l * No, t);
  else
    g_adjoint_tile<KM>(g, u, [=](int r) { return v[r]; }, g2 + (size_t)l * No, t);
}

// sums of the partial tiles of the tiles split over several work items
__global__ __launch_bounds__(ADJ_THREADS) void k_data_adjoint_reduce(const GDev* __restrict__ gd, int64_t No,
                                                                     float2* __restrict__ g2, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nred) return;
  g_adjoint_tile_reduce(g, u, g2 + (size_t)l * No);
}

void launch_data_adjoint(fhs_ctx* ctx, const fhs_data* D, const float2* v2, const float* lam, int dr) {
  const int L = D->L;
  if (dr && D->stencil) {
    const int units = (int)cdiv(D->maxM, 256);
    k_dr_stencil<DR_ADJ><<<xcd_grid(L, units), 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, v2, D->No, lam, D->g2,
                                                                      nullptr, L, units);
    check_launch();
    return;
  }
  if (D->tile_grid > 0) {
    auto go = [&](auto kern) {
      allow_dynamic_lds((const void*)kern, D->tile_lds);
      kern<<<xcd_grid(L, D->tile_grid), ADJ_THREADS, D->tile_lds, ctx->stream>>>(D->gdev, D->d_row_off, v2, lam,
                                                                                 D->No, D->g2, L, D->tile_grid);
    };
    if (D->kmax == 8)
      lam ? go(k_data_adjoint_tiled<true, 8>) : go(k_data_adjoint_tiled<false, 8>);
    else
      lam ? go(k_data_adjoint_tiled<true, 16>) : go(k_data_adjoint_tiled<false, 16>);
    check_launch();
    if (D->red_grid > 0) {
      k_data_adjoint_reduce<<<xcd_grid(L, D->red_grid), ADJ_THREADS, 0, ctx->stream>>>(D->gdev, D->No, D->g2, L,
                                                                                       D->red_grid);
      check_launch();
    }
  }
  if (D->adj_grid > 0) {
    k_data_adjoint<<<xcd_grid(L, D->adj_grid), ADJ_THREADS, 0, ctx->stream>>>(D->gdev, D->d_row_off, v2, lam, dr,
                                                                              D->No, D->g2, L, D->adj_grid);
    check_launch();
  }
}

// y = lam .* (G_l A x_l) for all channels (stacked rows)
__global__ __launch_bounds__(256) void k_data_apply(const GDev* __restrict__ gd,
                                                    const int64_t* __restrict__ row_off,
                                                    const float2* __restrict__ Fx, int64_t No,
                                                    const float* __restrict__ lam,
                                                    float2* __restrict__ y, int L, int units) {
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (g.tiled) return;   // k_data_apply_tiled
  const int m = u * blockDim.x + threadIdx.x;
  if (m >= g.M) return;
  const int64_t i = row_off[l] + m;
  float2 v = g_row_dot(g, m, Fx + (size_t)l * No);
  if (lam) v = cscale(v, lam[i]);
  y[i] = v;
}

// the same two operations for the channels with a tile-gridded separable G:
// one workgroup per forward work item, grid window staged in LDS
template <int KM>
__global__ __launch_bounds__(ADJ_THREADS) void k_data_forward_tiled(const GDev* __restrict__ gd,
                                                                    const int64_t* __restrict__ row_off,
                                                                    const float2* __restrict__ Fx, int64_t No,
                                                                    float2* __restrict__ gx_old,
                                                                    float2* __restrict__ r2, int L, int units) {
  __shared__ float2 win[fwd_win<KM>()];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nfitems) return;
  float2* __restrict__ go = gx_old + row_off[l];
  float2* __restrict__ rr = r2 + row_off[l];
  g_forward_tile<KM>(
      g, u, Fx + (size_t)l * No, win, [=](int r) { return go[r]; },
      [=](int r, float2 gx, float2 old) {
        rr[r] = make_float2(2.f * gx.x - old.x, 2.f * gx.y - old.y);
        go[r] = gx;
      });
}

template <int KM>
__global__ __launch_bounds__(ADJ_THREADS) void k_data_apply_tiled(const GDev* __restrict__ gd,
                                                                  const int64_t* __restrict__ row_off,
                                                                  const float2* __restrict__ Fx, int64_t No,
                                                                  const float* __restrict__ lam,
                                                                  float2* __restrict__ y, int L, int units) {
  __shared__ float2 win[fwd_win<KM>()];
  int l, u;
  if (!xcd_map(L, units, l, u)) return;
  const GDev g = gd[l];
  if (!g.tiled || u >= g.nfitems) return;
  float2* __restrict__ yo = y + row_off[l];
  const float* __restrict__ lm = lam ? lam + row_off[l] : nullptr;
  g_forward_tile<KM>(
      g, u, Fx + (size_t)l * No, win, [=](int r) { return lm ? lm[r] : 1.f; },
      [=](int r, float2 v, float s) { yo[r] = lm ? cscale(v, s) : v; });
}

// G Fx -> r2 = 2 G Fx - gx_old, gx_old <- G Fx (non-DR form) for all channels
void launch_data_forward(fhs_ctx* ctx, const fhs_data* D, float2* gx_old, unsigned row_grid, int row_units) {
  if (D->fwd_grid > 0) {
    if (D->kmax == 8)
      k_data_forward_tiled<8><<<xcd_grid(D->L, D->fwd_grid), ADJ_THREADS, 0, ctx->stream>>>(
          D->gdev, D->d_row_off, D->Fx, D->No, gx_old, D->r2, D->L, D->fwd_grid);
    else
      k_data_forward_tiled<16><<<xcd_grid(D->L, D->fwd_grid), ADJ_THREADS, 0, ctx->stream>>>(
          D->gdev, D->d_row_off, D->Fx, D->No, gx_old, D->r2, D->L, D->fwd_grid);
    check_launch();
  }
  if (D->any_rows) {
    k_data_forward<<<row_grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, gx_old, D->r2, D->L,
                                                      row_units);
    check_launch();
  }
}


// Block residuals from the FB parts (fb_res, S parts per block, added in
// part order; NULL when the residuals are already per block), then their
// sum and that of the squared radii into out2 (may be NULL).  A NaN residual
// sum reports a failed FB workgroup exchange (*fb_err set).
__global__ void k_sum_blocks(double* __restrict__ norm_res, const double* __restrict__ fb_res, int S,
                             const double* __restrict__ eps, int nb, const unsigned* __restrict__ fb_err,
                             double* __restrict__ out2) {
  if (fb_res)
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      double r = 0.0;
      for (int p = 0; p < S; ++p) r += fb_res[(size_t)b * FB_SPLIT_MAX + p];
      norm_res[b] = r;
    }
  __syncthreads();
  if (out2 && threadIdx.x == 0) {
    double s = 0.0, e = 0.0;
    for (int i = 0; i < nb; ++i) {
      s += norm_res[i];
      e += eps[i] * eps[i];
    }
    out2[0] = *fb_err ? __builtin_nan("") : s;
    out2[1] = e;
  }
}

template <class T>
T* dev_copy(const T* h, size_t n) {
  T* d = nullptr;
  FHS_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) FHS_HIP(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

// Device-resident step: all inputs/outputs on the device; norms (squared) in
// data->d_norm_res, sums in d_out2 = {sum norm_res^2, sum eps^2}.
void data_step(fhs_ctx* ctx, fhs_data* D, float2* v2, const float2* y, const float* x,
               float2* gx_old, float2* proj, const float* pU, int max_it, int min_it, double tol,
               int dr, const float* lam, float* Ftx, double* d_out2) {
  const int L = D->L;
  nufft_forward(ctx, D->A, x, L, D->Fx);
  {
    Scope sc(ctx, "data_forward");
    const int units = (int)cdiv(D->maxM, 256);
    const unsigned grid = xcd_grid(L, units);
    if (dr && D->stencil)
      k_dr_stencil<DR_STATE><<<grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, gx_old,
                                                            D->r2, L, units);
    else if (dr)
      k_data_forward_dr<true><<<grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, gx_old,
                                                             D->r2, L, units);
    else
      launch_data_forward(ctx, D, gx_old, grid, units);
    check_launch();
  }
  {
    Scope sc(ctx, dr ? "l2_dual" : "fb_dual");
    if (dr) {
      const unsigned g2grid = (unsigned)(D->nblocks * D->l2_split);
      k_l2_partial<<<g2grid, L2_THREADS, 0, ctx->stream>>>(D->d_blk_start, D->l2_split, v2, y, D->r2, gx_old,
                                                           D->l2_part);
      check_launch();
      k_l2_apply<<<g2grid, L2_THREADS, 0, ctx->stream>>>(D->d_blk_start, D->l2_split, D->d_eps, v2, y, D->r2,
                                                         D->l2_part, D->d_norm_res);
    } else
      k_fb_dual<<<D->nblocks * D->fb_split, FB_THREADS, FB_LDS, ctx->stream>>>(
          D->d_blk_start, D->d_eps, v2, y, pU, D->r2, proj, gx_old, max_it, min_it, tol, D->fb_split,
          fb_sync(D));
    check_launch();
  }
  {
    Scope sc(ctx, "data_adjoint");
    launch_data_adjoint(ctx, D, v2, lam, dr);
  }
  nufft_adjoint(ctx, D->A, D->g2, L, D->d_sigma22, Ftx);
  if (d_out2 || !dr) {
    k_sum_blocks<<<1, 256, 0, ctx->stream>>>(D->d_norm_res, dr ? nullptr : fb_sync(D).res, D->fb_split, D->d_eps,
                                             D->nblocks, D->fb_sync + 2 * D->nblocks, d_out2);
    check_launch();
  }
}

void data_set_sigma22(fhs_ctx* ctx, fhs_data* D, const float* h_sigma22) {
  FHS_HIP(hipMemcpyAsync(D->d_sigma22, h_sigma22, D->L * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
}

void data_set_epsilon(fhs_ctx* ctx, fhs_data* D, const double* h_eps) {
  FHS_HIP(hipMemcpyAsync(D->d_eps, h_eps, D->nblocks * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
}

// gx_old = G Fx for the grid vectors already in D->Fx: the forward rows
// with a zero previous state, keeping gx
static void data_gx_from_grid(fhs_ctx* ctx, fhs_data* D, float2* gx_old, int dr, const float* lam) {
  FHS_HIP(hipMemsetAsync(gx_old, 0, D->Mtot * sizeof(float2), ctx->stream));
  const int units = (int)cdiv(D->maxM, 256);
  const unsigned grid = xcd_grid(D->L, units);
  if (dr && D->stencil)
    k_dr_stencil<DR_STATE><<<grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, gx_old, D->r2,
                                                          D->L, units);
  else if (dr)
    k_data_forward_dr<true><<<grid, 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, gx_old, D->r2,
                                                           D->L, units);
  else
    launch_data_forward(ctx, D, gx_old, grid, units);
  check_launch();
}

void data_init_gx(fhs_ctx* ctx, fhs_data* D, const float* x, float2* gx_old, int dr, const float* lam) {
  nufft_forward(ctx, D->A, x, D->L, D->Fx);
  data_gx_from_grid(ctx, D, gx_old, dr, lam);
}

// y = Lambda .* G A x (all channels)
void data_forward(fhs_ctx* ctx, fhs_data* D, const float* x, const float* lam, float2* y) {
  nufft_forward(ctx, D->A, x, D->L, D->Fx);
  Scope sc(ctx, "data_forward");
  const int units = (int)cdiv(D->maxM, 256);
  if (D->fwd_grid > 0) {
    if (D->kmax == 8)
      k_data_apply_tiled<8><<<xcd_grid(D->L, D->fwd_grid), ADJ_THREADS, 0, ctx->stream>>>(
          D->gdev, D->d_row_off, D->Fx, D->No, lam, y, D->L, D->fwd_grid);
    else
      k_data_apply_tiled<16><<<xcd_grid(D->L, D->fwd_grid), ADJ_THREADS, 0, ctx->stream>>>(
          D->gdev, D->d_row_off, D->Fx, D->No, lam, y, D->L, D->fwd_grid);
    check_launch();
  }
  if (D->any_rows)
    k_data_apply<<<xcd_grid(D->L, units), 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No, lam, y, D->L,
                                                                 units);
  check_launch();
}

// y = Lambda .* (G + G') A x (all channels, holographic G)
void data_forward_dr(fhs_ctx* ctx, fhs_data* D, const float* x, const float* lam, float2* y) {
  nufft_forward(ctx, D->A, x, D->L, D->Fx);
  Scope sc(ctx, "data_forward");
  const int units = (int)cdiv(D->maxM, 256);
  if (D->stencil)
    k_dr_stencil<DR_APPLY><<<xcd_grid(D->L, units), 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx, D->No,
                                                                           lam, y, nullptr, D->L, units);
  else
    k_data_forward_dr<false><<<xcd_grid(D->L, units), 256, 0, ctx->stream>>>(D->gdev, D->d_row_off, D->Fx,
                                                                            D->No, lam, y, nullptr, D->L, units);
  check_launch();
}

// out_l = scale_l real(At(G_l' (lam .* v)))  (+ G_l (lam .* v) for the DR form)
void data_adjoint(fhs_ctx* ctx, fhs_data* D, const float2* v, const float* lam, const float* scale_dev,
                  float* out, int dr = 0) {
  {
    Scope sc(ctx, "data_adjoint");
    launch_data_adjoint(ctx, D, v, lam, dr);
  }
  nufft_adjoint(ctx, D->A, D->g2, D->L, scale_dev, out);
}

// Reports a timed-out workgroup exchange of the ellipsoid projection (its
// error flag set by k_fb_dual): the counters of every block are then out of
// step, so they are cleared (outside any graph capture: the caller has just
// synchronised) before the error is raised, and the next call starts clean.
void data_check_fb(fhs_ctx* ctx, fhs_data* D) {
  unsigned err = 0;
  FHS_HIP(hipStreamSynchronize(ctx->stream));
  FHS_HIP(hipMemcpy(&err, D->fb_sync + 2 * D->nblocks, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (err) {
    FHS_HIP(hipMemset(D->fb_sync, 0, (2 * (size_t)D->nblocks + 1) * sizeof(unsigned)));
    FHS_HIP(hipDeviceSynchronize());
    throw Error(FHS_ERR_INTERNAL,
                "data step: a workgroup of the ellipsoid projection timed out waiting for its block "
                "(results of that iteration are invalid; the exchange counters were reset)");
  }
}

void data_get_norm_res(fhs_ctx* ctx, fhs_data* D, double* h_out) {
  std::vector<double> nr(D->nblocks);
  FHS_HIP(hipMemcpyAsync(nr.data(), D->d_norm_res, D->nblocks * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  data_check_fb(ctx, D);
  for (int b = 0; b < D->nblocks; ++b) h_out[b] = std::sqrt(nr[b]);
}

// Test hook: sets the timed-out flag from the host, as a real time-out would
// (no kernel waits on the flag; the next residual sum reads it).
void data_inject_fb_timeout(fhs_ctx* ctx, fhs_data* D) {
  const unsigned one = 1;
  FHS_HIP(hipStreamSynchronize(ctx->stream));
  FHS_HIP(hipMemcpy(D->fb_sync + 2 * D->nblocks, &one, sizeof(unsigned), hipMemcpyHostToDevice));
}

int data_nblocks(const fhs_data* D) { return D->nblocks; }
bool data_square(const fhs_data* D) { return D->square; }
int data_channels(const fhs_data* D) { return D->L; }

}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_data_create(fhs_ctx* ctx, const fhs_nufft* A, int L, const fhs_gmat* const* G,
                    int nblocks, const int64_t* block_start, const int* block_channel,
                    fhs_data** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(A && G && L >= 1 && L <= 65535 && nblocks >= L, "data_create: bad argument");
    FHS_HIP(hipSetDevice(ctx->device));
    fhs_data* D = new fhs_data();
    D->device = ctx->device;
    D->L = L;
    D->A = A;
    D->nblocks = nblocks;
    D->No = (int64_t)A->Kty * A->Ktx;
    D->row_off.assign(L + 1, 0);
    std::vector<GDev> gd(L);
    bool square = true;
    for (int l = 0; l < L; ++l) {
      FHS_CHECK(G[l] != nullptr, "data_create: null G");
      FHS_CHECK(G[l]->ncols <= D->No, "data_create: G has more columns than the FFT grid");
      gd[l] = G[l]->d;
      D->row_off[l + 1] = D->row_off[l] + G[l]->M;
      D->maxM = std::max<int>(D->maxM, (int)G[l]->M);
      if (G[l]->d.tiled) {
        FHS_CHECK(D->kmax == 0 || D->kmax == G[l]->d.kmax,
                  "data_create: the channels' separable G must share their row width (8 or 16 slots)");
        D->kmax = G[l]->d.kmax;
        D->tile_grid = std::max<int>(D->tile_grid, G[l]->d.nunits);
        D->red_grid = std::max<int>(D->red_grid, G[l]->d.nred);
        D->tile_lds = std::max(D->tile_lds, tile_lds_bytes(G[l]->d.Ky, G[l]->d.Kx));
        D->fwd_grid = std::max<int>(D->fwd_grid, G[l]->d.nfitems);
      } else {
        D->any_rows = true;
        const int64_t tail = (D->No - G[l]->ncols + ADJ_COLS - 1) / ADJ_COLS;
        D->adj_grid = std::max<int>(D->adj_grid, G[l]->d.nunits + (int)tail);
      }
      square = square && (G[l]->ncols == G[l]->M) && !G[l]->d.tiled;   // H of the DR form is not tiled
    }
    D->square = square;
    bool sten = square;
    for (int l = 0; l < L; ++l) sten = sten && G[l]->d.stencil != nullptr && G[l]->M == D->No;
    D->stencil = sten;
    D->Mtot = D->row_off[L];
    D->blk_start.assign(block_start, block_start + nblocks + 1);
    D->blk_chan.assign(block_channel, block_channel + nblocks);
    FHS_CHECK(D->blk_start[0] == 0 && D->blk_start[nblocks] == D->Mtot,
              "data_create: blocks must cover all rows of all channels");
    for (int b = 0; b < nblocks; ++b) {
      const int c = D->blk_chan[b];
      FHS_CHECK(c >= 0 && c < L, "data_create: block channel out of range");
      FHS_CHECK(D->blk_start[b] < D->blk_start[b + 1], "data_create: empty block");
      FHS_CHECK(D->blk_start[b] >= D->row_off[c] && D->blk_start[b + 1] <= D->row_off[c + 1],
                "data_create: a block straddles channels");
    }
    D->gdev = dev_copy(gd.data(), L);
    D->gd = gd;
    for (int l = 0; l < L; ++l) D->rows32.push_back(gd[l].vals);
    D->rows16.assign(L, nullptr);
    D->d_row_off = dev_copy(D->row_off.data(), L + 1);
    D->d_blk_start = dev_copy(D->blk_start.data(), nblocks + 1);
    FHS_HIP(hipMalloc(&D->d_eps, nblocks * sizeof(double)));
    FHS_HIP(hipMemset(D->d_eps, 0, nblocks * sizeof(double)));
    FHS_HIP(hipMalloc(&D->d_norm_res, nblocks * sizeof(double)));
    FHS_HIP(hipMalloc(&D->d_sigma22, L * sizeof(float)));
    FHS_HIP(hipMalloc(&D->Fx, (size_t)L * D->No * sizeof(float2)));
    FHS_HIP(hipMalloc(&D->g2, (size_t)L * D->No * sizeof(float2)));
    FHS_HIP(hipMalloc(&D->r2, (size_t)D->Mtot * sizeof(float2)));
    FHS_HIP(hipMalloc(&D->fb_sync, (2 * (size_t)nblocks + 1) * sizeof(unsigned)));
    FHS_HIP(hipMemset(D->fb_sync, 0, (2 * (size_t)nblocks + 1) * sizeof(unsigned)));
    FHS_HIP(hipMalloc(&D->fb_part, (size_t)nblocks * (2 * FB_SPLIT_MAX * FB_PART + FB_SPLIT_MAX) * sizeof(double)));
    D->fb_split = fb_split_for(D, ctx->device);
    // rows the DR kernels skip (inactive modes) must read as zero
    FHS_HIP(hipMemset(D->r2, 0, (size_t)D->Mtot * sizeof(float2)));
    if (square) {
      int64_t longest = 0;
      for (int b = 0; b < nblocks; ++b) longest = std::max(longest, D->blk_start[b + 1] - D->blk_start[b]);
      D->l2_split = (int)std::min<int64_t>(L2_SPLIT_MAX, std::max<int64_t>(1, longest / (L2_THREADS * 16)));
      FHS_HIP(hipMalloc(&D->l2_part, (size_t)nblocks * L2_SPLIT_MAX * 2 * sizeof(double)));
    }
    *out = D;
  });
}

namespace fhs {
namespace {
// FP32 complex slots -> FP16 (re, im) pairs, round to nearest
__global__ __launch_bounds__(256) void k_rows_to_half(const float2* __restrict__ in, __half2* __restrict__ out,
                                                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = __float22half2_rn(in[i]);
}
}  // namespace
}  // namespace fhs

int fhs_data_set_row_precision(fhs_ctx* ctx, fhs_data* D, int bits) {
  if (!ctx || !D) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(bits == 16 || bits == 32, "data_set_row_precision: bits must be 16 or 32");
    FHS_HIP(hipSetDevice(D->device));
    if (bits == 16)   // every channel is checked before any is changed
      for (int l = 0; l < D->L; ++l)
        if (!D->gd[l].tiled || D->gd[l].layout != G_SEP)
          throw Error(FHS_ERR_UNSUPPORTED, "data_set_row_precision: FP16 rows need a tile-gridded separable G");
    for (int l = 0; l < D->L; ++l) {
      GDev& g = D->gd[l];
      if (bits == 16) {
        if (!D->rows16[l]) {
          const int64_t n = (int64_t)g.M * 2 * g.kmax;
          FHS_HIP(hipMalloc(&D->rows16[l], n * sizeof(__half2)));
          k_rows_to_half<<<cdiv(n, 256), 256, 0, ctx->stream>>>(D->rows32[l], (__half2*)D->rows16[l], n);
          check_launch();
        }
        g.vals = (const float2*)D->rows16[l];
        g.half = 1;
      } else {
        g.vals = D->rows32[l];
        g.half = 0;
      }
    }
    FHS_HIP(hipStreamSynchronize(ctx->stream));
    FHS_HIP(hipMemcpy(D->gdev, D->gd.data(), D->L * sizeof(GDev), hipMemcpyHostToDevice));
    if (bits == 32)
      for (void*& p : D->rows16)
        if (p) {
          FHS_HIP(hipFree(p));
          p = nullptr;
        }
  });
}

int fhs_data_destroy(fhs_data* D) {
  if (!D) return FHS_OK;
  (void)hipSetDevice(D->device);
  for (void* p : D->rows16)
    if (p) (void)hipFree(p);
  void* ps[] = {D->gdev, D->d_row_off, D->d_blk_start, D->d_eps, D->d_norm_res,
                D->d_sigma22, D->Fx, D->g2, D->r2, D->fb_sync, D->fb_part, D->l2_part};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete D;
  return FHS_OK;
}

int fhs_data_forward(fhs_ctx* ctx, fhs_data* D, const float* x, const float* Lambda, void* y) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && x && y, "data_forward: null argument");
    data_forward(ctx, D, x, Lambda, (float2*)y);
  });
}

int fhs_data_adjoint(fhs_ctx* ctx, fhs_data* D, const void* y, const float* Lambda, const float* scale,
                     float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && y && out, "data_adjoint: null argument");
    float* sd = nullptr;
    if (scale) {
      sd = (float*)ctx->ws.get("data_adj_scale", D->L * sizeof(float));
      FHS_HIP(hipMemcpyAsync(sd, scale, D->L * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    }
    data_adjoint(ctx, D, (const float2*)y, Lambda, sd, out);
  });
}

int fhs_data_forward_dr(fhs_ctx* ctx, fhs_data* D, const float* x, const float* Lambda, void* y) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && x && y, "data_forward_dr: null argument");
    FHS_CHECK(D->square, "data_forward_dr: needs square (holographic) G");
    data_forward_dr(ctx, D, x, Lambda, (float2*)y);
  });
}

int fhs_data_adjoint_dr(fhs_ctx* ctx, fhs_data* D, const void* y, const float* Lambda, const float* scale,
                        float* out) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && y && out, "data_adjoint_dr: null argument");
    FHS_CHECK(D->square, "data_adjoint_dr: needs square (holographic) G");
    float* sd = nullptr;
    if (scale) {
      sd = (float*)ctx->ws.get("data_adj_scale", D->L * sizeof(float));
      FHS_HIP(hipMemcpyAsync(sd, scale, D->L * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    }
    data_adjoint(ctx, D, (const float2*)y, Lambda, sd, out, 1);
  });
}

int fhs_data_init_fx_old(fhs_ctx* ctx, fhs_data* D, const float* x, const float* Lambda,
                         int flag_dimensionality_reduction, void* Gx_old) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && x && Gx_old, "data_init_fx_old: bad argument");
    FHS_CHECK(!flag_dimensionality_reduction || D->square, "data_init_fx_old: DR needs square G");
    data_init_gx(ctx, D, x, (float2*)Gx_old, flag_dimensionality_reduction, Lambda);
  });
}

int fhs_data_info(const fhs_data* D, int* Ny, int* Nx, int* L, int64_t* No, int64_t* Mtot, int* nblocks) {
  if (!D) return FHS_ERR_ARG;
  if (Ny) *Ny = D->A->Ny;
  if (Nx) *Nx = D->A->Nx;
  if (L) *L = D->L;
  if (No) *No = D->No;
  if (Mtot) *Mtot = D->Mtot;
  if (nblocks) *nblocks = D->nblocks;
  return FHS_OK;
}

int fhs_data_plan(const fhs_data* D, const fhs_nufft** plan) {
  if (!D || !plan) return FHS_ERR_ARG;
  *plan = D->A;
  return FHS_OK;
}

int fhs_data_fx_old_to_state(fhs_ctx* ctx, fhs_data* D, const void* Fx_old, const float* Lambda,
                             int flag_dimensionality_reduction, void* Gx_old) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && Fx_old && Gx_old, "data_fx_old_to_state: bad argument");
    FHS_CHECK(!flag_dimensionality_reduction || D->square, "data_fx_old_to_state: DR needs square G");
    FHS_HIP(hipMemcpyAsync(D->Fx, Fx_old, (size_t)D->L * D->No * sizeof(float2), hipMemcpyDeviceToDevice,
                           ctx->stream));
    data_gx_from_grid(ctx, D, (float2*)Gx_old, flag_dimensionality_reduction, Lambda);
  });
}

int fhs_update_dual_data_fidelity(fhs_ctx* ctx, fhs_data* D, void* v2, const void* y,
                                  const float* x, void* Gx_old, void* proj, const float* pU,
                                  const double* epsilon, int elipse_proj_max_iter,
                                  int elipse_proj_min_iter, double elipse_proj_eps,
                                  const float* sigma22, int flag_dimensionality_reduction,
                                  const float* Lambda, float* Ftx, double* norm_res,
                                  double* global_norm_res, double* norm_epsilon) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_CHECK(D && v2 && y && x && Gx_old && Ftx && epsilon && sigma22,
              "update_dual_data_fidelity: null argument");
    const int dr = flag_dimensionality_reduction != 0;
    FHS_CHECK(dr || proj, "update_dual_data_fidelity: proj required without DR");
    FHS_CHECK(!dr || D->square, "update_dual_data_fidelity: DR needs square (holographic) G");
    FHS_CHECK(elipse_proj_max_iter >= 1 && elipse_proj_min_iter >= 0,
              "update_dual_data_fidelity: bad ellipsoid iteration bounds");
    data_set_epsilon(ctx, D, epsilon);
    data_set_sigma22(ctx, D, sigma22);
    double* d2 = (double*)ctx->ws.get("data_out2", 2 * sizeof(double));
    data_step(ctx, D, (float2*)v2, (const float2*)y, x, (float2*)Gx_old, (float2*)proj, pU,
              elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps, dr, Lambda, Ftx, d2);
    std::vector<double> nr(D->nblocks);
    double h2[2];
    FHS_HIP(hipMemcpyAsync(nr.data(), D->d_norm_res, D->nblocks * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    FHS_HIP(hipMemcpyAsync(h2, d2, sizeof(h2), hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
    if (norm_res)
      for (int b = 0; b < D->nblocks; ++b) norm_res[b] = std::sqrt(nr[b]);
    if (global_norm_res) *global_norm_res = h2[0];
    if (norm_epsilon) *norm_epsilon = h2[1];
  });
}

}  // extern "C"
