// PDFB iteration engine shared by SARA, HyperSARA and Faceted HyperSARA:
// one call runs whole iterations on the GPU
// (src/fhs/facetHyperSARA.m:701-783, src/hs/hyperSARA.m:565-617,
//  src/sara/sara.m:521-633):
//   1. xsol <- max(xsol - g0 - g1 - g2, 0), xhat = 2 xsol - xprev  (+ FP64 norms)
//   2. low-rank dual step (global or faceted)           -> g0 = sigma00 w v0
//   3. SARA l21 dual step                                -> g1 = sigma11 Psi v1
//   4. data-fidelity step (A, G, ellipsoid, G', At)     -> g2 = sigma22 Re At g2
// Per iteration four FP64 scalars come back: sum (dx)^2, sum x^2,
// sum norm_res^2, sum eps^2 (the reference's rel_x, norm_x,
// norm_residual_check^2, norm_epsilon_check^2).
//
// The engine owns a stream created with default flags and priority, so it
// orders with work on the null stream.  After one eager iteration has sized
// every workspace buffer, the kernel sequence of an iteration is captured in
// a hipGraph and replayed; parameters that change (thresholds after a
// reweighting, epsilon) live in device memory, so replays need no re-capture.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

namespace fhs {
void lowrank_dual_step(fhs_ctx* ctx, fhs_lowrank* R, float* v0, const float* xhat, float gscale, float* g0);
void lowrank_set_thresh(fhs_ctx* ctx, fhs_lowrank* R, const double* h_thresh);
void sara_dual_step(fhs_ctx* ctx, const fhs_sara* S, float* v1, const float* xhat, int L, const float* w,
                    double beta, float sigma1, float* g1);
int64_t sara_ncoefs(const fhs_sara* S);
void data_step(fhs_ctx* ctx, fhs_data* D, float2* v2, const float2* y, const float* x, float2* gx_old,
               float2* proj, const float* pU, int max_it, int min_it, double tol, int dr, const float* lam,
               float* Ftx, double* d_out2);
void data_set_sigma22(fhs_ctx* ctx, fhs_data* D, const float* h_sigma22);
void data_set_epsilon(fhs_ctx* ctx, fhs_data* D, const double* h_eps);
void data_get_norm_res(fhs_ctx* ctx, fhs_data* D, double* h_out);
void data_check_fb(fhs_ctx* ctx, fhs_data* D);
void data_inject_fb_timeout(fhs_ctx* ctx, fhs_data* D);
int data_nblocks(const fhs_data* D);
int data_channels(const fhs_data* D);
bool data_square(const fhs_data* D);
int lowrank_facets(const fhs_lowrank* R);
int lowrank_channels(const fhs_lowrank* R);
}  // namespace fhs

struct fhs_pdfb {
  fhs_ctx* ctx = nullptr;
  fhs_ctx own{};            // private context state bound to the engine stream
  fhs_ctx side{};           // second stream: the sparsity prior step runs beside the data step
  fhs_ctx low{};            // third stream: the low-rank prior step
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join_low = nullptr;
  bool concurrent = true;
  fhs_data* D = nullptr;
  fhs_sara* S = nullptr;
  fhs_lowrank* R = nullptr;
  fhs_pdfb_buffers b{};
  int L = 0, Ny = 0, Nx = 0;
  int64_t N = 0;
  double beta1 = 0.0;
  float sigma00 = 0.f, sigma11 = 0.f;
  int fb_max = 20, fb_min = 1, dr = 0;
  double fb_eps = 1e-8;
  double* d_scal = nullptr;   // [5]: the four stopping-rule sums, then sum (truth - x)^2
  struct Pub {                // coherent pinned host memory, written by k_publish
    double scal[5];
    unsigned long long seq;
  };
  Pub* h_pub = nullptr;
  Pub* d_pub = nullptr;       // its device address
  unsigned long long* d_seq = nullptr;   // iterations published (device counter)
  unsigned long long seen = 0;           // last sequence number read by the host
  const float* truth = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  bool params_set = false;
  long iters = 0;
};

namespace fhs {
namespace {

// Last node of an iteration: copies the five sums into coherent host memory
// and then advances the published sequence number (system-scope release),
// so the host learns that the iteration ended by polling host memory instead
// of a stream synchronisation and a separate copy.
__global__ void k_publish(const double* __restrict__ d_scal, unsigned long long* __restrict__ d_seq,
                          fhs_pdfb::Pub* pub) {
  if (threadIdx.x == 0) {
    const unsigned long long s = *d_seq + 1;
    *d_seq = s;
#pragma unroll
    for (int k = 0; k < 5; ++k) pub->scal[k] = d_scal[k];
    __threadfence_system();
    __hip_atomic_store(&pub->seq, s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// After the primal update the two prior steps (reading xhat) and the data
// step (reading x) are independent: with `concurrent` the sparsity step is
// forked onto the side stream and the low-rank step onto a third one (its
// one-wavefront eigen-solver is latency, not work), and both are joined
// before the next iteration's primal update.
void run_iteration(fhs_pdfb* P) {
  fhs_ctx* c = &P->own;
  fhs_ctx* p = P->concurrent ? &P->side : c;
  fhs_ctx* r = P->concurrent ? &P->low : c;
  const auto& b = P->b;
  primal_update(c, b.x, b.g1, b.g2, b.g0, b.xhat, P->N * P->L, P->d_scal, P->truth, P->d_scal + 4);
  if (P->concurrent) {
    FHS_HIP(hipEventRecord(P->ev_fork, c->stream));
    FHS_HIP(hipStreamWaitEvent(p->stream, P->ev_fork, 0));
    if (P->R) FHS_HIP(hipStreamWaitEvent(r->stream, P->ev_fork, 0));
  }
  if (P->R) lowrank_dual_step(r, P->R, b.v0, b.xhat, P->sigma00, b.g0);
  sara_dual_step(p, P->S, b.v1, b.xhat, P->L, b.weights1, P->beta1, P->sigma11, b.g1);
  data_step(c, P->D, (float2*)b.v2, (const float2*)b.y, b.x, (float2*)b.Gx_old, (float2*)b.proj, b.pU,
            P->fb_max, P->fb_min, P->fb_eps, P->dr, b.Lambda, b.g2, P->d_scal + 2);
  if (P->concurrent) {
    FHS_HIP(hipEventRecord(P->ev_join, p->stream));
    FHS_HIP(hipStreamWaitEvent(c->stream, P->ev_join, 0));
    if (P->R) {
      FHS_HIP(hipEventRecord(P->ev_join_low, r->stream));
      FHS_HIP(hipStreamWaitEvent(c->stream, P->ev_join_low, 0));
    }
  }
  k_publish<<<1, 64, 0, c->stream>>>(P->d_scal, P->d_seq, P->d_pub);
  check_launch();
}

// Waits until iteration `want` is published.  Polls the host word; every
// 1024 polls the stream is queried, so a failed launch raises its error and
// an idle stream without the publication is reported instead of waited for.
void wait_published(fhs_pdfb* P, unsigned long long want) {
  for (unsigned long spin = 1;; ++spin) {
    if (__atomic_load_n(&P->h_pub->seq, __ATOMIC_ACQUIRE) >= want) return;
    if ((spin & 1023) == 0) {
      const hipError_t q = hipStreamQuery(P->own.stream);
      if (q == hipSuccess) {
        if (__atomic_load_n(&P->h_pub->seq, __ATOMIC_ACQUIRE) >= want) return;
        throw Error(FHS_ERR_HIP, "pdfb_iterate: the stream is idle but the iteration was not published");
      }
      if (q != hipErrorNotReady) FHS_HIP(q);
    }
  }
}

void merge_timing(fhs_ctx* from, fhs_ctx* into) {
  from->collect();
  for (auto& kv : from->totals) {
    auto& t = into->totals[kv.first];
    t.first += kv.second.first;
    t.second += kv.second.second;
  }
  from->totals.clear();
}

void drop_graph(fhs_pdfb* P) {
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  P->exec = nullptr;
  P->graph = nullptr;
}

}  // namespace
}  // namespace fhs

using namespace fhs;

extern "C" {

int fhs_pdfb_create(fhs_ctx* ctx, fhs_data* D, fhs_sara* S, fhs_lowrank* R,
                    const fhs_pdfb_buffers* bufs, fhs_pdfb** out) {
  if (!ctx || !out) return FHS_ERR_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(D && S && bufs, "pdfb_create: data and sparsity prior are required");
    const fhs_pdfb_buffers& b = *bufs;
    FHS_CHECK(b.x && b.xhat && b.g1 && b.g2 && b.v1 && b.v2 && b.Gx_old && b.y,
              "pdfb_create: missing buffer");
    FHS_CHECK(!R || (b.g0 && b.v0), "pdfb_create: the low-rank prior needs g0 and v0");
    fhs_pdfb* P = new fhs_pdfb();
    P->ctx = ctx;
    P->own.device = ctx->device;
    FHS_HIP(hipSetDevice(ctx->device));
    FHS_HIP(hipStreamCreate(&P->own.stream));
    FHS_HIP(hipStreamCreate(&P->side.stream));
    P->side.device = ctx->device;
    FHS_HIP(hipStreamCreate(&P->low.stream));
    P->low.device = ctx->device;
    FHS_HIP(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
    FHS_HIP(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
    FHS_HIP(hipEventCreateWithFlags(&P->ev_join_low, hipEventDisableTiming));
    P->D = D;
    P->S = S;
    P->R = R;
    P->b = b;
    P->L = data_channels(D);
    P->Ny = b.Ny;
    P->Nx = b.Nx;
    P->N = (int64_t)b.Ny * b.Nx;
    if (R) FHS_CHECK(lowrank_channels(R) == P->L, "pdfb_create: low-rank prior has a different L");
    FHS_HIP(hipMalloc(&P->d_scal, 5 * sizeof(double)));
    FHS_HIP(hipMemset(P->d_scal, 0, 5 * sizeof(double)));
    FHS_HIP(hipMalloc(&P->d_seq, sizeof(unsigned long long)));
    FHS_HIP(hipMemset(P->d_seq, 0, sizeof(unsigned long long)));
    FHS_HIP(hipHostMalloc((void**)&P->h_pub, sizeof(fhs_pdfb::Pub), hipHostMallocMapped | hipHostMallocCoherent));
    for (int k = 0; k < 5; ++k) P->h_pub->scal[k] = 0.0;
    P->h_pub->seq = 0;
    FHS_HIP(hipHostGetDevicePointer((void**)&P->d_pub, P->h_pub, 0));
    *out = P;
  });
}

int fhs_pdfb_destroy(fhs_pdfb* P) {
  if (!P) return FHS_OK;
  (void)hipSetDevice(P->own.device);
  (void)hipStreamSynchronize(P->own.stream);
  drop_graph(P);
  (void)hipFree(P->d_scal);
  (void)hipFree(P->d_seq);
  (void)hipHostFree(P->h_pub);
  (void)hipStreamSynchronize(P->side.stream);
  (void)hipStreamSynchronize(P->low.stream);
  P->own.ws.release();
  P->side.ws.release();
  P->low.ws.release();
  for (auto e : P->own.event_pool) (void)hipEventDestroy(e);
  for (auto e : P->side.event_pool) (void)hipEventDestroy(e);
  for (auto e : P->low.event_pool) (void)hipEventDestroy(e);
  (void)hipEventDestroy(P->ev_fork);
  (void)hipEventDestroy(P->ev_join);
  (void)hipEventDestroy(P->ev_join_low);
  (void)hipStreamDestroy(P->own.stream);
  (void)hipStreamDestroy(P->side.stream);
  (void)hipStreamDestroy(P->low.stream);
  delete P;
  return FHS_OK;
}

int fhs_pdfb_set_params(fhs_pdfb* P, const double* thresh0, double beta1, float sigma00, float sigma11,
                        const float* sigma22, const double* epsilon, int elipse_proj_max_iter,
                        int elipse_proj_min_iter, double elipse_proj_eps, int flag_dimensionality_reduction) {
  if (!P) return FHS_ERR_ARG;
  fhs_ctx* c = &P->own;
  int st = guard(P->ctx, [&] {
    FHS_CHECK(sigma22 && epsilon, "pdfb_set_params: sigma22 and epsilon are required");
    FHS_CHECK(!P->R || thresh0, "pdfb_set_params: thresh0 required with the low-rank prior");
    FHS_CHECK(elipse_proj_max_iter >= 1 && elipse_proj_min_iter >= 0, "pdfb_set_params: bad ellipsoid bounds");
    if (P->R) {
      const int Q = lowrank_facets(P->R);
      std::vector<double> th((size_t)Q * 16, 0.0);
      for (int q = 0; q < Q; ++q)
        for (int i = 0; i < P->L; ++i) th[(size_t)q * 16 + i] = thresh0[(size_t)q * P->L + i];
      lowrank_set_thresh(c, P->R, th.data());
    }
    FHS_CHECK(!flag_dimensionality_reduction || data_square(P->D),
              "pdfb_set_params: DR needs square (holographic) G");
    data_set_sigma22(c, P->D, sigma22);
    data_set_epsilon(c, P->D, epsilon);
    FHS_HIP(hipStreamSynchronize(c->stream));
    const bool structural = P->params_set &&
                            (P->beta1 != beta1 || P->sigma00 != sigma00 || P->sigma11 != sigma11 ||
                             P->fb_max != elipse_proj_max_iter || P->fb_min != elipse_proj_min_iter ||
                             P->fb_eps != elipse_proj_eps || P->dr != flag_dimensionality_reduction);
    if (structural) drop_graph(P);   // scalars baked into kernel arguments
    P->beta1 = beta1;
    P->sigma00 = sigma00;
    P->sigma11 = sigma11;
    P->fb_max = elipse_proj_max_iter;
    P->fb_min = elipse_proj_min_iter;
    P->fb_eps = elipse_proj_eps;
    P->dr = flag_dimensionality_reduction != 0;
    P->params_set = true;
  });
  return st;
}

int fhs_pdfb_set_graph(fhs_pdfb* P, int enable) {
  if (!P) return FHS_ERR_ARG;
  P->use_graph = enable != 0;
  if (!enable) drop_graph(P);
  return FHS_OK;
}

int fhs_pdfb_set_concurrent(fhs_pdfb* P, int enable) {
  if (!P) return FHS_ERR_ARG;
  if ((enable != 0) != P->concurrent) drop_graph(P);
  P->concurrent = enable != 0;
  return FHS_OK;
}

}  // extern "C"

namespace fhs {
namespace {

// One iteration: graph replay (captured on first use) or eager launches, then
// the wait for its published scalars.
void step(fhs_pdfb* P) {
  fhs_ctx* c = &P->own;
  const bool graph_now = P->use_graph && !c->timing && P->iters >= 1;
  if (graph_now && !P->exec) {
    FHS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    run_iteration(P);
    FHS_HIP(hipStreamEndCapture(c->stream, &P->graph));
    FHS_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
  }
  if (graph_now)
    FHS_HIP(hipGraphLaunch(P->exec, c->stream));
  else
    run_iteration(P);
  wait_published(P, ++P->seen);
  ++P->iters;
}

// A NaN residual sum is how a timed-out exchange of the ellipsoid projection
// (or non-finite data) shows in the published scalars: raise instead of
// letting the stopping rule read it as "not converged".
void check_published(fhs_pdfb* P) {
  const double* sc = P->h_pub->scal;
  if (std::isnan(sc[2]) || std::isnan(sc[3])) {
    data_check_fb(&P->own, P->D);   // raises (and resets the counters) on a time-out
    throw Error(FHS_ERR_INTERNAL, "pdfb: non-finite data residual (NaN in the data or the iterates)");
  }
}

void begin_calls(fhs_pdfb* P) {
  P->own.timing = P->ctx->timing;
  P->side.timing = P->ctx->timing;
  P->own.autotune = P->ctx->autotune;
  P->side.autotune = P->ctx->autotune;
  P->low.timing = P->ctx->timing;
}

void end_calls(fhs_pdfb* P) {
  if (P->own.timing) {
    merge_timing(&P->own, P->ctx);
    merge_timing(&P->side, P->ctx);
    merge_timing(&P->low, P->ctx);
  }
}

}  // namespace
}  // namespace fhs

extern "C" {

int fhs_pdfb_iterate(fhs_pdfb* P, int n_iter, double* scalars) {
  if (!P) return FHS_ERR_ARG;
  return guard(P->ctx, [&] {
    FHS_CHECK(P->params_set, "pdfb_iterate: call fhs_pdfb_set_params first");
    FHS_CHECK(n_iter >= 1, "pdfb_iterate: n_iter >= 1");
    begin_calls(P);
    for (int it = 0; it < n_iter; ++it) {
      step(P);
      check_published(P);
      if (scalars)
        for (int k = 0; k < 4; ++k) scalars[4 * it + k] = P->h_pub->scal[k];
    }
    end_calls(P);
  });
}

int fhs_pdfb_stopping_rule(int rule, int k, double rel_val, double norm_res, double norm_eps, int min_iter,
                           int max_iter, double rel_var, double fid_tol, double rel_var_low) {
  // comparisons with NaN are false, as in MATLAB
  if (!(k >= min_iter)) return 0;
  if (k >= max_iter) return 1;
  bool small = rel_val <= rel_var;
  if (rule == 1) small = small && norm_res <= fid_tol * norm_eps;
  return (small || rel_val <= rel_var_low) ? 1 : 0;
}

int fhs_pdfb_run(fhs_pdfb* P, int rule, int k0, int min_iter, int max_iter, double rel_var, double fid_tol,
                 double rel_var_low, int n_max, double err2_stop, double seconds_max, double* out, int* n_done,
                 int* converged) {
  if (!P || !out || !n_done || !converged) return FHS_ERR_ARG;
  return guard(P->ctx, [&] {
    FHS_CHECK(P->params_set, "pdfb_run: call fhs_pdfb_set_params first");
    FHS_CHECK(n_max >= 1 && (rule == 0 || rule == 1), "pdfb_run: bad argument");
    FHS_CHECK(err2_stop < 0.0 || P->truth, "pdfb_run: err2_stop needs fhs_pdfb_set_truth");
    begin_calls(P);
    const auto t0 = std::chrono::steady_clock::now();
    *n_done = 0;
    *converged = 0;
    for (int i = 0; i < n_max; ++i) {
      step(P);
      check_published(P);
      const double* sc = P->h_pub->scal;
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      for (int j = 0; j < 5; ++j) out[6 * i + j] = sc[j];
      out[6 * i + 5] = secs;
      *n_done = i + 1;
      const double rel = sc[1] > 0.0 ? std::sqrt(sc[0]) / std::sqrt(sc[1]) : std::nan("");
      if (fhs_pdfb_stopping_rule(rule, k0 + i + 1, rel, std::sqrt(sc[2]), std::sqrt(sc[3]), min_iter, max_iter,
                                 rel_var, fid_tol, rel_var_low)) {
        *converged = 1;
        break;
      }
      if (err2_stop >= 0.0 && sc[4] <= err2_stop) break;
      if (seconds_max > 0.0 && secs > seconds_max) break;
    }
    end_calls(P);
  });
}

int fhs_pdfb_set_truth(fhs_pdfb* P, const float* X0) {
  if (!P) return FHS_ERR_ARG;
  return guard(P->ctx, [&] {
    FHS_HIP(hipStreamSynchronize(P->own.stream));
    if (X0 != P->truth) drop_graph(P);   // the pointer is a kernel argument of the captured graph
    P->truth = X0;
  });
}

int fhs_pdfb_truth_error(fhs_pdfb* P, double* err2) {
  if (!P || !err2) return FHS_ERR_ARG;
  return guard(P->ctx, [&] {
    FHS_CHECK(P->truth, "pdfb_truth_error: no reference cube set (fhs_pdfb_set_truth)");
    FHS_CHECK(P->iters > 0, "pdfb_truth_error: no iteration has run");
    *err2 = P->h_pub->scal[4];
  });
}

int fhs_pdfb_norm_res(fhs_pdfb* P, double* norm_res) {
  if (!P || !norm_res) return FHS_ERR_ARG;
  return guard(P->ctx, [&] { data_get_norm_res(&P->own, P->D, norm_res); });
}

int fhs_pdfb_inject_fb_timeout(fhs_pdfb* P) {
  if (!P) return FHS_ERR_ARG;
  return guard(P->ctx, [&] { data_inject_fb_timeout(&P->own, P->D); });
}

int fhs_pdfb_stream(fhs_pdfb* P, void** stream) {
  if (!P || !stream) return FHS_ERR_ARG;
  *stream = (void*)P->own.stream;
  return FHS_OK;
}

}  // extern "C"
