// Shared definitions of libfhs: context, status codes, error capture,
// device workspace, complex helpers and FP64 reductions for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fhs.h"

namespace fhs {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FHS_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::fhs::Error(FHS_ERR_HIP, std::string(#expr) + ": " +            \
                                          hipGetErrorString(e_));            \
  } while (0)

#define FHS_CHECK(cond, msg)                                                 \
  do {                                                                       \
    if (!(cond)) throw ::fhs::Error(FHS_ERR_ARG, msg);                       \
  } while (0)

// Grow-only device scratch owned by a context, keyed by purpose so that
// buffers needed at the same time never alias.
struct Workspace {
  std::map<std::string, std::pair<void*, size_t>> bufs;
  void* get(const std::string& key, size_t bytes) {
    auto it = bufs.find(key);
    if (it != bufs.end() && it->second.second >= bytes) return it->second.first;
    if (it != bufs.end()) {
      FHS_HIP(hipFree(it->second.first));
      bufs.erase(it);
    }
    void* p = nullptr;
    FHS_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
    bufs[key] = {p, bytes};
    return p;
  }
  void release() {
    for (auto& kv : bufs) (void)hipFree(kv.second.first);
    bufs.clear();
  }
};

struct KernelTimer {
  hipEvent_t start, stop;
  std::string name;
};

}  // namespace fhs

struct fhs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  fhs::Workspace ws;
  // live kernel timing (HIP events on the context stream)
  bool timing = false;
  bool autotune = false;               // time the launch-shape candidates of tunable kernels on first use
  std::vector<fhs::KernelTimer> pending;
  std::map<std::string, std::pair<double, long>> totals;  // ms, count
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  void begin(const char* name, fhs::KernelTimer& t);
  void end(fhs::KernelTimer& t);
  void collect();
};

namespace fhs {

// RAII helper: times the kernels launched in its scope when timing is on.
struct Scope {
  fhs_ctx* ctx;
  KernelTimer t;
  bool on;
  Scope(fhs_ctx* c, const char* name) : ctx(c), on(c->timing) {
    if (on) ctx->begin(name, t);
  }
  ~Scope() {
    if (on) ctx->end(t);
  }
};

// ---- autotuner ---------------------------------------------------------
// Launch shapes that change only how the work is spread over workgroups,
// never the arithmetic of an output (rows per workgroup of an FFT pass, rows
// staged per chunk of a gridding tile), are chosen per device and per
// problem shape: with autotuning on, the first call of a tunable kernel
// outside graph capture times every candidate three times with HIP events
// on the context stream (the kernels are pure functions of their inputs, so
// re-running them leaves the same outputs) and keeps the fastest; the choice
// is then reused by every context on that device.  With autotuning off, or
// under capture before a choice exists, the default candidate runs.
int tuned_choice(int device, const std::string& key);   // -1: not tuned yet
void set_tuned(int device, const std::string& key, int choice, const std::string& label);
std::string tuned_report(int device);

template <class F>
int tune_launch(fhs_ctx* ctx, const std::string& key, int ncand, int dflt, const char* const* labels, F&& launch) {
  if (ncand <= 1 || !ctx->autotune) {
    launch(dflt);
    return dflt;
  }
  const int known = tuned_choice(ctx->device, key);
  if (known >= 0) {
    launch(known);
    return known;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  FHS_HIP(hipStreamIsCapturing(ctx->stream, &cap));
  if (cap != hipStreamCaptureStatusNone) {
    launch(dflt);
    return dflt;
  }
  constexpr int ROUNDS = 3;
  std::vector<hipEvent_t> ev(2 * ROUNDS * ncand);
  for (auto& e : ev) FHS_HIP(hipEventCreate(&e));
  for (int r = 0; r < ROUNDS; ++r)
    for (int c = 0; c < ncand; ++c) {
      FHS_HIP(hipEventRecord(ev[2 * (r * ncand + c)], ctx->stream));
      launch(c);
      FHS_HIP(hipEventRecord(ev[2 * (r * ncand + c) + 1], ctx->stream));
    }
  FHS_HIP(hipEventSynchronize(ev.back()));
  int best = dflt;
  float best_ms = 0.f;
  for (int c = 0; c < ncand; ++c) {
    float m = 0.f;
    for (int r = 0; r < ROUNDS; ++r) {
      float t = 0.f;
      FHS_HIP(hipEventElapsedTime(&t, ev[2 * (r * ncand + c)], ev[2 * (r * ncand + c) + 1]));
      m = r == 0 ? t : std::min(m, t);
    }
    if (c == 0 || m < best_ms) {
      best_ms = m;
      best = c;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  set_tuned(ctx->device, key, best, labels ? labels[best] : std::to_string(best));
  return best;
}

// Runs body, converting C++ exceptions to a status code and a message.
template <class F>
int guard(fhs_ctx* ctx, F&& body) {
  try {
    body();
    return FHS_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return FHS_ERR_INTERNAL;
  }
}

inline void check_launch() { FHS_HIP(hipGetLastError()); }

// Lets `kernel` take `bytes` of dynamic LDS when that is above the 64 KiB
// default (a CU has 160 KiB).  Cheap to repeat; done before each launch
// that may need it.
inline void allow_dynamic_lds(const void* kernel, size_t bytes) {
  if (bytes > 65536) FHS_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------- device --
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // conj(a) * b
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum of NV doubles across a workgroup; every thread gets the totals.
// scratch: at least NV * (blockDim.x / 64) doubles of LDS.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) scratch[i * nw + wid] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += scratch[i * nw + w];
    v[i] = s;
  }
  __syncthreads();
}

}  // namespace fhs
