// Context lifetime, error reporting and live kernel timing.
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace fhs {

// The autotuner's choices, per device and problem key (shared by every
// context of the process; the mutex covers contexts used from several host
// threads).
namespace {
std::mutex g_tune_mu;
std::map<std::pair<int, std::string>, std::pair<int, std::string>> g_tuned;
}  // namespace

int tuned_choice(int device, const std::string& key) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find({device, key});
  return it == g_tuned.end() ? -1 : it->second.first;
}

void set_tuned(int device, const std::string& key, int choice, const std::string& label) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[{device, key}] = {choice, label};
}

std::string tuned_report(int device) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::string r;
  for (const auto& kv : g_tuned)
    if (kv.first.first == device) r += kv.first.second + "=" + kv.second.second + "\n";
  return r;
}

}  // namespace fhs

hipEvent_t fhs_ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  FHS_HIP(hipEventCreate(&e));
  return e;
}

void fhs_ctx::begin(const char* name, fhs::KernelTimer& t) {
  t.name = name;
  t.start = take_event();
  t.stop = take_event();
  FHS_HIP(hipEventRecord(t.start, stream));
}

void fhs_ctx::end(fhs::KernelTimer& t) {
  if (hipEventRecord(t.stop, stream) == hipSuccess) pending.push_back(t);
}

void fhs_ctx::collect() {
  for (auto& t : pending) {
    FHS_HIP(hipEventSynchronize(t.stop));
    float ms = 0.f;
    FHS_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    auto& tot = totals[t.name];
    tot.first += ms;
    tot.second += 1;
    event_pool.push_back(t.start);
    event_pool.push_back(t.stop);
  }
  pending.clear();
}

extern "C" {

int fhs_ctx_create(int device, fhs_ctx** out) {
  if (!out) return FHS_ERR_ARG;
  *out = nullptr;
  fhs_ctx* ctx = new fhs_ctx();
  int st = fhs::guard(ctx, [&] {
    int n = 0;
    FHS_HIP(hipGetDeviceCount(&n));
    FHS_CHECK(device >= 0 && device < n, "fhs_ctx_create: device out of range");
    FHS_HIP(hipSetDevice(device));
    ctx->device = device;
    fhs::noalloc_register_device(device);
  });
  if (st != FHS_OK) {
    delete ctx;
    return st;
  }
  *out = ctx;
  return FHS_OK;
}

int fhs_ctx_destroy(fhs_ctx* ctx) {
  if (!ctx) return FHS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->pending) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->ws.release();
  delete ctx;
  return FHS_OK;
}

const char* fhs_last_error(const fhs_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int fhs_ctx_set_stream(fhs_ctx* ctx, void* stream) {
  if (!ctx) return FHS_ERR_ARG;
  ctx->stream = reinterpret_cast<hipStream_t>(stream);
  return FHS_OK;
}

int fhs_ctx_synchronize(fhs_ctx* ctx) {
  if (!ctx) return FHS_ERR_ARG;
  return fhs::guard(ctx, [&] { FHS_HIP(hipStreamSynchronize(ctx->stream)); });
}

int fhs_ctx_set_autotune(fhs_ctx* ctx, int enable) {
  if (!ctx) return FHS_ERR_ARG;
  ctx->autotune = enable != 0;
  return FHS_OK;
}

int fhs_autotune_report(fhs_ctx* ctx, char* buf, int cap, int* len) {
  if (!ctx) return FHS_ERR_ARG;
  return fhs::guard(ctx, [&] {
    const std::string r = fhs::tuned_report(ctx->device);
    if (len) *len = (int)r.size();
    if (buf && cap > 0) {
      const int n = std::min<int>(cap - 1, (int)r.size());
      std::memcpy(buf, r.data(), n);
      buf[n] = 0;
    }
  });
}

int fhs_ctx_set_timing(fhs_ctx* ctx, int enable) {
  if (!ctx) return FHS_ERR_ARG;
  return fhs::guard(ctx, [&] {
    if (!enable) ctx->collect();
    ctx->timing = enable != 0;
  });
}

int fhs_ctx_timing_report(fhs_ctx* ctx, char* names_buf, int names_cap,
                          double* total_ms, long* counts, int cap, int* n_out) {
  if (!ctx) return FHS_ERR_ARG;
  return fhs::guard(ctx, [&] {
    ctx->collect();
    int i = 0;
    std::string names;
    for (auto& kv : ctx->totals) {
      if (i >= cap) break;
      total_ms[i] = kv.second.first;
      counts[i] = kv.second.second;
      names += kv.first + "\n";
      ++i;
    }
    FHS_CHECK((int)names.size() < names_cap, "timing report: names buffer too small");
    std::memcpy(names_buf, names.c_str(), names.size() + 1);
    *n_out = i;
  });
}

int fhs_ctx_timing_reset(fhs_ctx* ctx) {
  if (!ctx) return FHS_ERR_ARG;
  return fhs::guard(ctx, [&] {
    ctx->collect();
    ctx->totals.clear();
  });
}

}  // extern "C"
