// Host-side entry points: adaptive epsilon, device memory helpers for FFI
// callers without an allocator of their own (MEX, ctypes), and G from a
// MATLAB (CSC) sparse matrix.
#include <vector>

#include "common.hpp"

using namespace fhs;

extern "C" {

int fhs_update_epsilon(double* epsilon, int t, int* t_block, const double* norm_res, int n,
                       double tol_in, double tol_out, int steps, double change, int* n_updated) {
  if (!epsilon || !t_block || !norm_res || n < 0) return FHS_ERR_ARG;
  int upd = 0;
  for (int b = 0; b < n; ++b) {
    if (t > t_block[b] + steps) {
      if (norm_res[b] < tol_in * epsilon[b]) {
        epsilon[b] = change * norm_res[b] + (1.0 - change) * epsilon[b];
        t_block[b] = t;
        ++upd;
      }
      if (norm_res[b] > tol_out * epsilon[b]) {
        epsilon[b] = change * norm_res[b] + (1.0 - change) * epsilon[b];
        t_block[b] = t;
        ++upd;
      }
    }
  }
  if (n_updated) *n_updated = upd;
  return FHS_OK;
}

int fhs_device_alloc(fhs_ctx* ctx, int64_t bytes, void** ptr) {
  if (!ctx || !ptr) return FHS_ERR_ARG;
  *ptr = nullptr;
  return guard(ctx, [&] {
    FHS_CHECK(bytes >= 0, "device_alloc: negative size");
    FHS_HIP(hipSetDevice(ctx->device));
    FHS_HIP(hipMalloc(ptr, bytes > 0 ? bytes : 1));
  });
}

int fhs_device_free(fhs_ctx* ctx, void* ptr) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] { FHS_HIP(hipFree(ptr)); });
}

int fhs_memcpy_h2d(fhs_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int fhs_memcpy_d2h(fhs_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx) return FHS_ERR_ARG;
  return guard(ctx, [&] {
    FHS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FHS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int fhs_gmat_create_csr_host(fhs_ctx* ctx, int64_t M, int64_t ncols, const int64_t* rowptr,
                             const int64_t* colidx, const float* vals, fhs_gmat** out);

int fhs_gmat_create_csc_host(fhs_ctx* ctx, int64_t M, int64_t ncols, const int64_t* colptr,
                             const int64_t* rowidx, const float* vals, fhs_gmat** out) {
  if (!ctx || !out || !colptr || !rowidx || !vals) return FHS_ERR_ARG;
  int64_t nnz = colptr[ncols];
  std::vector<int64_t> rowptr(M + 1, 0), col(nnz);
  std::vector<float> v(2 * nnz);
  int st = guard(ctx, [&] {
    FHS_CHECK(M > 0 && ncols > 0 && nnz >= 0, "gmat_create_csc: bad shape");
    for (int64_t e = 0; e < nnz; ++e) {
      FHS_CHECK(rowidx[e] >= 0 && rowidx[e] < M, "gmat_create_csc: row index out of range");
      rowptr[rowidx[e] + 1]++;
    }
    for (int64_t m = 0; m < M; ++m) rowptr[m + 1] += rowptr[m];
    std::vector<int64_t> pos(rowptr.begin(), rowptr.end() - 1);
    for (int64_t k = 0; k < ncols; ++k) {
      for (int64_t e = colptr[k]; e < colptr[k + 1]; ++e) {
        const int64_t p = pos[rowidx[e]]++;
        col[p] = k;
        v[2 * p] = vals[2 * e];
        v[2 * p + 1] = vals[2 * e + 1];
      }
    }
  });
  if (st != FHS_OK) return st;
  return fhs_gmat_create_csr_host(ctx, M, ncols, rowptr.data(), col.data(), v.data(), out);
}

}  // extern "C"
