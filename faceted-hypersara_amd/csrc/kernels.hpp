// Internal (non-ABI) entry points shared between translation units.  All
// launch on ctx->stream and keep results on the device.
#pragma once

#include "common.hpp"

namespace fhs {

// xsol <- max(xsol - (ga+gb+gc), 0); xhat = 2 xsol - prev;
// d_out2[0] = sum (prev - xsol)^2, d_out2[1] = sum xsol^2; with truth
// (NULL = none) *d_truth = sum (truth - xsol)^2
void primal_update(fhs_ctx* ctx, float* x, const float* ga, const float* gb, const float* gc,
                   float* xhat, int64_t n, double* d_out2, const float* truth = nullptr,
                   double* d_truth = nullptr);

// Fx[l] = A(x[l]), l < L
void nufft_forward(fhs_ctx* ctx, const fhs_nufft* p, const float* x, int L, float2* Fx);
// out[l] = scale[l] * real(At(g[l])); scale_dev: device float[L] or null
void nufft_adjoint(fhs_ctx* ctx, const fhs_nufft* p, const float2* g, int L, const float* scale_dev,
                   float* out);

// Registers this device's address of the fhs_proj_l2ball_noalloc scratch
// (with `device` current; once per device, from fhs_ctx_create).
void noalloc_register_device(int device);

}  // namespace fhs
