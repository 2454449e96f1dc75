/*
 * libfhs: MI355X (gfx950) kernels of the Faceted HyperSARA primal-dual
 * (PDFB) iteration behind a plain C ABI.
 *
 * Conventions
 *  - Every function returns an int status (FHS_OK = 0); on failure
 *    fhs_last_error(ctx) holds a message.  No exception crosses the ABI.
 *  - Array arguments of compute calls are DEVICE pointers (hipMalloc or any
 *    allocator of the same device); set-up calls named *_host take host
 *    pointers.  Outputs are caller-allocated.
 *  - Layout is MATLAB's: column-major, complex values interleaved
 *    (re, im) in float32 ("complex64").  An image cube Ny x Nx x L is
 *    L contiguous column-major Ny x Nx slices.
 *  - Reductions (norms, residuals, l2-ball radii) are accumulated in FP64;
 *    the operators (G, FFT, wavelets) compute in FP32.
 *  - All work of a context is issued on its stream (default: the null
 *    stream); calls returning host scalars synchronise that stream.
 *
 * Each entry point cites the reference interface it replaces
 * (basp-group/Faceted-HyperSARA, path:line).
 */
#ifndef FHS_H
#define FHS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FHS_OK 0
#define FHS_ERR_ARG 1
#define FHS_ERR_HIP 2
#define FHS_ERR_INTERNAL 3
#define FHS_ERR_UNSUPPORTED 4

typedef struct fhs_ctx fhs_ctx;
typedef struct fhs_gmat fhs_gmat;     /* device interpolation matrix G (one channel) */
typedef struct fhs_nufft fhs_nufft;   /* oversampled scaled FFT plan (A / At) */
typedef struct fhs_sara fhs_sara;     /* SARA dictionary plan (Psit / Psi) */
typedef struct fhs_data fhs_data;     /* data-fidelity state of L channels */

/* ------------------------------------------------------------ context -- */
int fhs_ctx_create(int device, fhs_ctx** out);
int fhs_ctx_destroy(fhs_ctx* ctx);
const char* fhs_last_error(const fhs_ctx* ctx);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL = null stream. */
int fhs_ctx_set_stream(fhs_ctx* ctx, void* stream);
int fhs_ctx_synchronize(fhs_ctx* ctx);
/* Live per-kernel timing with HIP events on the context stream. */
int fhs_ctx_set_timing(fhs_ctx* ctx, int enable);
/* Copies up to cap entries: names as "name\n" concatenated into names_buf
 * (size names_cap), total milliseconds and launch counts. */
int fhs_ctx_timing_report(fhs_ctx* ctx, char* names_buf, int names_cap,
                          double* total_ms, long* counts, int cap, int* n_out);
int fhs_ctx_timing_reset(fhs_ctx* ctx);

/* ------------------------------------------------ interpolation matrix -- */
/* G from MATLAB-style CSR rows (rowptr[M+1], colidx[nnz], vals[2*nnz]
 * interleaved complex; host pointers).  The sparse G{i}{j} of the
 * reference (lib/operators/util_gen_measurement_operator_synth.m:85, CSC in
 * MATLAB) is passed row-compressed: a CSC of G' is a CSR of G.
 * Stored on device transposed (ELL, coalesced forward gather) and as CSC
 * (atomic-free adjoint gather). */
int fhs_gmat_create_csr_host(fhs_ctx* ctx, int64_t M, int64_t ncols,
                             const int64_t* rowptr, const int64_t* colidx,
                             const float* vals, fhs_gmat** out);
/* G of the min-max NUFFT with a dense Ky x Kx window per row anchored at
 * (ky0, kx0) (wrapping on the Kty x Ktx grid): anchor[m] = ky0 + Kty*kx0,
 * coef[m*Ky*Kx + jx*Ky + jy] (interleaved complex, host pointers). */
int fhs_gmat_create_window_host(fhs_ctx* ctx, int64_t M, int Kty, int Ktx,
                                int Ky, int Kx, const int32_t* anchor,
                                const float* coef, fhs_gmat** out);
int fhs_gmat_destroy(fhs_gmat* g);
int fhs_gmat_info(const fhs_gmat* g, int64_t* M, int64_t* ncols, int64_t* nnz);

/* y = Lambda .* (G * Fx)     src/apply_direct_operator.m:1-22
 * Fx: complex[ncols]; Lambda: float[M] or NULL (= 1); y: complex[M]. */
int fhs_apply_direct_operator(fhs_ctx* ctx, const fhs_gmat* G, const void* Fx,
                              const float* Lambda, void* y);
/* Fx = G' * (Lambda .* y)    src/apply_adjoint_operator.m:1-23 */
int fhs_apply_adjoint_operator(fhs_ctx* ctx, const fhs_gmat* G, const void* y,
                               const float* Lambda, void* Fx);

/* ------------------------------------------------- scaled Fourier (A) -- */
/* Plan for A = fft2(pad(sn .* x)) with sn = sy * sx' (real scaling factors
 * of the NUFFT, host arrays sy[Ny], sx[Nx]); grid Kty x Ktx, powers of two
 * in [16, 4096], Kty >= Ny, Ktx >= Nx. */
int fhs_nufft_create_host(fhs_ctx* ctx, int Ny, int Nx, int Kty, int Ktx,
                          const double* sy, const double* sx, fhs_nufft** out);
int fhs_nufft_destroy(fhs_nufft* p);
/* Fx(:, l) = A(x(:, :, l)), l = 1..L
 * src/apply_scaled_fourier_transform.m:21-24.  x: float[L][Nx][Ny];
 * Fx: complex[L][Kty*Ktx]. */
int fhs_apply_scaled_fourier_transform(fhs_ctx* ctx, const fhs_nufft* p,
                                       const float* x, int L, void* Fx);
/* out(:, :, l) = scale[l] * real(At(g(:, l)))  (the form used at
 * src/update_dual_data_fidelity.m:145).  scale: host float[L] or NULL. */
int fhs_apply_scaled_fourier_adjoint(fhs_ctx* ctx, const fhs_nufft* p,
                                     const void* g, int L, const float* scale,
                                     float* out);

/* ---------------------------------------------------- l2 ball / duals -- */
/* p = y + (x - y) * min(eps / ||x - y||, 1)      src/proj_l2ball.m:20-22
 * x, y, p: complex[n] (p may alias x). */
int fhs_proj_l2ball(fhs_ctx* ctx, const void* x, double eps, const void* y,
                    int64_t n, void* p);

/* ------------------------------------------------------------- primal -- */
/* xsol <- max(xsol - (g_a + g_b + g_c), 0); xhat = 2 xsol - xsol_prev;
 * rel_x = sum (xsol_prev - xsol)^2, norm_x = sum xsol^2.
 * src/fhs/fhs_update_primal.m:33-38 (g_b, g_c may be NULL). */
int fhs_update_primal(fhs_ctx* ctx, float* xsol, const float* g_a,
                      const float* g_b, const float* g_c, float* xhat,
                      int64_t n, double* rel_x, double* norm_x);

#ifdef __cplusplus
}
#endif
#endif /* FHS_H */
