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
typedef struct fhs_lowrank fhs_lowrank; /* facet layout of the nuclear-norm prior */
typedef struct fhs_sara_facets fhs_sara_facets; /* faceted SARA transform (per-facet Psit / Psi) */

/* ------------------------------------------------------------ context -- */
int fhs_ctx_create(int device, fhs_ctx** out);
int fhs_ctx_destroy(fhs_ctx* ctx);
const char* fhs_last_error(const fhs_ctx* ctx);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL = null stream. */
int fhs_ctx_set_stream(fhs_ctx* ctx, void* stream);
int fhs_ctx_synchronize(fhs_ctx* ctx);
/* Autotuning of launch shapes that never change the arithmetic of an output
 * (rows per workgroup of the FFT x-passes, rows staged per chunk of the
 * gridding tiles): with enable != 0, the first call of such a kernel outside
 * graph capture times every candidate with HIP events and keeps the fastest
 * for that device and problem shape; results are bit-identical to the
 * defaults.  No reference counterpart (the reference's MATLAB has no launch
 * shapes).  The report lists "kernel/shape=choice" lines (len: its length). */
int fhs_ctx_set_autotune(fhs_ctx* ctx, int enable);
int fhs_autotune_report(fhs_ctx* ctx, char* buf, int cap, int* len);
/* Live per-kernel timing with HIP events on the context stream. */
int fhs_ctx_set_timing(fhs_ctx* ctx, int enable);
/* Copies up to cap entries: names as "name\n" concatenated into names_buf
 * (size names_cap), total milliseconds and launch counts. */
int fhs_ctx_timing_report(fhs_ctx* ctx, char* names_buf, int names_cap,
                          double* total_ms, long* counts, int cap, int* n_out);
int fhs_ctx_timing_reset(fhs_ctx* ctx);
/* Device memory for FFI callers without an allocator (MEX, ctypes). */
int fhs_device_alloc(fhs_ctx* ctx, int64_t bytes, void** ptr);
int fhs_device_free(fhs_ctx* ctx, void* ptr);
int fhs_memcpy_h2d(fhs_ctx* ctx, void* dst, const void* src, int64_t bytes);
int fhs_memcpy_d2h(fhs_ctx* ctx, void* dst, const void* src, int64_t bytes);

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
/* Separable min-max G of a 2-D NUFFT (no w-term): G[m, (jy, jx)] =
 * cy[m*Ky + jy] * cx[m*Kx + jx] on the wrapped window anchored at
 * anchor[m] = ky0 + Kty*kx0 (interleaved complex, host pointers; any
 * per-row scaling such as the noise whitening folded into cx).
 * Ky, Kx <= 16 (rows of 8 or 16 slots per axis), M < 2^26; more than 64
 * taps per row (w-projected rows) needs a grid of whole 32 x 32 tiles.
 * Same products as the window form, 3.5x less coefficient traffic. */
int fhs_gmat_create_separable_host(fhs_ctx* ctx, int64_t M, int Kty, int Ktx, int Ky, int Kx,
                                   const int32_t* anchor, const float* cy, const float* cx,
                                   fhs_gmat** out);
/* G straight from a MATLAB sparse matrix (CSC: mxGetJc / mxGetIr as
 * int64, interleaved complex values of the R2018a API, host pointers). */
int fhs_gmat_create_csc_host(fhs_ctx* ctx, int64_t M, int64_t ncols,
                             const int64_t* colptr, const int64_t* rowidx,
                             const float* vals, fhs_gmat** out);
int fhs_gmat_destroy(fhs_gmat* g);
int fhs_gmat_info(const fhs_gmat* g, int64_t* M, int64_t* ncols, int64_t* nnz);

/* y = Lambda .* (G * Fx)     src/apply_direct_operator.m:1-22
 * Fx: complex[ncols]; Lambda: float[M] or NULL (= 1); y: complex[M]. */
int fhs_apply_direct_operator(fhs_ctx* ctx, const fhs_gmat* G, const void* Fx,
                              const float* Lambda, void* y);
/* Fx = G' * (Lambda .* y)    src/apply_adjoint_operator.m:1-23 */
int fhs_apply_adjoint_operator(fhs_ctx* ctx, const fhs_gmat* G, const void* y,
                               const float* Lambda, void* Fx);

/* ------------------------------------- holographic (DR) data operator -- */
/* Dimensionality-reduced form of one channel's separable G, built on the
 * GPU (lib/operators/util_load_data_gen_dr_measurement_operator.m:196-225):
 * H = tril(G'G) restricted to the active Fourier modes W = {p : (G'G)_pp >
 * 0} (:203) with its diagonal halved (:215-222).  H is square over the
 * whole Kty*Ktx grid with the rows and columns off W empty (W ascending, so
 * tril over W's order is tril over grid order); data vectors of the DR form
 * are grid-sized and zero off W, their entries on W being the reference's
 * compact vectors.  Deterministic (no atomics). */
int fhs_gmat_create_holographic(fhs_ctx* ctx, const fhs_gmat* G, fhs_gmat** H);
/* n_modes = |W|; Lambda (device float[Kty*Ktx], may be NULL) <- 1 ./
 * sqrt(diag(G'G)) on W (:206), 0 off W. */
int fhs_gmat_holographic_info(fhs_ctx* ctx, const fhs_gmat* H, int64_t* n_modes, float* Lambda);
/* yh = Lambda .* (G' y) on the grid (device complex[M] -> complex[Kty*Ktx]):
 * the gridded data (:171, :209), or with y a noise vector the gridded noise
 * whose norm is the DR l2 bound (:190-192, :211-212). */
int fhs_holographic_grid(fhs_ctx* ctx, const fhs_gmat* G, const fhs_gmat* H, const void* y, void* yh);

/* ------------------------------------------------- scaled Fourier (A) -- */
/* Plan for A = fft2(pad(sn .* x)) with sn = sy * sx' (real scaling factors
 * of the NUFFT, host arrays sy[Ny], sx[Nx]); grid Kty x Ktx of the form
 * 2^a 3^b 5^c in [16, 8192] (powers of two up to 4096 take the tuned
 * fixed-length kernels, other sizes a runtime mixed-radix path), Kty >= Ny,
 * Ktx >= Nx. */
int fhs_nufft_create_host(fhs_ctx* ctx, int Ny, int Nx, int Kty, int Ktx,
                          const double* sy, const double* sx, fhs_nufft** out);
int fhs_nufft_destroy(fhs_nufft* p);
/* Sizes of a plan (any pointer may be NULL): image Ny x Nx, grid Kty x Ktx. */
int fhs_nufft_info(const fhs_nufft* p, int* Ny, int* Nx, int* Kty, int* Ktx);
/* Min-max interpolation coefficients of M visibilities, computed on the GPU
 * in FP64: the per-visibility loop of op_p_nufft that builds G
 * (lib/operators/util_gen_measurement_operator_synth.m:85).  Per axis: image
 * size N, grid K, J <= 8 taps, na <= 8 cosine-series scaling terms a[na]
 * and the J x J matrix T (host arrays); y is paired with v, x with u.
 * u, v, row_scale (or NULL): device double[M]; outputs (device):
 * anchor int32[M] = (k0y+1 mod Kty) + Kty (k0x+1 mod Ktx), cy complex[M][Jy],
 * cx complex[M][Jx] (times row_scale), so G(m, (jy, jx)) = cy[m][jy] cx[m][jx]
 * (the layout of fhs_gmat_create_separable_host). */
int fhs_nufft_interp_coefficients(fhs_ctx* ctx, int Ny, int Kty, int Jy, int nay, const double* ay,
                                  const double* Ty, int Nx, int Ktx, int Jx, int nax, const double* ax,
                                  const double* Tx, const double* u, const double* v,
                                  const double* row_scale, int64_t M, int32_t* anchor, float* cy, float* cx);
/* W-projected rows of G from min-max rows (anchor_in, cy_in [M][Jy],
 * cx_in [M][Jx] of fhs_nufft_interp_coefficients, device): the per-
 * visibility step of op_p_nufft_wproj_dde
 * (lib/operators/util_load_data_gen_measurement_operator.m:119-121) with
 * the Fresnel w-phase exp(i pi w (l^2 + m^2)), which keeps rows separable:
 * per axis cy' = (1/Kty) cy correlated with the w-kernel Chat(w)[d], |d| <= S,
 * then cut to the largest taps holding keep_y (keep_x) of the row's energy
 * (the reference's GEnergyL2 split over the two axes, keep = sqrt(GEnergyL2)).
 * table_y / table_x: host FP64 complex [nbins][2S + 1], Chat at w = b dw
 * (b = 0..nbins-1, |w| <= (nbins - 1) dw; CEnergyL2 applied by the caller);
 * w: device double[M] in wavelengths.  Outputs (device): anchor_out
 * (S points before the input anchor on each axis), cy_out [M][Jy + 2 Sy],
 * cx_out [M][Jx + 2 Sx] (Jy + 2 Sy, Jx + 2 Sx <= 16): the layout of
 * fhs_gmat_create_separable_host. */
int fhs_nufft_wproj_rows(fhs_ctx* ctx, int Kty, int Ktx, int Jy, int Jx, int Sy, int Sx, int nbins, double dw,
                         const double* table_y, const double* table_x, double keep_y, double keep_x,
                         const double* w, int64_t M, const int32_t* anchor_in, const float* cy_in,
                         const float* cx_in, int32_t* anchor_out, float* cy_out, float* cx_out);
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

/* ------------------------------------------------------ preconditioner -- */
/* aW = util_gen_preconditioning_matrix(u, v, param_precond) with uniform
 * weighting (gen_uniform_weight_matrix, imaging/default_parameters.json:47-51;
 * called at lib/operators/util_gen_measurement_operator_synth.m:74 and
 * lib/operators/util_load_data_gen_measurement_operator.m:116): 1 / (number
 * of points in the same cell of the Kty x Ktx grid).  u, v: device double[M]
 * in radians ([-pi, pi]); aW: device float[M] out.  Deterministic. */
int fhs_precond_uniform_weights(fhs_ctx* ctx, const double* u, const double* v, int64_t M, int Kty, int Ktx,
                                float* aW);

/* ---------------------------------------------------- l2 ball / duals -- */
/* p = y + (x - y) * min(eps / ||x - y||, 1)      src/proj_l2ball.m:20-22
 * x, y, p: complex[n] (p may alias x). */
int fhs_proj_l2ball(fhs_ctx* ctx, const void* x, double eps, const void* y,
                    int64_t n, void* p);

/* fhs_proj_l2ball without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same kernels
 * and results as fhs_proj_l2ball; its reduction partials live in a fixed
 * 16 KiB array per device (2048 FP64 partials, enough for every n: the
 * reduction grid is capped at 2048 workgroups), registered by
 * fhs_ctx_create.  The context must come from fhs_ctx_create.  Calls that
 * overlap on one GPU share that array and race: issue them in order (one
 * stream, or streams ordered by events). */
int fhs_proj_l2ball_noalloc(fhs_ctx* ctx, const void* x, double eps, const void* y, int64_t n,
                            void* p);

/* [epsilon, t_block] = update_epsilon(epsilon, t, t_block, norm_res,
 *   adapt_eps_tol_in, adapt_eps_tol_out, adapt_eps_steps,
 *   adapt_eps_change_percentage)   src/update_epsilon.m:1-73
 * Cells {L}{nblocks} flattened to n blocks (host arrays, updated in place);
 * n_updated (may be NULL) counts the changes the reference prints. */
int fhs_update_epsilon(double* epsilon, int t, int* t_block, const double* norm_res, int n,
                       double adapt_eps_tol_in, double adapt_eps_tol_out, int adapt_eps_steps,
                       double adapt_eps_change_percentage, int* n_updated);

/* ------------------------------------------------------------- primal -- */
/* xsol <- max(xsol - (g_a + g_b + g_c), 0); xhat = 2 xsol - xsol_prev;
 * rel_x = sum (xsol_prev - xsol)^2, norm_x = sum xsol^2.
 * src/fhs/fhs_update_primal.m:33-38 (g_b, g_c may be NULL). */
int fhs_update_primal(fhs_ctx* ctx, float* xsol, const float* g_a,
                      const float* g_b, const float* g_c, float* xhat,
                      int64_t n, double* rel_x, double* norm_x);

/* --------------------------------------------------- data fidelity -- */
/* State of the data term of L channels sharing one scaled FFT A.  G[l]
 * (rows of channel l; its columns index A's Kty*Ktx grid, i.e. W is the
 * whole grid) is stacked: channel l owns rows row_offset[l]..row_offset[l+1]
 * of every M-vector (v2, y, proj, pU, Gx_old, Lambda).  Data blocks
 * (y{l}{b} of the reference) are row ranges block_start[b]..block_start[b+1]
 * of channel block_channel[b]; they tile all rows in channel order.
 * Host arrays; G and A must outlive the state. */
int fhs_data_create(fhs_ctx* ctx, const fhs_nufft* A, int L, const fhs_gmat* const* G,
                    int nblocks, const int64_t* block_start, const int* block_channel,
                    fhs_data** out);
int fhs_data_destroy(fhs_data* data);
/* Storage precision of the interpolation rows the data state's kernels read
 * (forward, adjoint, FB data step): bits = 32, the G rows as created, or
 * 16, an FP16 copy of every channel's rows (4 B per complex slot: half the
 * row bytes of the dominant gridding kernels; relative rounding 2^-11 per
 * coefficient).  FP16 needs tile-gridded separable G on every channel
 * (FHS_ERR_UNSUPPORTED otherwise).  No counterpart in the reference, whose
 * G is double: a storage option of this library's mixed-precision form.
 * The G objects themselves are unchanged; call it between solves, with no
 * launch on the state in flight. */
int fhs_data_set_row_precision(fhs_ctx* ctx, fhs_data* data, int bits);
/* Sizes of a data state (any pointer may be NULL): image Ny x Nx, L
 * channels, No = Kty*Ktx grid points, Mtot stacked rows, data blocks. */
int fhs_data_info(const fhs_data* data, int* Ny, int* Nx, int* L, int64_t* No, int64_t* Mtot, int* nblocks);
/* The scaled-FFT plan A the state was created with. */
int fhs_data_plan(const fhs_data* data, const fhs_nufft** plan);
/* Gx_old = G A(x) (Lambda (G + G') A(x) for the DR form): the state that
 * replaces Fx_old = A(x) (src/hs/hyperSARA.m, Fxi_old = apply_scaled_
 * fourier_transform(xsol, A, No)); by linearity G(2 Fx - Fx_old) =
 * 2 G Fx - Gx_old, so M complex per channel are kept instead of No. */
int fhs_data_init_fx_old(fhs_ctx* ctx, fhs_data* data, const float* x, const float* Lambda,
                         int flag_dimensionality_reduction, void* Gx_old);
/* The same state from the reference's own Fx_old = A(x_old) array
 * (src/fhs/facetHyperSARA.m:574-576, src/hs/hyperSARA.m:446, consumed at
 * src/update_dual_data_fidelity.m:92,132): Fx_old complex[L][Kty*Ktx] on the
 * device (MATLAB's [No, L] column-major array); Gx_old = G_l Fx_old(:, l)
 * per channel (Lambda (G + G') Fx_old(:, l) for the DR form). */
int fhs_data_fx_old_to_state(fhs_ctx* ctx, fhs_data* data, const void* Fx_old, const float* Lambda,
                             int flag_dimensionality_reduction, void* Gx_old);
/* y = Lambda .* (G_l A x_l) for every channel (stacked): the measurement
 * operator of HS_forward_operator_G.m / util_operator_norm.m:45 (Lambda
 * NULL = 1, or sqrt(aW) for the preconditioned operator). */
int fhs_data_forward(fhs_ctx* ctx, fhs_data* data, const float* x, const float* Lambda, void* y);
/* out_l = scale_l real(At(G_l' (Lambda .* y))): HS_adjoint_operator_G.m
 * (scale: host float[L] or NULL). */
int fhs_data_adjoint(fhs_ctx* ctx, fhs_data* data, const void* y, const float* Lambda,
                     const float* scale, float* out);
/* The same pair with flag_visibility_gridding on a lower-triangular
 * (holographic) G: y = Lambda .* (G + G') A x and out_l = scale_l
 * real(At((G' + G)(Lambda .* y))) (lib/operators/HS_forward_operator_G.m,
 * HS_adjoint_operator_G.m, the istril branch).  Square G only. */
int fhs_data_forward_dr(fhs_ctx* ctx, fhs_data* data, const float* x, const float* Lambda, void* y);
int fhs_data_adjoint_dr(fhs_ctx* ctx, fhs_data* data, const void* y, const float* Lambda,
                        const float* scale, float* out);
/* [v2, Ftx, Fx_old, proj, norm_res, global_norm_res, norm_epsilon] =
 *   update_dual_data_fidelity(v2, y, x, Fx_old, proj, A, At, G, W, pU,
 *     epsilon, elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps,
 *     sigma22, flag_dimensionality_reduction, Lambda)
 * src/update_dual_data_fidelity.m:1-149.  Device: v2, y, proj, Gx_old
 * (complex[Mtot], updated in place), x float[L][Nx][Ny], pU float[Mtot] (NULL
 * = 1), Lambda float[Mtot] (DR only, NULL = 1), Ftx float[L][Nx][Ny] out.
 * Host: epsilon double[nblocks], sigma22 float[L], norm_res double[nblocks]
 * out (may be NULL), global_norm_res and norm_epsilon (sums of squares) out. */
int fhs_update_dual_data_fidelity(fhs_ctx* ctx, fhs_data* data, void* v2, const void* y,
                                  const float* x, void* Gx_old, void* proj, const float* pU,
                                  const double* epsilon, int elipse_proj_max_iter,
                                  int elipse_proj_min_iter, double elipse_proj_eps,
                                  const float* sigma22, int flag_dimensionality_reduction,
                                  const float* Lambda, float* Ftx, double* norm_res,
                                  double* global_norm_res, double* norm_epsilon);

/* ------------------------------------------------------ SARA prior -- */
/* SARA dictionary (op_sp_wlt_basis of the SARA-dictionary submodule,
 * built at imaging/imaging.m:797-810 and src/hs/hyperSARA.m): nbases
 * Daubechies bases (reconstruction low-pass filters lo_r concatenated,
 * lengths filter_len[b], even, <= 16) + the Dirac basis when with_dirac,
 * nlevel levels, 'zpd' extension, each basis scaled by 1/sqrt(P).
 * Per-channel coefficient vector: MATLAB wavedec2 order per basis, bases in
 * order, Dirac last. */
int fhs_sara_create_host(fhs_ctx* ctx, int Ny, int Nx, int nlevel, int nbases,
                         const int* filter_len, const double* lo_r, int with_dirac,
                         fhs_sara** out);
int fhs_sara_destroy(fhs_sara* S);
int fhs_sara_info(const fhs_sara* S, int64_t* n_coefs);
/* coefs(:, l) (+)= Psit(x(:, :, l)); coefs float[L][s] */
int fhs_sara_analysis(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, float* coefs,
                      int accumulate);
/* out(:, :, l) = scale * Psi(coefs(:, l)) */
int fhs_sara_synthesis(fhs_ctx* ctx, const fhs_sara* S, const float* coefs, int L, float scale,
                       float* out);
/* [v1, g1] = hs_update_dual_sparsity_distributed(v1, Psit, Psi, xhat,
 *   weights1, beta1, sigma1)  src/hs/hs_update_dual_sparsity_distributed.m:4-17
 * (also fhs_update_dual_sparsity.m:67-93 on the whole image, and the L = 1
 * SARA step src/sara/sara.m:921-926).  v1 float[L][s] in place, xhat
 * float[L][Nx][Ny], weights1 float[s] (NULL = 1), g1 out. */
int fhs_update_dual_sparsity(fhs_ctx* ctx, const fhs_sara* S, float* v1, const float* xhat, int L,
                             const float* weights1, double beta1, float sigma1, float* g1);
/* weights1 = upsilon ./ (upsilon + ||Psit(x)||_row): the l21 reweighting
 * of src/hs/hs_update_weights_sparsity_distributed.m:8-13 and
 * src/fhs/fhs_update_weights.m:82-92 (upsilon = sig * alpha); float[s] out. */
int fhs_sara_update_weights(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, double upsilon,
                            float* weights1);
/* f = compute_sara_prior(x, Psit, n_wavelet_coefficients)
 * src/compute_sara_prior.m:33-41 (host double out). */
int fhs_compute_sara_prior(fhs_ctx* ctx, const fhs_sara* S, const float* x, int L, double* out);

/* Faceted SARA transform over a grid of non-overlapping facets (the
 * reference's faceted wavelet submodule, absent from its tree: sdwt2_setup
 * src/fhs/facetHyperSARA.m:325-326, sdwt2_sara_faceting /
 * isdwt2_sara_faceting src/fhs/fhs_update_dual_sparsity.m:71-93).  Facet
 * q = qx*Qy + qy owns pixels [ybounds[qy], ybounds[qy+1]) x [xbounds[qx],
 * xbounds[qx+1]); dictionary arguments as fhs_sara_create_host.  Each facet
 * reads an extended image (its pixels plus a top/left border from its
 * neighbours, zero outside the image) and owns an exact share of the
 * whole-image coefficients; the sum of the facets' syntheses is the
 * whole-image synthesis. */
int fhs_sara_facets_create(fhs_ctx* ctx, int Ny, int Nx, int nlevel, int nbases, const int* filter_len,
                           const double* lo_r, int with_dirac, int Qy, int Qx, const int* ybounds,
                           const int* xbounds, fhs_sara_facets** out);
int fhs_sara_facets_destroy(fhs_sara_facets* F);
/* ext: int[4] = origin (row, col; may be negative) and size (rows, cols) of
 * facet q's extended image; n_coefs: its coefficients per channel;
 * ncoefs_rows (or NULL): int[2 * (nbases (nlevel + 1) + with_dirac)], per
 * basis the (rows, cols) of its levels 1..nlevel and of the approximation,
 * the Dirac part last (the rows of the reference's Ncoefs{q}). */
int fhs_sara_facets_info(const fhs_sara_facets* F, int q, int* ext, int64_t* n_coefs, int* ncoefs_rows);
/* z = sdwt2_sara_faceting(x_ext): x_ext float[L][enx][eny], z float[L][n_coefs] */
int fhs_sara_facet_analysis(fhs_ctx* ctx, const fhs_sara_facets* F, int q, const float* x_ext, int L, float* z);
/* g_ext = scale * isdwt2_sara_faceting(z): float[L][enx][eny], facet q's
 * contribution over its extended image */
int fhs_sara_facet_synthesis(fhs_ctx* ctx, const fhs_sara_facets* F, int q, const float* z, int L, float scale,
                             float* g_ext);
/* [v1, g] = fhs_update_dual_sparsity(v1, x_facet, weights1, beta1, ...)
 * src/fhs/fhs_update_dual_sparsity.m:65-93 for facet q: w = v1 + Psit_q
 * x_ext; v1 <- w - max(||w|| - beta1 weights1, 0)/(||w|| + eps) w (norms
 * over the L channels); g_ext = Psi_q v1.  v1 float[L][n_coefs], weights1
 * float[n_coefs] or NULL (= 1). */
int fhs_update_dual_sparsity_facet(fhs_ctx* ctx, const fhs_sara_facets* F, int q, float* v1, const float* x_ext,
                                   int L, const float* weights1, double beta1, float* g_ext);

/* -------------------------------------------------- low-rank prior -- */
/* Facets q = qx*Qy + qy (src/fhs/facetHyperSARA.m:310-321 order) with origin
 * (oy, ox) and size (py, px); windows: concatenated py*px apodisation
 * windows (column-major) or NULL (= 1).  L <= 16.  A single facet covering
 * the image is the HyperSARA global prior.  v0 is stored per facet as
 * L x (py*px) blocks, facet after facet (total v0_size floats). */
int fhs_lowrank_create_host(fhs_ctx* ctx, int Ny, int Nx, int L, int Qy, int Qx, const int* oy,
                            const int* ox, const int* py, const int* px, const float* windows,
                            fhs_lowrank** out);
int fhs_lowrank_destroy(fhs_lowrank* R);
int fhs_lowrank_info(const fhs_lowrank* R, int64_t* v0_size);
/* Precision of the Gram Z'Z inside the nuclear-norm steps (no reference
 * counterpart: the reference takes an FP64 SVD of Z).  mode 0 (default):
 * FP64 sums of exact products.  3: BF16 hi + mid + lo parts of z (exact)
 * multiplied pairwise on the matrix cores, FP32 sums over 256 pixels, FP64
 * across them.  2: BF16 hi + lo parts (16 mantissa bits).  The eigen-
 * decomposition stays FP64.  A PDFB engine records its kernels at its first
 * captured iteration: set the mode before that. */
int fhs_lowrank_set_gram_precision(fhs_lowrank* R, int mode);
/* [v0, g0] = fhs_update_dual_lowrankness(v0, xhat, apodization_window,
 *   weights0, beta0)   src/fhs/fhs_update_dual_lowrankness.m:40-45, and
 * hs_update_dual_lowrankness_serial.m:5-10 for one facet.  g0 =
 * sigma0 * sum_q w_q .* v0_q (overlap-added image, float[L][Nx][Ny]).
 * weights0: host double[Q][L] (descending-singular-value order). */
int fhs_update_dual_lowrankness(fhs_ctx* ctx, fhs_lowrank* R, float* v0, const float* xhat,
                                const double* weights0, double beta0, float sigma0, float* g0);
/* singular values (descending) of w_q .* x_q per facet: host double[Q][L] */
int fhs_lowrank_singular_values(fhs_ctx* ctx, fhs_lowrank* R, const float* x, double* sv);
/* f = nuclear_norm(x)   src/nuclear_norm.m:25-26; x float[L][Nx][Ny] */
int fhs_nuclear_norm(fhs_ctx* ctx, const float* x, int Ny, int Nx, int L, double* out);

/* ------------------------------------------------------ PDFB engine -- */
/* Whole PDFB iterations of SARA (lowrank = NULL, L = 1), HyperSARA (one
 * facet) and Faceted HyperSARA (Q facets): src/sara/sara.m:521-633,
 * src/hs/hyperSARA.m:565-617, src/fhs/facetHyperSARA.m:701-783.
 * Caller-owned device buffers (float = float32, complex = complex64):
 *   x, xhat, g0 (NULL without low-rank prior), g1, g2: float[L][Nx][Ny]
 *   v0: lowrank v0_size floats (NULL without), v1: float[L][s]
 *   v2, proj, Gx_old: complex[Mtot] (in/out), y: complex[Mtot]
 *   pU: float[Mtot] (NULL = 1), weights1: float[s] (NULL = 1),
 *   Lambda: float[Mtot] (DR only, NULL = 1).
 * The engine runs on its own stream (default flags and priority). */
typedef struct fhs_pdfb fhs_pdfb;
typedef struct {
  int Ny, Nx;
  float* x;
  float* xhat;
  float* g0;
  float* g1;
  float* g2;
  float* v0;
  float* v1;
  void* v2;
  void* proj;
  void* Gx_old;
  const void* y;
  const float* pU;
  const float* weights1;
  const float* Lambda;
} fhs_pdfb_buffers;
int fhs_pdfb_create(fhs_ctx* ctx, fhs_data* data, fhs_sara* sara, fhs_lowrank* lowrank,
                    const fhs_pdfb_buffers* bufs, fhs_pdfb** out);
int fhs_pdfb_destroy(fhs_pdfb* p);
/* thresh0: host double[Q][L] = beta0_q * weights0_q (descending singular
 * value order; NULL without low-rank prior); beta1 = gamma / sigma1
 * (src/fhs/facetHyperSARA.m:598-603); sigma00 = tau sigma0, sigma11 = tau
 * sigma1, sigma22: host float[L] = tau sigma2 (:582-596);
 * epsilon: host double[nblocks]. */
int fhs_pdfb_set_params(fhs_pdfb* p, const double* thresh0, double beta1, float sigma00, float sigma11,
                        const float* sigma22, const double* epsilon, int elipse_proj_max_iter,
                        int elipse_proj_min_iter, double elipse_proj_eps,
                        int flag_dimensionality_reduction);
/* replay iterations from a captured hipGraph (default on) */
int fhs_pdfb_set_graph(fhs_pdfb* p, int enable);
/* run the prior steps on a second stream beside the data step (default on) */
int fhs_pdfb_set_concurrent(fhs_pdfb* p, int enable);
/* runs n_iter iterations; scalars: host double[4*n_iter] or NULL:
 * sum (x_prev - x)^2, sum x^2, sum_b norm_res_b^2, sum_b eps_b^2 */
int fhs_pdfb_iterate(fhs_pdfb* p, int n_iter, double* scalars);
/* Reference cube X0 (device float[L][Nx][Ny], NULL = none) whose distance
 * to x the primal update sums in FP64 each iteration, fused with its other
 * reductions. This replaces the per-iteration SNR evaluation
 * `snr = 20*log10(norm(X0(:))/norm(X0(:)-xsol(:)))` of the solvers' logging
 * (src/hs/hyperSARA.m:549, src/fhs/facetHyperSARA.m:864) without a second pass over
 * the cube or a second synchronisation. X0 must stay valid while set. */
int fhs_pdfb_set_truth(fhs_pdfb* p, const float* X0);
/* last iteration's sum (X0 - x)^2 (host copy, no GPU call) */
int fhs_pdfb_truth_error(fhs_pdfb* p, double* err2);
/* The PDFB (inner) stopping rule of the solvers, evaluated on the host:
 *   k >= min_iter && (k >= max_iter || small || rel_val <= rel_var_low),
 *   small = rel_val <= rel_var, and for rule 1 also norm_res <= fid_tol * norm_eps.
 * rule 0: src/sara/sara.m:666-670 and src/hs/hyperSARA.m:682-686 (no
 * data-fidelity term); rule 1: src/fhs/facetHyperSARA.m:877-881.  k counts
 * iterations since the last reweighting; NaN compares false, as in MATLAB.
 * Returns 1 when the cycle has converged, else 0. */
int fhs_pdfb_stopping_rule(int rule, int k, double rel_val, double norm_res, double norm_eps, int min_iter,
                           int max_iter, double rel_var, double fid_tol, double rel_var_low);
/* Runs iterations until fhs_pdfb_stopping_rule(rule, k, ...) holds for
 * iteration k = k0 + i, rel_val = sqrt(sum dx^2 / sum x^2), norm_res and
 * norm_eps the square roots of the published sums; or until n_max iterations, or until
 * sum (X0 - x)^2 <= err2_stop (err2_stop >= 0; needs fhs_pdfb_set_truth),
 * or after seconds_max seconds of wall-clock (> 0).  The per-iteration loop
 * of the solvers without a host round trip per iteration.
 * out: host double[6 * n_max]: per iteration the four sums of
 * fhs_pdfb_iterate, sum (X0 - x)^2 (0 without a reference cube) and the
 * seconds since the call began; n_done: iterations run; converged: 1 if the
 * stopping rule held at the last one. */
int fhs_pdfb_run(fhs_pdfb* p, int rule, int k0, int min_iter, int max_iter, double rel_var, double fid_tol,
                 double rel_var_low, int n_max, double err2_stop, double seconds_max, double* out,
                 int* n_done, int* converged);
/* fhs_pdfb_iterate and fhs_pdfb_run fail (FHS_ERR_INTERNAL) when an
 * iteration's data residual is NaN: a workgroup exchange of the ellipsoid
 * projection that timed out (its counters are then reset, so the next call
 * starts clean), or non-finite data.  Test hook: marks the exchange as
 * timed out from the host, as a real time-out would. */
int fhs_pdfb_inject_fb_timeout(fhs_pdfb* p);
/* last iteration's per-block residual norms ||G_b A x - y_b||: host double[nblocks] */
int fhs_pdfb_norm_res(fhs_pdfb* p, double* norm_res);
int fhs_pdfb_stream(fhs_pdfb* p, void** stream);

/* ------------------------------------------- learned warm start (U-Net) -- */
/* No reference counterpart: the initial cube the reference loads through
 * xinit (imaging/imaging.m:881-891) is produced here by a U-Net trained
 * with these kernels.  Activations FP32 [C][N][H][W]; GEMM operands rounded
 * to BF16, FP32 accumulation on the matrix cores. */
/* C = alpha op(A) op(B) + beta C (+ bias[m]) (ReLU); op(A) = A^T when
 * trans_a (A stored K x M, leading dim lda), op(B) = B^T when trans_b. */
int fhs_nn_gemm(fhs_ctx* ctx, int M, int N, int K, const float* A, int lda, int trans_a, const float* B,
                int ldb, int trans_b, float* C, int ldc, float alpha, float beta, const float* bias, int relu);
/* col[(c*9 + ky*3 + kx)][n,y,x] = in[c][n][y+ky-1][x+kx-1] (zero padding) */
int fhs_nn_im2col3(fhs_ctx* ctx, const float* in, int C, int NB, int H, int W, float* col);
/* adjoint of im2col3 (gather form, no atomics) */
int fhs_nn_col2im3(fhs_ctx* ctx, const float* dcol, int C, int NB, int H, int W, float* din);
/* g = 0 where y <= 0 */
int fhs_nn_relu_bwd(fhs_ctx* ctx, const float* y, float* g, int64_t n);
int fhs_nn_maxpool2(fhs_ctx* ctx, const float* in, int64_t planes, int H, int W, float* out);
int fhs_nn_maxpool2_bwd(fhs_ctx* ctx, const float* in, const float* gout, int64_t planes, int H, int W,
                        float* gin);
int fhs_nn_upsample2(fhs_ctx* ctx, const float* in, int64_t planes, int H, int W, float* out);
int fhs_nn_upsample2_bwd(fhs_ctx* ctx, const float* gout, int64_t planes, int H, int W, float* gin);
/* db[c] = sum_q g[c][q] */
int fhs_nn_bias_grad(fhs_ctx* ctx, const float* g, int C, int64_t Q, float* db);
/* loss = sum (pred - target)^2 (host double), grad = grad_scale (pred - target) (NULL: none) */
int fhs_nn_mse(fhs_ctx* ctx, const float* pred, const float* target, int64_t n, float grad_scale, float* grad,
               double* loss);
/* Adam step t (bias-corrected) */
int fhs_nn_adam(fhs_ctx* ctx, float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                float beta2, float eps, int step);

#ifdef __cplusplus
}
#endif
#endif /* FHS_H */
