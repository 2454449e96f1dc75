"""ctypes binding of libfhs.so (the C ABI declared in include/fhs.h).

The library is loaded from this directory (built in-tree by
``faceted-hypersara_amd/build.py``).  There is no fallback: if the library is
missing or was built without a symbol, loading fails with an error naming it.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libfhs.so')

c_void_p = C.c_void_p
c_int = C.c_int
c_i64 = C.c_int64
c_double = C.c_double
c_float_p = C.POINTER(C.c_float)
c_double_p = C.POINTER(C.c_double)
c_i64_p = C.POINTER(C.c_int64)
c_i32_p = C.POINTER(C.c_int32)
c_int_p = C.POINTER(C.c_int)
c_long_p = C.POINTER(C.c_long)

# name -> argtypes (all return int status unless listed in _RESTYPE)
PROTOTYPES = {
    'fhs_ctx_create': [c_int, C.POINTER(c_void_p)],
    'fhs_ctx_destroy': [c_void_p],
    'fhs_last_error': [c_void_p],
    'fhs_ctx_set_stream': [c_void_p, c_void_p],
    'fhs_ctx_synchronize': [c_void_p],
    'fhs_ctx_set_timing': [c_void_p, c_int],
    'fhs_ctx_set_autotune': [c_void_p, c_int],
    'fhs_autotune_report': [c_void_p, C.c_char_p, c_int, c_int_p],
    'fhs_ctx_timing_report': [c_void_p, C.c_char_p, c_int, c_double_p, c_long_p, c_int, c_int_p],
    'fhs_ctx_timing_reset': [c_void_p],
    'fhs_gmat_create_csr_host': [c_void_p, c_i64, c_i64, c_i64_p, c_i64_p, c_float_p, C.POINTER(c_void_p)],
    'fhs_gmat_create_window_host': [c_void_p, c_i64, c_int, c_int, c_int, c_int, c_i32_p, c_float_p,
                                    C.POINTER(c_void_p)],
    'fhs_gmat_create_separable_host': [c_void_p, c_i64, c_int, c_int, c_int, c_int, c_i32_p, c_float_p,
                                       c_float_p, C.POINTER(c_void_p)],
    'fhs_gmat_create_csc_host': [c_void_p, c_i64, c_i64, c_i64_p, c_i64_p, c_float_p, C.POINTER(c_void_p)],
    'fhs_gmat_destroy': [c_void_p],
    'fhs_device_alloc': [c_void_p, c_i64, C.POINTER(c_void_p)],
    'fhs_device_free': [c_void_p, c_void_p],
    'fhs_memcpy_h2d': [c_void_p, c_void_p, c_void_p, c_i64],
    'fhs_memcpy_d2h': [c_void_p, c_void_p, c_void_p, c_i64],
    'fhs_update_epsilon': [c_double_p, c_int, c_int_p, c_double_p, c_int, c_double, c_double, c_int, c_double,
                           c_int_p],
    'fhs_gmat_info': [c_void_p, c_i64_p, c_i64_p, c_i64_p],
    'fhs_apply_direct_operator': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    'fhs_apply_adjoint_operator': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    'fhs_gmat_create_holographic': [c_void_p, c_void_p, C.POINTER(c_void_p)],
    'fhs_gmat_holographic_info': [c_void_p, c_void_p, c_i64_p, c_void_p],
    'fhs_holographic_grid': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    'fhs_nufft_create_host': [c_void_p, c_int, c_int, c_int, c_int, c_double_p, c_double_p,
                              C.POINTER(c_void_p)],
    'fhs_nufft_interp_coefficients': [c_void_p, c_int, c_int, c_int, c_int, c_double_p, c_double_p, c_int, c_int,
                                      c_int, c_int, c_double_p, c_double_p, c_void_p, c_void_p, c_void_p, c_i64,
                                      c_void_p, c_void_p, c_void_p],
    'fhs_nufft_wproj_rows': [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_double_p,
                             c_double_p, c_double, c_double, c_void_p, c_i64, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p],
    'fhs_nufft_destroy': [c_void_p],
    'fhs_apply_scaled_fourier_transform': [c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    'fhs_apply_scaled_fourier_adjoint': [c_void_p, c_void_p, c_void_p, c_int, c_float_p, c_void_p],
    'fhs_proj_l2ball': [c_void_p, c_void_p, c_double, c_void_p, c_i64, c_void_p],
    'fhs_proj_l2ball_noalloc': [c_void_p, c_void_p, c_double, c_void_p, c_i64, c_void_p],
    'fhs_data_create': [c_void_p, c_void_p, c_int, C.POINTER(c_void_p), c_int, c_i64_p, c_int_p,
                        C.POINTER(c_void_p)],
    'fhs_data_destroy': [c_void_p],
    'fhs_data_set_row_precision': [c_void_p, c_void_p, c_int],
    'fhs_data_init_fx_old': [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    'fhs_data_fx_old_to_state': [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    'fhs_data_info': [c_void_p, c_int_p, c_int_p, c_int_p, c_i64_p, c_i64_p, c_int_p],
    'fhs_precond_uniform_weights': [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_int, c_void_p],
    'fhs_data_plan': [c_void_p, C.POINTER(c_void_p)],
    'fhs_nufft_info': [c_void_p, c_int_p, c_int_p, c_int_p, c_int_p],
    'fhs_update_dual_data_fidelity': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_double_p, c_int, c_int, c_double, c_float_p, c_int,
                                      c_void_p, c_void_p, c_double_p, c_double_p, c_double_p],
    'fhs_sara_create_host': [c_void_p, c_int, c_int, c_int, c_int, c_int_p, c_double_p, c_int,
                             C.POINTER(c_void_p)],
    'fhs_sara_destroy': [c_void_p],
    'fhs_sara_info': [c_void_p, c_i64_p],
    'fhs_sara_analysis': [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int],
    'fhs_sara_synthesis': [c_void_p, c_void_p, c_void_p, c_int, C.c_float, c_void_p],
    'fhs_sara_facets_create': [c_void_p, c_int, c_int, c_int, c_int, c_int_p, c_double_p, c_int, c_int, c_int,
                               c_int_p, c_int_p, C.POINTER(c_void_p)],
    'fhs_sara_facets_destroy': [c_void_p],
    'fhs_sara_facets_info': [c_void_p, c_int, c_int_p, c_i64_p, c_int_p],
    'fhs_sara_facet_analysis': [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p],
    'fhs_sara_facet_synthesis': [c_void_p, c_void_p, c_int, c_void_p, c_int, C.c_float, c_void_p],
    'fhs_update_dual_sparsity_facet': [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_double,
                                       c_void_p],
    'fhs_update_dual_sparsity': [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_double,
                                 C.c_float, c_void_p],
    'fhs_compute_sara_prior': [c_void_p, c_void_p, c_void_p, c_int, c_double_p],
    'fhs_lowrank_create_host': [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int_p, c_int_p, c_int_p,
                                c_int_p, c_float_p, C.POINTER(c_void_p)],
    'fhs_lowrank_destroy': [c_void_p],
    'fhs_lowrank_info': [c_void_p, c_i64_p],
    'fhs_lowrank_set_gram_precision': [c_void_p, c_int],
    'fhs_update_dual_lowrankness': [c_void_p, c_void_p, c_void_p, c_void_p, c_double_p, c_double,
                                    C.c_float, c_void_p],
    'fhs_lowrank_singular_values': [c_void_p, c_void_p, c_void_p, c_double_p],
    'fhs_nuclear_norm': [c_void_p, c_void_p, c_int, c_int, c_int, c_double_p],
    'fhs_pdfb_create': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(c_void_p)],
    'fhs_pdfb_destroy': [c_void_p],
    'fhs_pdfb_set_params': [c_void_p, c_double_p, c_double, C.c_float, C.c_float, c_float_p, c_double_p,
                            c_int, c_int, c_double, c_int],
    'fhs_pdfb_set_graph': [c_void_p, c_int],
    'fhs_pdfb_set_concurrent': [c_void_p, c_int],
    'fhs_pdfb_iterate': [c_void_p, c_int, c_double_p],
    'fhs_pdfb_norm_res': [c_void_p, c_double_p],
    'fhs_pdfb_set_truth': [c_void_p, c_void_p],
    'fhs_pdfb_run': [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double, c_double, c_int, c_double,
                     c_double, c_double_p, c_int_p, c_int_p],
    'fhs_pdfb_truth_error': [c_void_p, c_double_p],
    'fhs_pdfb_stopping_rule': [c_int, c_int, c_double, c_double, c_double, c_int, c_int, c_double, c_double,
                               c_double],
    'fhs_pdfb_inject_fb_timeout': [c_void_p],
    'fhs_pdfb_stream': [c_void_p, C.POINTER(c_void_p)],
    'fhs_data_forward': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    'fhs_data_adjoint': [c_void_p, c_void_p, c_void_p, c_void_p, c_float_p, c_void_p],
    'fhs_data_forward_dr': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    'fhs_data_adjoint_dr': [c_void_p, c_void_p, c_void_p, c_void_p, c_float_p, c_void_p],
    'fhs_sara_update_weights': [c_void_p, c_void_p, c_void_p, c_int, c_double, c_void_p],
    'fhs_nn_gemm': [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_void_p, c_int,
                    C.c_float, C.c_float, c_void_p, c_int],
    'fhs_nn_im2col3': [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p],
    'fhs_nn_col2im3': [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p],
    'fhs_nn_relu_bwd': [c_void_p, c_void_p, c_void_p, c_i64],
    'fhs_nn_maxpool2': [c_void_p, c_void_p, c_i64, c_int, c_int, c_void_p],
    'fhs_nn_maxpool2_bwd': [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_int, c_void_p],
    'fhs_nn_upsample2': [c_void_p, c_void_p, c_i64, c_int, c_int, c_void_p],
    'fhs_nn_upsample2_bwd': [c_void_p, c_void_p, c_i64, c_int, c_int, c_void_p],
    'fhs_nn_bias_grad': [c_void_p, c_void_p, c_int, c_i64, c_void_p],
    'fhs_nn_mse': [c_void_p, c_void_p, c_void_p, c_i64, C.c_float, c_void_p, c_double_p],
    'fhs_nn_adam': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_i64, C.c_float, C.c_float, C.c_float,
                    C.c_float, c_int],
    'fhs_update_primal': [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_i64,
                          c_double_p, c_double_p],
}
_RESTYPE = {'fhs_last_error': C.c_char_p}


class FhsError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library (raises FhsError if it is absent or incomplete)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FhsError(f'libfhs.so not found at {LIB_PATH}: build it with '
                       f'python faceted-hypersara_amd/build.py')
    L = C.CDLL(LIB_PATH)
    for name, args in PROTOTYPES.items():
        try:
            f = getattr(L, name)
        except AttributeError as e:
            raise FhsError(f'libfhs.so lacks symbol {name}; rebuild it') from e
        f.argtypes = args
        f.restype = _RESTYPE.get(name, c_int)
    _lib = L
    return L


def check(status, ctx_ptr=None, what=''):
    if status != 0:
        msg = ''
        if ctx_ptr is not None:
            raw = lib().fhs_last_error(ctx_ptr)
            msg = raw.decode() if raw else ''
        raise FhsError(f'{what} failed (status {status}): {msg}')
