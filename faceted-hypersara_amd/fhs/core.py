"""Device handles over the C ABI: context, interpolation matrix G, scaled-FFT plan.

Tensors are torch tensors on the context's GPU; torch is used only to own
device memory.  Image cubes use the MATLAB memory layout: a MATLAB
``Ny x Nx x L`` array is a torch tensor of shape ``(L, Nx, Ny)``
(contiguous), a complex column vector is a complex64 tensor.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import check, lib

_default_ctx = {}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _require(t, dtype, device, what):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f'{what}: expected a torch tensor')
    if t.dtype != dtype:
        raise TypeError(f'{what}: expected {dtype}, got {t.dtype}')
    if t.device != device:
        raise ValueError(f'{what}: tensor on {t.device}, context on {device}')
    if not t.is_contiguous():
        raise ValueError(f'{what}: tensor must be contiguous')
    return t


class Context:
    """An ``fhs_ctx`` bound to one GPU; work goes on torch's current stream."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError('fhs needs a ROCm GPU (torch.cuda.is_available() is False)')
        self.device = torch.device('cuda', device)
        h = C.c_void_p()
        check(lib().fhs_ctx_create(device, C.byref(h)), None, 'fhs_ctx_create')
        self.h = h
        self.bind_stream()

    def bind_stream(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        check(lib().fhs_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream)), self.h, 'set_stream')

    def check(self, status, what):
        check(status, self.h, what)

    def synchronize(self):
        self.check(lib().fhs_ctx_synchronize(self.h), 'synchronize')

    def set_timing(self, on):
        self.check(lib().fhs_ctx_set_timing(self.h, 1 if on else 0), 'set_timing')

    def timing_report(self):
        cap = 64
        names = C.create_string_buffer(8192)
        ms = (C.c_double * cap)()
        cnt = (C.c_long * cap)()
        n = C.c_int()
        self.check(lib().fhs_ctx_timing_report(self.h, names, 8192, ms, cnt, cap, C.byref(n)),
                   'timing_report')
        keys = names.value.decode().split('\n')[:n.value]
        return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}

    def set_autotune(self, on):
        """Launch-shape autotuning (``fhs_ctx_set_autotune``): results are
        bit-identical either way."""
        self.check(lib().fhs_ctx_set_autotune(self.h, int(bool(on))), 'set_autotune')

    def autotune_report(self):
        """{kernel/shape: choice} of the launch shapes tuned on this device."""
        n = C.c_int()
        self.check(lib().fhs_autotune_report(self.h, None, 0, C.byref(n)), 'autotune_report')
        buf = C.create_string_buffer(n.value + 1)
        self.check(lib().fhs_autotune_report(self.h, buf, n.value + 1, C.byref(n)), 'autotune_report')
        out = {}
        for line in buf.value.decode().splitlines():
            k, _, v = line.partition('=')
            out[k] = v
        return out

    def timing_reset(self):
        self.check(lib().fhs_ctx_timing_reset(self.h), 'timing_reset')

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and lib is not None:
            try:
                lib().fhs_ctx_destroy(h)
            except Exception:
                pass
            self.h = None


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


class GMatrix:
    """Device interpolation matrix of one channel (``fhs_gmat``)."""

    def __init__(self, ctx, handle, M, ncols, nnz):
        self.ctx, self.h = ctx, handle
        self.M, self.ncols, self.nnz = M, ncols, nnz
        self.shape = (M, ncols)

    @classmethod
    def from_scipy(cls, ctx, G):
        """From any scipy sparse matrix (complex), row-compressed on the host."""
        G = G.tocsr()
        G.sort_indices()
        rowptr = np.ascontiguousarray(G.indptr, dtype=np.int64)
        col = np.ascontiguousarray(G.indices, dtype=np.int64)
        vals = np.ascontiguousarray(np.asarray(G.data, dtype=np.complex64).view(np.float32))
        h = C.c_void_p()
        ctx.check(lib().fhs_gmat_create_csr_host(
            ctx.h, G.shape[0], G.shape[1],
            rowptr.ctypes.data_as(C.POINTER(C.c_int64)), col.ctypes.data_as(C.POINTER(C.c_int64)),
            vals.ctypes.data_as(C.POINTER(C.c_float)), C.byref(h)), 'gmat_create_csr')
        return cls(ctx, h, G.shape[0], G.shape[1], G.nnz)

    @classmethod
    def from_window(cls, ctx, anchor, coef, Kty, Ktx, Ky, Kx):
        """Dense Ky x Kx window rows: anchor[m] = ky0 + Kty*kx0, coef (M, Ky*Kx)
        with tap index jx*Ky + jy."""
        anchor = np.ascontiguousarray(anchor, dtype=np.int32)
        coef = np.ascontiguousarray(np.asarray(coef, dtype=np.complex64).view(np.float32))
        M = anchor.shape[0]
        h = C.c_void_p()
        ctx.check(lib().fhs_gmat_create_window_host(
            ctx.h, M, Kty, Ktx, Ky, Kx, anchor.ctypes.data_as(C.POINTER(C.c_int32)),
            coef.ctypes.data_as(C.POINTER(C.c_float)), C.byref(h)), 'gmat_create_window')
        return cls(ctx, h, M, Kty * Ktx, M * Ky * Kx)

    @classmethod
    def from_separable(cls, ctx, anchor, cy, cx, Kty, Ktx):
        """Kronecker rows G[m, (jy, jx)] = cy[m, jy] cx[m, jx] (cy (M, Ky), cx (M, Kx))."""
        anchor = np.ascontiguousarray(anchor, dtype=np.int32)
        M, Ky = cy.shape
        Kx = cx.shape[1]
        cyf = np.ascontiguousarray(np.asarray(cy, dtype=np.complex64).view(np.float32))
        cxf = np.ascontiguousarray(np.asarray(cx, dtype=np.complex64).view(np.float32))
        h = C.c_void_p()
        ctx.check(lib().fhs_gmat_create_separable_host(
            ctx.h, M, Kty, Ktx, Ky, Kx, anchor.ctypes.data_as(C.POINTER(C.c_int32)),
            cyf.ctypes.data_as(C.POINTER(C.c_float)), cxf.ctypes.data_as(C.POINTER(C.c_float)), C.byref(h)),
            'gmat_create_separable')
        g = cls(ctx, h, M, Kty * Ktx, M * Ky * Kx)
        g.Ky, g.Kx = Ky, Kx
        return g

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_gmat_destroy(h)
            except Exception:
                pass
            self.h = None


class NufftPlan:
    """Scaled oversampled FFT ``A`` / ``At`` (``fhs_nufft``)."""

    def __init__(self, ctx, Ny, Nx, Kty, Ktx, sy, sx):
        self.ctx = ctx
        self.Ny, self.Nx, self.Kty, self.Ktx = Ny, Nx, Kty, Ktx
        self.No = Kty * Ktx
        sy = np.ascontiguousarray(sy, dtype=np.float64)
        sx = np.ascontiguousarray(sx, dtype=np.float64)
        h = C.c_void_p()
        ctx.check(lib().fhs_nufft_create_host(
            ctx.h, Ny, Nx, Kty, Ktx, sy.ctypes.data_as(C.POINTER(C.c_double)),
            sx.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)), 'nufft_create')
        self.h = h

    def A(self, x):
        """Fx (L, No) complex64 from x (L, Nx, Ny) or (Nx, Ny) float32."""
        squeeze = x.dim() == 2
        xx = x.reshape(-1, self.Nx, self.Ny)
        _require(xx, torch.float32, self.ctx.device, 'A: x')
        L = xx.shape[0]
        Fx = torch.empty((L, self.No), dtype=torch.complex64, device=self.ctx.device)
        self.ctx.check(lib().fhs_apply_scaled_fourier_transform(self.ctx.h, self.h, _ptr(xx), L, _ptr(Fx)),
                       'apply_scaled_fourier_transform')
        return Fx[0] if squeeze else Fx

    def At_real(self, g, scale=None):
        """scale[l] * real(At(g[l])) as (L, Nx, Ny) float32."""
        squeeze = g.dim() == 1
        gg = g.reshape(-1, self.No)
        _require(gg, torch.complex64, self.ctx.device, 'At: g')
        L = gg.shape[0]
        out = torch.empty((L, self.Nx, self.Ny), dtype=torch.float32, device=self.ctx.device)
        sc = None
        if scale is not None:
            sc_arr = (C.c_float * L)(*[float(s) for s in np.broadcast_to(np.asarray(scale), (L,))])
            sc = sc_arr
        self.ctx.check(lib().fhs_apply_scaled_fourier_adjoint(self.ctx.h, self.h, _ptr(gg), L, sc, _ptr(out)),
                       'apply_scaled_fourier_adjoint')
        return out[0] if squeeze else out

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_nufft_destroy(h)
            except Exception:
                pass
            self.h = None
