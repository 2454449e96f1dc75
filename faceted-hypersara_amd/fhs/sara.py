"""SARA dictionary (Daubechies db1..db8 + Dirac) on the GPU.

Filters are computed here in float64 (Daubechies' construction: the
half-band product filter (1+z)^N (1+z^-1)^N Q(z) factored with the zeros of
Q inside the unit circle, extremal phase), then uploaded; the transforms run
in libfhs (``fhs_sara_*``).
"""
import ctypes as C
from math import comb

import numpy as np
import torch

from ._lib import lib
from .core import _ptr, _require

DEFAULT_BASES = ('db1', 'db2', 'db3', 'db4', 'db5', 'db6', 'db7', 'db8', 'self')
DEFAULT_NLEVEL = 4


def daubechies_scaling_filter(N):
    """dbN reconstruction low-pass filter (2N taps, sum sqrt(2), energy first)."""
    if N == 1:
        return np.full(2, 1.0 / np.sqrt(2.0))
    # Q(z) z^{N-1} as a polynomial in z: sum_k C(N-1+k,k) ((2 - z - 1/z)/4)^k z^{N-1}
    q = np.zeros(2 * N - 1)
    base = np.array([-0.25, 0.5, -0.25])          # (2 - z - 1/z)/4 times z, powers z^2..z^0
    pk = np.array([1.0])
    for k in range(N):
        term = np.polynomial.polynomial.polymul(np.zeros(N - 1 - k).tolist() + [1.0], pk[::-1])
        q[:len(term)] += comb(N - 1 + k, k) * np.asarray(term)
        pk = np.convolve(pk, base)
    roots = np.polynomial.polynomial.polyroots(q)
    inner = roots[np.abs(roots) < 1]
    h = np.real(np.poly(inner))
    for _ in range(N):
        h = np.convolve(h, [1.0, 1.0])
    h *= np.sqrt(2.0) / h.sum()
    if abs(h[0]) < abs(h[-1]):
        h = h[::-1]
    return h


class SaraDictionary:
    """``[Psi, Psit] = op_sp_wlt_basis(wavelet, nlevel, Ny, Nx)`` on the GPU."""

    def __init__(self, ctx, Ny, Nx, bases=DEFAULT_BASES, nlevel=DEFAULT_NLEVEL):
        self.ctx, self.Ny, self.Nx, self.nlevel = ctx, Ny, Nx, nlevel
        bases = list(bases)
        dirac = 'self' in bases
        if dirac and bases[-1] != 'self':
            raise ValueError("SARA: the Dirac basis 'self' must come last")
        db = [b for b in bases if b != 'self']
        self.bases = tuple(bases)
        self.filters = [daubechies_scaling_filter(int(b[2:])) for b in db]
        lens = np.asarray([len(f) for f in self.filters], dtype=np.int32)
        lo = np.ascontiguousarray(np.concatenate(self.filters) if self.filters else np.zeros(1))
        h = C.c_void_p()
        ctx.check(lib().fhs_sara_create_host(
            ctx.h, Ny, Nx, nlevel, len(db), lens.ctypes.data_as(C.POINTER(C.c_int)),
            lo.ctypes.data_as(C.POINTER(C.c_double)), int(dirac), C.byref(h)), 'sara_create')
        self.h = h
        s = C.c_int64()
        ctx.check(lib().fhs_sara_info(h, C.byref(s)), 'sara_info')
        self.s = s.value

    def psit(self, x, out=None, accumulate=False):
        """(L, s) coefficients of x (L, Nx, Ny)."""
        xx = x.reshape(-1, self.Nx, self.Ny)
        _require(xx, torch.float32, self.ctx.device, 'Psit: x')
        L = xx.shape[0]
        if out is None:
            out = torch.empty((L, self.s), dtype=torch.float32, device=self.ctx.device)
            accumulate = False
        self.ctx.check(lib().fhs_sara_analysis(self.ctx.h, self.h, _ptr(xx), L, _ptr(out), int(accumulate)),
                       'sara_analysis')
        return out

    def psi(self, c, scale=1.0):
        cc = c.reshape(-1, self.s)
        _require(cc, torch.float32, self.ctx.device, 'Psi: c')
        L = cc.shape[0]
        out = torch.empty((L, self.Nx, self.Ny), dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(lib().fhs_sara_synthesis(self.ctx.h, self.h, _ptr(cc), L, float(scale), _ptr(out)),
                       'sara_synthesis')
        return out

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_sara_destroy(h)
            except Exception:
                pass
            self.h = None


class SaraFacets:
    """Faceted SARA transform (``fhs_sara_facets``): the role of the
    reference's ``sdwt2_setup`` / ``sdwt2_sara_faceting`` /
    ``isdwt2_sara_faceting`` (``src/fhs/facetHyperSARA.m:325-326``,
    ``src/fhs/fhs_update_dual_sparsity.m:71-93``).  ``ybounds`` /
    ``xbounds``: facet boundaries from 0 to the image size; facet
    q = qx * Qy + qy.  Every facet reads an extended image (``ext(q)``: its
    pixels plus a top/left border, zero outside the image) and owns an exact
    share of the whole-image coefficients."""

    def __init__(self, ctx, Ny, Nx, ybounds, xbounds, bases=DEFAULT_BASES, nlevel=DEFAULT_NLEVEL):
        self.ctx, self.Ny, self.Nx, self.nlevel = ctx, Ny, Nx, nlevel
        bases = list(bases)
        dirac = 'self' in bases
        if dirac and bases[-1] != 'self':
            raise ValueError("SARA: the Dirac basis 'self' must come last")
        db = [b for b in bases if b != 'self']
        filt = [daubechies_scaling_filter(int(b[2:])) for b in db]
        lens = np.asarray([len(f) for f in filt], dtype=np.int32)
        lo = np.ascontiguousarray(np.concatenate(filt) if filt else np.zeros(1))
        yb = np.ascontiguousarray(ybounds, dtype=np.int32)
        xb = np.ascontiguousarray(xbounds, dtype=np.int32)
        self.Qy, self.Qx = len(yb) - 1, len(xb) - 1
        self.Q = self.Qy * self.Qx
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))   # noqa: E731
        h = C.c_void_p()
        ctx.check(lib().fhs_sara_facets_create(ctx.h, Ny, Nx, nlevel, len(db), ip(lens),
                                               lo.ctypes.data_as(C.POINTER(C.c_double)), int(dirac), self.Qy,
                                               self.Qx, ip(yb), ip(xb), C.byref(h)), 'sara_facets_create')
        self.h = h
        nrows = len(db) * (nlevel + 1) + int(dirac)
        self._ext, self._n, self.ncoefs = [], [], []
        for q in range(self.Q):
            e = np.zeros(4, dtype=np.int32)
            n = C.c_int64()
            rows = np.zeros(2 * nrows, dtype=np.int32)
            lib().fhs_sara_facets_info(h, q, ip(e), C.byref(n), ip(rows))
            self._ext.append(tuple(int(v) for v in e))
            self._n.append(n.value)
            self.ncoefs.append(rows.reshape(nrows, 2))

    def ext(self, q):
        """(row origin, col origin, rows, cols) of facet q's extended image."""
        return self._ext[q]

    def n_coefs(self, q):
        return self._n[q]

    def extended(self, q, x):
        """Facet q's extended image cut from a whole cube x (L, Nx, Ny), zero
        outside the image: (L, enx, eny)."""
        ey0, ex0, eny, enx = self._ext[q]
        xx = x.reshape(-1, self.Nx, self.Ny)
        out = torch.zeros((xx.shape[0], enx, eny), dtype=torch.float32, device=xx.device)
        y0, x0 = max(ey0, 0), max(ex0, 0)
        out[:, x0 - ex0:, y0 - ey0:] = xx[:, x0:ex0 + enx, y0:ey0 + eny]
        return out

    def add_extended(self, q, g, out):
        """out (L, Nx, Ny) += facet q's extended-region image g, cropped."""
        ey0, ex0, eny, enx = self._ext[q]
        y0, x0 = max(ey0, 0), max(ex0, 0)
        out[:, x0:ex0 + enx, y0:ey0 + eny] += g[:, x0 - ex0:, y0 - ey0:]
        return out

    def analysis(self, q, x_ext):
        """z = sdwt2_sara_faceting(x_ext): (L, n_coefs(q))."""
        _require(x_ext, torch.float32, self.ctx.device, 'facet analysis: x_ext')
        ey0, ex0, eny, enx = self._ext[q]
        L = x_ext.numel() // (eny * enx)
        if L * eny * enx != x_ext.numel():
            raise ValueError('facet analysis: x_ext must be (L, enx, eny)')
        z = torch.empty((L, self._n[q]), dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(lib().fhs_sara_facet_analysis(self.ctx.h, self.h, q, _ptr(x_ext), L, _ptr(z)),
                       'sara_facet_analysis')
        return z

    def synthesis(self, q, z, scale=1.0):
        """scale * isdwt2_sara_faceting(z): (L, enx, eny)."""
        _require(z, torch.float32, self.ctx.device, 'facet synthesis: z')
        ey0, ex0, eny, enx = self._ext[q]
        L = z.numel() // self._n[q]
        g = torch.empty((L, enx, eny), dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(lib().fhs_sara_facet_synthesis(self.ctx.h, self.h, q, _ptr(z), L, float(scale), _ptr(g)),
                       'sara_facet_synthesis')
        return g

    def update_dual(self, q, v1, x_ext, weights1, beta1):
        """[v1, g] = fhs_update_dual_sparsity(v1, x_facet, weights1, beta1, ...)
        for facet q (``src/fhs/fhs_update_dual_sparsity.m:65-93``): v1
        (L, n_coefs(q)) updated in place; returns (v1, g_ext)."""
        _require(v1, torch.float32, self.ctx.device, 'facet dual: v1')
        _require(x_ext, torch.float32, self.ctx.device, 'facet dual: x_ext')
        ey0, ex0, eny, enx = self._ext[q]
        L = x_ext.numel() // (eny * enx)
        if v1.numel() != L * self._n[q]:
            raise ValueError('facet dual: v1 must be (L, n_coefs)')
        if weights1 is not None:
            _require(weights1, torch.float32, self.ctx.device, 'facet dual: weights1')
        g = torch.empty((L, enx, eny), dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(lib().fhs_update_dual_sparsity_facet(self.ctx.h, self.h, q, _ptr(v1), _ptr(x_ext), L,
                                                            _ptr(weights1), float(beta1), _ptr(g)),
                       'update_dual_sparsity_facet')
        return v1, g

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_sara_facets_destroy(h)
            except Exception:
                pass
            self.h = None


def hs_update_dual_sparsity_distributed(v1, Psit, Psi, xhat, weights1, beta1, sigma1):
    """``[v1, g1] = hs_update_dual_sparsity_distributed(v1, Psit, Psi, xhat,
    weights1, beta1, sigma1)`` — ``src/hs/hs_update_dual_sparsity_distributed.m:4-17``.

    ``Psit``/``Psi``: the same ``SaraDictionary``; v1 (L, s) float32 updated in
    place; weights1 (s,) float32 or None; returns (v1, g1)."""
    S = Psit
    if Psi is not None and Psi is not S:
        raise ValueError('Psit and Psi must be the same SaraDictionary')
    ctx = S.ctx
    L = xhat.reshape(-1, S.Nx, S.Ny).shape[0]
    _require(v1, torch.float32, ctx.device, 'update_dual_sparsity: v1')
    if v1.numel() != L * S.s:
        raise ValueError('update_dual_sparsity: v1 must be (L, s)')
    _require(xhat, torch.float32, ctx.device, 'update_dual_sparsity: xhat')
    if weights1 is not None:
        _require(weights1, torch.float32, ctx.device, 'update_dual_sparsity: weights1')
    g1 = torch.empty((L, S.Nx, S.Ny), dtype=torch.float32, device=ctx.device)
    ctx.check(lib().fhs_update_dual_sparsity(ctx.h, S.h, _ptr(v1), _ptr(xhat), L, _ptr(weights1),
                                             float(beta1), float(sigma1), _ptr(g1)),
              'update_dual_sparsity')
    return v1, g1


def compute_sara_prior(x, Psit, n_wavelet_coefficients=None):
    """``f = compute_sara_prior(x, Psit, n_wavelet_coefficients)`` —
    ``src/compute_sara_prior.m:33-41``: sum over coefficients of the l2 norm
    across channels of Psit(x)."""
    S = Psit
    if n_wavelet_coefficients is not None and int(n_wavelet_coefficients) != S.s:
        raise ValueError(f'compute_sara_prior: dictionary has {S.s} coefficients')
    xx = x.reshape(-1, S.Nx, S.Ny)
    _require(xx, torch.float32, S.ctx.device, 'compute_sara_prior: x')
    out = C.c_double()
    S.ctx.check(lib().fhs_compute_sara_prior(S.ctx.h, S.h, _ptr(xx), xx.shape[0], C.byref(out)),
                'compute_sara_prior')
    return out.value
