"""Measurement-operator construction: min-max NUFFT interpolation (host, FP64).

Builds, per channel, the scaling factors of the scaled FFT ``A`` and the
dense-window interpolation matrix ``G`` (anchor + Ky*Kx coefficients per
visibility) that ``op_p_nufft([{v} {u}], [Ny Nx], [Ky Kx], [oy*Ny ox*Nx],
[Ny/2 Nx/2], ...)`` returns in the reference
(``lib/operators/util_gen_measurement_operator_synth.m:85``).

Min-max interpolator (Fessler & Sutton, IEEE TSP 2003): with scaling factors
``s_n = sum_l a_l cos(gamma l (n - n0))`` (least-squares fit of the inverse
Kaiser-Bessel transform), the J coefficients of frequency w are
``T r(w)`` where ``T^-1_jk = sum_n s_n^2 cos(gamma (j-k)(n-n0))`` and
``r_j(w) = sum_l a_l/2 [D(t_j + gamma l) + D(t_j - gamma l)]``,
``t_j = gamma k_j - w``, ``D`` the Dirichlet kernel; phases move the
origin to pixel N/2.  The coefficients are computed here once per
visibility set; the iteration only reads them.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import lib
from .core import GMatrix, NufftPlan, _ptr

KB_ALPHA_PER_J = 2.34
SCALING_TERMS = 6


class AxisInterp:
    """1-D min-max interpolation setup for image size N on grid K."""

    def __init__(self, N, K, J):
        self.N, self.K, self.J = N, K, J
        self.gamma = 2 * np.pi / K
        self.n0 = (N - 1) / 2.0
        n = np.arange(N) - self.n0
        # inverse Kaiser-Bessel transform, then cosine-series fit
        a_kb = KB_ALPHA_PER_J * J
        arg = np.sqrt((a_kb ** 2 - (np.pi * J * n / K) ** 2).astype(complex))
        target = 1.0 / np.real(np.sinh(arg) / arg)
        C = np.cos(self.gamma * np.outer(n, np.arange(SCALING_TERMS)))
        a = np.linalg.lstsq(C, target, rcond=None)[0]
        self.alpha = a / a[0]
        self.s = C @ self.alpha
        d = np.subtract.outer(np.arange(J), np.arange(J))
        gram = np.cos(self.gamma * d[..., None] * n) @ (self.s ** 2)
        self.T = np.linalg.inv(gram)

    def _dirichlet(self, t):
        half = np.sin(0.5 * t)
        out = np.full(t.shape, float(self.N))
        ok = np.abs(half) >= 1e-12
        out[ok] = np.sin(0.5 * self.N * t[ok]) / half[ok]
        return out

    def coefficients(self, w):
        """(first grid index mod K, (M, J) complex coefficients) for frequencies w."""
        w = np.asarray(w, dtype=np.float64)
        k0 = np.floor(w / self.gamma - 0.5 * self.J)
        k = k0[:, None] + np.arange(1, self.J + 1)
        t = self.gamma * k - w[:, None]
        r = self.alpha[0] * self._dirichlet(t)
        for l in range(1, len(self.alpha)):
            r += 0.5 * self.alpha[l] * (self._dirichlet(t + self.gamma * l) + self._dirichlet(t - self.gamma * l))
        c = r @ self.T   # T symmetric
        phase = np.exp(1j * (self.gamma * k * self.n0 - w[:, None] * (self.n0 - self.N // 2)))
        return np.mod(k0 + 1, self.K).astype(np.int64), c * phase


class Nufft2D:
    """Scaled-FFT plan and per-channel G builder for an Ny x Nx image."""

    def __init__(self, Ny, Nx, Ky=7, Kx=7, oy=2, ox=2):
        self.Ny, self.Nx, self.Ky, self.Kx = Ny, Nx, Ky, Kx
        self.Kty, self.Ktx = oy * Ny, ox * Nx
        self.No = self.Kty * self.Ktx
        self.ay = AxisInterp(Ny, self.Kty, Ky)
        self.ax = AxisInterp(Nx, self.Ktx, Kx)

    def window(self, u, v, row_scale=None):
        """anchor (M,), coef (M, Ky*Kx) with tap jx*Ky + jy; G(m, :) scaled by row_scale."""
        ay, cy = self.ay.coefficients(v)
        ax, cx = self.ax.coefficients(u)
        anchor = ay + self.Kty * ax
        coef = (cx[:, :, None] * cy[:, None, :]).reshape(len(u), -1)
        if row_scale is not None:
            coef = coef * np.asarray(row_scale)[:, None]
        return anchor, coef

    def separable(self, u, v, row_scale=None):
        """anchor (M,), cy (M, Ky), cx (M, Kx): G(m, (jy, jx)) = cy[m, jy] cx[m, jx];
        row_scale folded into cx."""
        ay, cy = self.ay.coefficients(v)
        ax, cx = self.ax.coefficients(u)
        if row_scale is not None:
            cx = cx * np.asarray(row_scale)[:, None]
        return ay + self.Kty * ax, cy, cx

    def separable_gpu(self, ctx, u, v, row_scale=None):
        """``separable`` computed on the GPU (``fhs_nufft_interp_coefficients``,
        FP64 arithmetic): anchor int32 (M,), cy (M, Ky), cx (M, Kx) complex64,
        returned to the host for ``GMatrix.from_separable``."""
        M = len(u)
        dev = ctx.device
        ud = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64)).to(dev)
        vd = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
        rs = None if row_scale is None else torch.from_numpy(
            np.ascontiguousarray(np.broadcast_to(np.asarray(row_scale, dtype=np.float64), (M,)))).to(dev)
        anchor = torch.empty(M, dtype=torch.int32, device=dev)
        cy = torch.empty((M, self.Ky), dtype=torch.complex64, device=dev)
        cx = torch.empty((M, self.Kx), dtype=torch.complex64, device=dev)
        ay, ax = self.ay, self.ax
        pa = [np.ascontiguousarray(t, dtype=np.float64) for t in (ay.alpha, ay.T, ax.alpha, ax.T)]
        dp = [t.ctypes.data_as(C.POINTER(C.c_double)) for t in pa]
        ctx.check(lib().fhs_nufft_interp_coefficients(
            ctx.h, self.Ny, self.Kty, self.Ky, len(ay.alpha), dp[0], dp[1],
            self.Nx, self.Ktx, self.Kx, len(ax.alpha), dp[2], dp[3],
            _ptr(ud), _ptr(vd), _ptr(rs), M, _ptr(anchor), _ptr(cy), _ptr(cx)), 'nufft_interp_coefficients')
        return anchor.cpu().numpy(), cy.cpu().numpy(), cx.cpu().numpy()

    def separable_wproj_gpu(self, ctx, u, v, w, fov, S=4, c_energy=0.9999, g_energy=0.9999, row_scale=None,
                            dw=0.05):
        """W-projected separable rows on the GPU (``fhs_nufft_wproj_rows``;
        ``op_p_nufft_wproj_dde`` of the reference with the Fresnel w-phase):
        anchor int32 (M,), cy (M, Ky + 2S), cx (M, Kx + 2S) complex64.
        ``fov`` is the field of view in direction cosines (the reference's
        FoV = sin(pixelSize N), ``imaging/imaging.m:428-429``), w in
        wavelengths; the energy fractions are the reference's CEnergyL2 and
        GEnergyL2, split evenly over the two axes (square roots).  Raises if
        the kernels kept by CEnergyL2 reach beyond S grid points; with
        c_energy = 1 the kernels are cut to |d| <= S."""
        M = len(u)
        dev = ctx.device
        w = np.asarray(w, dtype=np.float64)
        wmax = float(np.max(np.abs(w))) if M else 0.0
        tabs = []
        for N, K in ((self.Ny, self.Kty), (self.Nx, self.Ktx)):
            tab, step, dmax = wkernel_table(N, K, fov, wmax, S, np.sqrt(c_energy), dw)
            if c_energy < 1.0 and dmax > S:
                raise ValueError(f'w-kernels reach {dmax} grid points (S = {S}): lower fov * |w|, or raise S')
            tabs.append(np.ascontiguousarray(tab))
        ud = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64)).to(dev)
        vd = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
        wd = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
        rs = None if row_scale is None else torch.from_numpy(
            np.ascontiguousarray(np.broadcast_to(np.asarray(row_scale, dtype=np.float64), (M,)))).to(dev)
        a0 = torch.empty(M, dtype=torch.int32, device=dev)
        cy0 = torch.empty((M, self.Ky), dtype=torch.complex64, device=dev)
        cx0 = torch.empty((M, self.Kx), dtype=torch.complex64, device=dev)
        ay, ax = self.ay, self.ax
        pa = [np.ascontiguousarray(t_, dtype=np.float64) for t_ in (ay.alpha, ay.T, ax.alpha, ax.T)]
        dp = [t_.ctypes.data_as(C.POINTER(C.c_double)) for t_ in pa]
        ctx.check(lib().fhs_nufft_interp_coefficients(
            ctx.h, self.Ny, self.Kty, self.Ky, len(ay.alpha), dp[0], dp[1],
            self.Nx, self.Ktx, self.Kx, len(ax.alpha), dp[2], dp[3],
            _ptr(ud), _ptr(vd), _ptr(rs), M, _ptr(a0), _ptr(cy0), _ptr(cx0)), 'nufft_interp_coefficients')
        a1 = torch.empty(M, dtype=torch.int32, device=dev)
        cy1 = torch.empty((M, self.Ky + 2 * S), dtype=torch.complex64, device=dev)
        cx1 = torch.empty((M, self.Kx + 2 * S), dtype=torch.complex64, device=dev)
        tp = [t_.ctypes.data_as(C.POINTER(C.c_double)) for t_ in tabs]
        ctx.check(lib().fhs_nufft_wproj_rows(
            ctx.h, self.Kty, self.Ktx, self.Ky, self.Kx, S, S, tabs[0].shape[0], dw, tp[0], tp[1],
            float(np.sqrt(g_energy)), float(np.sqrt(g_energy)), _ptr(wd), M, _ptr(a0), _ptr(cy0), _ptr(cx0),
            _ptr(a1), _ptr(cy1), _ptr(cx1)), 'nufft_wproj_rows')
        return a1.cpu().numpy(), cy1.cpu().numpy(), cx1.cpu().numpy()

    def plan(self, ctx):
        return NufftPlan(ctx, self.Ny, self.Nx, self.Kty, self.Ktx, self.ay.s, self.ax.s)

    def gmatrix(self, ctx, u, v, row_scale=None, layout='separable'):
        if layout == 'separable':
            anchor, cy, cx = self.separable(u, v, row_scale)
            return GMatrix.from_separable(ctx, anchor, cy, cx, self.Kty, self.Ktx)
        anchor, coef = self.window(u, v, row_scale)
        return GMatrix.from_window(ctx, anchor, coef, self.Kty, self.Ktx, self.Ky, self.Kx)


def extended_chirp_phase(t, w, fov):
    """Phase of the Fresnel w-chirp along one oversampled axis (direction
    cosine t, period 2 fov): pi w t^2 over the image (|t| <= fov/2); over the
    padding, where the scaled-FFT input is zero and any values are allowed,
    the even quartic pi w (2.5 h^2 - 2 u^2 + u^4 / (2 h^2)), h = fov/2,
    u = t - fov (mod 2 fov), which meets the chirp with equal value, slope
    and curvature at both image edges.  The smooth periodic extension keeps
    the kernel's Fourier transform compact (the zero-padded chirp has sinc
    tails)."""
    h = fov / 2
    tt = np.mod(t + fov, 2 * fov) - fov
    u = np.where(tt > 0, tt - fov, tt + fov)
    bridge = np.pi * w * (2.5 * h * h - 2 * u * u + u ** 4 / (2 * h * h))
    return np.where(np.abs(tt) <= h, np.pi * w * tt * tt, bridge)


_WTABLES = {}


def wkernel_table(N, K, fov, w_max, S, keep=1.0, dw=0.05):
    """W-kernels of one axis for w = 0, dw, 2 dw, ... >= w_max: per w the
    K-point DFT of the extended chirp (pixel n at (n - N/2) fov/N, the
    padding after the image), cut to its largest entries holding ``keep``
    of its energy (the reference's CEnergyL2 per axis; the energy is K^2 by
    Parseval), then to offsets |d| <= S.  Returns (table (nb, 2S+1)
    complex128, dw, largest kept |d| before the window cut)."""
    key = (N, K, float(fov), float(w_max), S, float(keep), float(dw))
    if key in _WTABLES:
        return _WTABLES[key]
    nb = int(np.ceil(w_max / dw)) + 2
    t = (np.arange(K) - N // 2) * (fov / N)
    d = np.arange(-S, S + 1) % K
    table = np.zeros((nb, 2 * S + 1), dtype=np.complex128)
    dmax = 0
    dk = np.arange(K)
    dk = np.where(dk > K // 2, dk - K, dk)
    for b0 in range(0, nb, 256):
        ws = dw * np.arange(b0, min(nb, b0 + 256))
        C = np.fft.fft(np.exp(1j * extended_chirp_phase(t[None, :], ws[:, None], fov)), axis=1)
        if keep < 1.0:
            e = np.abs(C) ** 2
            order = np.argsort(-e, axis=1, kind='stable')
            cum = np.cumsum(np.take_along_axis(e, order, axis=1), axis=1)
            n_keep = np.sum(cum < keep * (K * K), axis=1) + 1
            mask = np.zeros(C.shape, dtype=bool)
            for r in range(len(ws)):
                mask[r, order[r, :n_keep[r]]] = True
            C = np.where(mask, C, 0)
            dmax = max(dmax, int(np.max(np.abs(np.where(mask, dk[None, :], 0)))))
        else:
            dmax = K // 2
        table[b0:b0 + len(ws)] = C[:, d]
    out = (table, dw, dmax)
    _WTABLES[key] = out
    return out


def uniform_weights_gpu(ctx, u, v, Kty, Ktx):
    """``uniform_weights`` on the GPU (``fhs_precond_uniform_weights``):
    float32 device tensor of 1/(points in the cell)."""
    dev = ctx.device
    ud = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64)).to(dev)
    vd = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
    out = torch.empty(len(ud), dtype=torch.float32, device=dev)
    ctx.check(lib().fhs_precond_uniform_weights(ctx.h, _ptr(ud), _ptr(vd), len(ud), int(Kty), int(Ktx), _ptr(out)),
              'precond_uniform_weights')
    return out


def uniform_weights(u, v, Kty, Ktx):
    """Preconditioning weights aW = 1/(number of uv points in the same cell of
    the oversampled grid): util_gen_preconditioning_matrix with
    gen_uniform_weight_matrix = true (``imaging/default_parameters.json:47-51``)."""
    iy = np.floor((np.asarray(v) + np.pi) * Kty / (2 * np.pi)).astype(np.int64) % Kty
    ix = np.floor((np.asarray(u) + np.pi) * Ktx / (2 * np.pi)).astype(np.int64) % Ktx
    cell = iy + Kty * ix
    return 1.0 / np.bincount(cell, minlength=Kty * Ktx)[cell]
