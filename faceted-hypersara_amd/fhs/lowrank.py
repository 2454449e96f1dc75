"""Low-rank (weighted nuclear norm) prior on the GPU, global or faceted."""
import ctypes as C

import numpy as np
import torch

from ._lib import lib
from .core import _ptr, _require


def facet_geometry(Ny, Nx, Qy, Qx, overlap_fraction=(0.0, 0.0)):
    """Overlapping nuclear-norm facets of an Ny x Nx image (facet q = qx*Qy + qy).

    Plain facet boundaries are round(i N / Q) (MATLAB rounding) as the
    reference's ``split_range`` gives them (``src/fhs/facetHyperSARA.m:297-321``);
    every facet after the first along an axis starts ``floor(f N / Q)``
    pixels earlier (``get_overlap_size``, ``imaging/imaging.m:421``), and the
    apodisation window is a separable ramp that rises over the leading
    overlap strip and falls over the trailing one, so overlapping windows sum
    to one (the 'triangular' ``generate_weights`` of
    ``src/fhs/fhs_setup_priors.m:189``).  Returns the dict ``LowRankPrior``
    takes: Qy, Qx, oy, ox, py, px, windows, overlap."""
    def axis(N, Q, f):
        d = int(np.floor(f * N / Q)) if Q > 1 else 0
        edge = [int(np.floor(i * N / Q + 0.5)) for i in range(Q + 1)]
        spans = [(max(edge[q] - (d if q > 0 else 0), 0), edge[q + 1]) for q in range(Q)]
        wins = []
        for q, (a, b) in enumerate(spans):
            w = np.ones(b - a)
            lead = d if q > 0 else 0
            trail = d if q < Q - 1 else 0
            w[:lead] = np.arange(1, lead + 1) / (lead + 1)
            if trail:
                w[b - a - trail:] = 1.0 - np.arange(1, trail + 1) / (trail + 1)
            wins.append(w)
        return d, spans, wins

    dy, sy, wy = axis(Ny, Qy, overlap_fraction[0])
    dx, sx, wx = axis(Nx, Qx, overlap_fraction[1])
    g = {'Qy': Qy, 'Qx': Qx, 'oy': [], 'ox': [], 'py': [], 'px': [], 'windows': [], 'overlap': (dy, dx)}
    for qx in range(Qx):
        for qy in range(Qy):
            g['oy'].append(sy[qy][0])
            g['ox'].append(sx[qx][0])
            g['py'].append(sy[qy][1] - sy[qy][0])
            g['px'].append(sx[qx][1] - sx[qx][0])
            g['windows'].append(np.outer(wy[qy], wx[qx]))
    return g


class LowRankPrior:
    """Facet layout of the nuclear-norm prior (``fhs_lowrank``).

    facets: None (one facet = the whole image, HyperSARA) or a dict with
    Qy, Qx and per-facet arrays oy, ox, py, px (facet q = qx*Qy + qy) and
    optional ``windows`` (list of (py, px) arrays, column-major upload)."""

    def __init__(self, ctx, Ny, Nx, L, facets=None):
        self.ctx, self.Ny, self.Nx, self.L = ctx, Ny, Nx, L
        if facets is None:
            facets = {'Qy': 1, 'Qx': 1, 'oy': [0], 'ox': [0], 'py': [Ny], 'px': [Nx]}
        self.Qy, self.Qx = int(facets['Qy']), int(facets['Qx'])
        self.Q = self.Qy * self.Qx
        arr = {k: np.ascontiguousarray(facets[k], dtype=np.int32) for k in ('oy', 'ox', 'py', 'px')}
        self.geom = arr
        win = facets.get('windows')
        wptr = None
        if win is not None:
            flat = np.concatenate([np.asarray(w, dtype=np.float32).ravel(order='F') for w in win])
            self._win = np.ascontiguousarray(flat)
            wptr = self._win.ctypes.data_as(C.POINTER(C.c_float))
        h = C.c_void_p()
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))   # noqa: E731
        ctx.check(lib().fhs_lowrank_create_host(ctx.h, Ny, Nx, L, self.Qy, self.Qx, ip(arr['oy']), ip(arr['ox']),
                                                ip(arr['py']), ip(arr['px']), wptr, C.byref(h)),
                  'lowrank_create')
        self.h = h
        n = C.c_int64()
        ctx.check(lib().fhs_lowrank_info(h, C.byref(n)), 'lowrank_info')
        self.v0_size = n.value

    GRAM_MODES = {'fp64': 0, 'bf16x3': 3, 'bf16x2': 2}

    def set_gram_precision(self, mode):
        """'fp64' (default): FP64 Gram of exact products; 'bf16x3' / 'bf16x2':
        BF16 split parts (hi + mid + lo / hi + lo) multiplied on the matrix
        cores with FP32 sums (``fhs_lowrank_set_gram_precision``)."""
        if mode not in self.GRAM_MODES:
            raise ValueError(f'gram precision must be one of {sorted(self.GRAM_MODES)}')
        self.ctx.check(lib().fhs_lowrank_set_gram_precision(self.h, self.GRAM_MODES[mode]),
                       'lowrank_set_gram_precision')
        self.gram = mode

    def zeros_v0(self):
        return torch.zeros(self.v0_size, dtype=torch.float32, device=self.ctx.device)

    def singular_values(self, x):
        """(Q, L) singular values (descending) of w_q .* x_q."""
        _require(x, torch.float32, self.ctx.device, 'singular_values: x')
        sv = np.zeros(self.Q * self.L)
        self.ctx.check(lib().fhs_lowrank_singular_values(self.ctx.h, self.h, _ptr(x),
                                                         sv.ctypes.data_as(C.POINTER(C.c_double))),
                       'lowrank_singular_values')
        return sv.reshape(self.Q, self.L)

    def update_dual(self, v0, xhat, weights0, beta0, sigma0):
        """``[v0, g0] = fhs_update_dual_lowrankness(v0, xhat, apodization_window,
        weights0, beta0)`` for all facets; g0 = sigma0 * overlap-added w .* v0
        (``src/fhs/fhs_update_dual_lowrankness.m:40-45``,
        ``src/hs/hs_update_dual_lowrankness_serial.m:5-10``)."""
        _require(v0, torch.float32, self.ctx.device, 'update_dual_lowrankness: v0')
        _require(xhat, torch.float32, self.ctx.device, 'update_dual_lowrankness: xhat')
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(weights0, dtype=np.float64), (self.Q, self.L)))
        g0 = torch.empty((self.L, self.Nx, self.Ny), dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(lib().fhs_update_dual_lowrankness(self.ctx.h, self.h, _ptr(v0), _ptr(xhat),
                                                         w.ctypes.data_as(C.POINTER(C.c_double)),
                                                         float(beta0), float(sigma0), _ptr(g0)),
                       'update_dual_lowrankness')
        return v0, g0

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_lowrank_destroy(h)
            except Exception:
                pass
            self.h = None


def nuclear_norm(x):
    """``f = nuclear_norm(x)`` — ``src/nuclear_norm.m:25-26``; x (L, Nx, Ny)."""
    from .core import default_context
    ctx = default_context(x.device.index or 0)
    _require(x, torch.float32, ctx.device, 'nuclear_norm: x')
    L, Nx, Ny = x.shape
    out = C.c_double()
    ctx.check(lib().fhs_nuclear_norm(ctx.h, _ptr(x), Ny, Nx, L, C.byref(out)), 'nuclear_norm')
    return out.value


def hs_update_dual_lowrankness_serial(v0, xhat, weights0, beta0, sigma0, prior: LowRankPrior):
    """``[v0, g0] = hs_update_dual_lowrankness_serial(v0, xhat, weights0,
    beta0, sigma0)`` — ``src/hs/hs_update_dual_lowrankness_serial.m:5-10``
    (``prior`` is the single-facet ``LowRankPrior`` holding the warm-start
    eigenvectors)."""
    return prior.update_dual(v0, xhat, weights0, beta0, sigma0)
