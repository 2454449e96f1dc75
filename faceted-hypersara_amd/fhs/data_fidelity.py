"""Data-fidelity term of L channels: ``update_dual_data_fidelity`` on the GPU.

The reference keeps cells ``{L}{nblocks}`` of M_b-vectors; here every
M-vector of all channels and blocks is one stacked complex64 tensor of
length ``Mtot`` (channel-major, blocks in order) and ``DataOperator`` holds the
per-channel G, the shared scaled FFT and the block layout.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import lib
from .core import GMatrix, NufftPlan, _ptr, _require


class DataOperator:
    """``fhs_data``: G{l} (row-stacked blocks) + A/At for L channels.

    block_sizes: list (per channel) of block row counts."""

    def __init__(self, ctx, A: NufftPlan, G, block_sizes):
        self.ctx, self.A = ctx, A
        self.G = list(G)
        self.L = len(self.G)
        if len(block_sizes) != self.L:
            raise ValueError('DataOperator: one block list per channel')
        starts, chans = [0], []
        self.row_offset = [0]
        for l, (g, bs) in enumerate(zip(self.G, block_sizes)):
            if int(sum(bs)) != g.M:
                raise ValueError(f'DataOperator: channel {l} blocks sum to {sum(bs)}, G has {g.M} rows')
            for b in bs:
                starts.append(starts[-1] + int(b))
                chans.append(l)
            self.row_offset.append(self.row_offset[-1] + g.M)
        self.nblocks = len(chans)
        self.Mtot = self.row_offset[-1]
        self.block_start = np.asarray(starts, dtype=np.int64)
        self.block_channel = np.asarray(chans, dtype=np.int32)
        self.block_sizes = [list(map(int, b)) for b in block_sizes]
        arr = (C.c_void_p * self.L)(*[g.h.value for g in self.G])
        h = C.c_void_p()
        bs = self.block_start
        bc = self.block_channel
        ctx.check(lib().fhs_data_create(ctx.h, A.h, self.L, arr, self.nblocks,
                                        bs.ctypes.data_as(C.POINTER(C.c_int64)),
                                        bc.ctypes.data_as(C.POINTER(C.c_int)), C.byref(h)),
                  'data_create')
        self.h = h
        self.row_bits = 32

    def set_row_precision(self, bits):
        """Storage precision of the G rows the data kernels read: 32 (as
        created) or 16 (an FP16 copy, half the row bytes; tile-gridded
        separable G only).  Call it between solves."""
        self.ctx.check(lib().fhs_data_set_row_precision(self.ctx.h, self.h, int(bits)), 'data_set_row_precision')
        self.row_bits = int(bits)

    def zeros(self):
        return torch.zeros(self.Mtot, dtype=torch.complex64, device=self.ctx.device)

    def init_fx_old(self, x, Lambda=None, flag_dimensionality_reduction=False):
        """Gx_old = G A(x): the state standing for Fx_old = A(x)."""
        _require(x, torch.float32, self.ctx.device, 'init_fx_old: x')
        gx = self.zeros()
        self.ctx.check(lib().fhs_data_init_fx_old(self.ctx.h, self.h, _ptr(x), _ptr(Lambda),
                                                  int(bool(flag_dimensionality_reduction)), _ptr(gx)),
                       'data_init_fx_old')
        return gx

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_data_destroy(h)
            except Exception:
                pass
            self.h = None


def update_dual_data_fidelity(v2, y, x, Fx_old, proj, A, At, G: DataOperator, W, pU, epsilon,
                              elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps,
                              sigma22, flag_dimensionality_reduction=False, Lambda=None):
    """``src/update_dual_data_fidelity.m:1-149`` (same arguments and outputs).

    ``G`` is a ``DataOperator`` (it carries A/At and the blocks; ``A``, ``At``
    and ``W`` are accepted for signature parity and must be ``G.A`` / None).
    ``Fx_old`` is the linearity-reduced state ``G A(x_prev)`` (see
    ``DataOperator.init_fx_old``).  v2, y, proj, Fx_old: stacked complex64
    (updated in place); x: (L, Nx, Ny) float32; pU, Lambda: float32 or None;
    epsilon: per-block bounds; sigma22: per-channel steps.
    Returns (v2, Ftx, Fx_old, proj, norm_res, global_norm_res, norm_epsilon).
    """
    ctx = G.ctx
    if A is not None and A is not G.A:
        raise ValueError('update_dual_data_fidelity: A must be the DataOperator plan')
    if W is not None:
        raise ValueError('update_dual_data_fidelity: W is the whole grid (pass None)')
    for t, name in ((v2, 'v2'), (y, 'y'), (Fx_old, 'Fx_old')):
        _require(t, torch.complex64, ctx.device, f'update_dual_data_fidelity: {name}')
        if t.numel() != G.Mtot:
            raise ValueError(f'update_dual_data_fidelity: {name} has {t.numel()} entries, expected {G.Mtot}')
    dr = bool(flag_dimensionality_reduction)
    if not dr:
        _require(proj, torch.complex64, ctx.device, 'update_dual_data_fidelity: proj')
    _require(x, torch.float32, ctx.device, 'update_dual_data_fidelity: x')
    if pU is not None:
        _require(pU, torch.float32, ctx.device, 'update_dual_data_fidelity: pU')
    if Lambda is not None:
        _require(Lambda, torch.float32, ctx.device, 'update_dual_data_fidelity: Lambda')
    eps = np.ascontiguousarray(np.broadcast_to(np.asarray(epsilon, dtype=np.float64), (G.nblocks,)))
    s22 = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma22, dtype=np.float32), (G.L,)))
    Ftx = torch.empty((G.L, G.A.Nx, G.A.Ny), dtype=torch.float32, device=ctx.device)
    norm_res = np.zeros(G.nblocks)
    gnr = C.c_double()
    neps = C.c_double()
    ctx.check(lib().fhs_update_dual_data_fidelity(
        ctx.h, G.h, _ptr(v2), _ptr(y), _ptr(x), _ptr(Fx_old), _ptr(proj), _ptr(pU),
        eps.ctypes.data_as(C.POINTER(C.c_double)), int(elipse_proj_max_iter), int(elipse_proj_min_iter),
        float(elipse_proj_eps), s22.ctypes.data_as(C.POINTER(C.c_float)), int(dr), _ptr(Lambda),
        _ptr(Ftx), norm_res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(gnr), C.byref(neps)),
        'update_dual_data_fidelity')
    return v2, Ftx, Fx_old, proj, norm_res, gnr.value, neps.value
