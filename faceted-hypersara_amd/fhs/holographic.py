"""Dimensionality-reduced (holographic) data: the visibilities of a channel
gridded onto the Fourier modes they touch,
``lib/operators/util_load_data_gen_dr_measurement_operator.m:160-225``.

Per channel, from its separable min-max G (rows = visibilities):

  H       = tril(G'G) on the active modes W, diagonal halved (``:196, :215-222``)
  W       = {p : (G'G)_pp > 0}                              (``:200-203``)
  Lambda  = 1 ./ sqrt(diag(G'G)) on W                        (``:206``)
  y_grid  = Lambda .* (G' y) on W                            (``:171, :209``)
  l2bound = || Lambda .* (G' n) on W ||, n ~ CN(0, 1)        (``:186-192, :211-212``)

built on the GPU (``fhs_gmat_create_holographic``).  The DR vectors are
kept on the whole oversampled grid (zero off W), so ``DataOperator`` takes H
like any square G and ``update_dual_data_fidelity(...,
flag_dimensionality_reduction=True, Lambda=...)`` runs the reference's DR
branch (``src/update_dual_data_fidelity.m:85-125``); ``compact`` gives the
reference's |W|-long vectors.  ``HolographicMeasurement`` is the DR form of
a whole measurement for the solvers.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import lib
from .core import GMatrix, _ptr, _require


class Holographic:
    """H, W and Lambda of one channel (its ``GMatrix`` G must have separable rows)."""

    def __init__(self, ctx, G):
        self.ctx, self.G = ctx, G
        h = C.c_void_p()
        ctx.check(lib().fhs_gmat_create_holographic(ctx.h, G.h, C.byref(h)), 'gmat_create_holographic')
        M, ncols, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        lib().fhs_gmat_info(h, C.byref(M), C.byref(ncols), C.byref(nnz))
        self.H = GMatrix(ctx, h, M.value, ncols.value, nnz.value)
        self.No = M.value
        self.Lambda = torch.empty(self.No, dtype=torch.float32, device=ctx.device)
        n = C.c_int64()
        ctx.check(lib().fhs_gmat_holographic_info(ctx.h, h, C.byref(n), _ptr(self.Lambda)),
                  'gmat_holographic_info')
        self.n_modes = n.value
        self._W = None

    @property
    def W(self):
        """Active modes: ascending grid indices (the reference's logical mask ``W``)."""
        if self._W is None:
            self._W = np.flatnonzero(self.Lambda.cpu().numpy() > 0)
        return self._W

    def grid(self, y):
        """Lambda .* (G' y) on the whole grid (complex64, zero off W)."""
        _require(y, torch.complex64, self.ctx.device, 'holographic grid: y')
        if y.numel() != self.G.M:
            raise ValueError(f'holographic grid: y has {y.numel()} entries, G has {self.G.M} rows')
        out = torch.empty(self.No, dtype=torch.complex64, device=self.ctx.device)
        self.ctx.check(lib().fhs_holographic_grid(self.ctx.h, self.G.h, self.H.h, _ptr(y.contiguous()), _ptr(out)),
                       'holographic_grid')
        return out

    def compact(self, v):
        """The reference's |W|-long vector of a grid-sized one."""
        return v[torch.from_numpy(self.W).to(v.device)]


class HolographicMeasurement:
    """The DR form of a ``solvers.Measurement``, for the solvers
    (``imaging/imaging.m:619-631``: one data block per channel after
    visibility gridding, no preconditioning, ``aW = 1``,
    ``util_load_data_gen_dr_measurement_operator.m:111``):

      D        DataOperator over the holographic H of every channel
      y        gridded data, stacked (grid-sized per channel, zero off W)
      Lambda   stacked weights (``fhs_pdfb_buffers.Lambda``)
      epsilon  l2 bounds of a seeded unit complex Gaussian noise vector
               gridded the same way; noise_std its per-channel std
               (``imaging/imaging.m:628-631``)
      forward / adjoint   the DR measurement operator Lambda (H + H') A and
               its adjoint (``HS_forward_operator_G.m`` / ``HS_adjoint_operator_G.m``)

    ``meas`` keeps the visibilities' G and stays referenced."""

    dr = True
    pU = None

    def __init__(self, meas, noise_seed=0):
        self.meas = meas
        self.ctx, self.problem = meas.ctx, meas.problem
        self.L, self.Ny, self.Nx = meas.L, meas.Ny, meas.Nx
        self.plan = meas.plan
        off = meas.D.row_offset
        ys = [meas.y[off[l]:off[l + 1]].contiguous() for l in range(self.L)]
        self.forms, self.y, self.epsilon, self.noise_std, self.noise_host = [], [], [], [], []
        gen = torch.Generator(device='cpu').manual_seed(int(noise_seed))
        for Gl, yl in zip(meas.G, ys):
            f = Holographic(self.ctx, Gl)
            self.forms.append(f)
            self.y.append(f.grid(yl))
            n = (torch.randn(Gl.M, generator=gen, dtype=torch.float64)
                 + 1j * torch.randn(Gl.M, generator=gen, dtype=torch.float64)) / np.sqrt(2.0)
            self.noise_host.append(n.numpy())
            gn = f.compact(f.grid(n.to(torch.complex64).to(self.ctx.device))).to(torch.complex128)
            self.epsilon.append(float(torch.linalg.vector_norm(gn).item()))
            d = gn - gn.mean()
            self.noise_std.append(float(torch.sqrt((d.abs() ** 2).sum() / max(d.numel() - 1, 1)).item()))
        from .data_fidelity import DataOperator
        self.D = DataOperator(self.ctx, self.plan, [f.H for f in self.forms], [[f.No] for f in self.forms])
        self.y = torch.cat(self.y)
        self.Lambda = torch.cat([f.Lambda for f in self.forms])
        self.epsilon = np.asarray(self.epsilon)
        self.noise_std = np.asarray(self.noise_std)

    def forward(self, x, lam=None):
        """Stacked Lambda .* (H + H') A x_l (``lam`` is ignored: no preconditioning)."""
        out = torch.empty(self.D.Mtot, dtype=torch.complex64, device=self.ctx.device)
        self.ctx.check(lib().fhs_data_forward_dr(self.ctx.h, self.D.h, _ptr(x), _ptr(self.Lambda), _ptr(out)),
                       'data_forward_dr')
        return out

    def adjoint(self, y, lam=None, scale=None):
        out = torch.empty((self.L, self.Nx, self.Ny), dtype=torch.float32, device=self.ctx.device)
        sc = None if scale is None else (C.c_float * self.L)(*[float(s) for s in scale])
        self.ctx.check(lib().fhs_data_adjoint_dr(self.ctx.h, self.D.h, _ptr(y), _ptr(self.Lambda), sc, _ptr(out)),
                       'data_adjoint_dr')
        return out

    def op_norm(self, precond=False, tol=1e-8, max_iter=200, seed=0):
        """Per-channel squared norm of the DR operator (``util_operator_norm.m``
        with flag_visibility_gridding; precond changes nothing, aW = 1)."""
        from .solvers import Measurement
        return Measurement.op_norm(self, False, tol, max_iter, seed)

    def residual_images(self, x):
        """``src/compute_residual_images.m:39-103``, DR branch."""
        from .solvers import Measurement
        return Measurement.residual_images(self, x)
