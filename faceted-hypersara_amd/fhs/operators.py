"""The reference's hot-path functions under their MATLAB names.

Each function keeps the argument order and meaning of the ``.m`` file it
replaces and runs on the GPU through the C ABI; there is no CPU path.
Tensors follow the layout of ``fhs.core`` (MATLAB column-major memory).
"""
import ctypes as C

import torch

from ._lib import lib
from .core import GMatrix, _ptr, _require


def apply_direct_operator(Fx, G: GMatrix, Lambda=None):
    """``y = Lambda .* (G * Fx)`` — ``src/apply_direct_operator.m:1-22``."""
    ctx = G.ctx
    _require(Fx, torch.complex64, ctx.device, 'apply_direct_operator: Fx')
    if Fx.numel() != G.ncols:
        raise ValueError(f'apply_direct_operator: Fx has {Fx.numel()} entries, G has {G.ncols} columns')
    lam = None
    if Lambda is not None:
        lam = _require(Lambda, torch.float32, ctx.device, 'apply_direct_operator: Lambda')
        if lam.numel() != G.M:
            raise ValueError('apply_direct_operator: Lambda must have one entry per row of G')
    y = torch.empty(G.M, dtype=torch.complex64, device=ctx.device)
    ctx.check(lib().fhs_apply_direct_operator(ctx.h, G.h, _ptr(Fx), _ptr(lam), _ptr(y)),
              'apply_direct_operator')
    return y


def apply_adjoint_operator(y, G: GMatrix, Lambda=None):
    """``Fx = G' * (Lambda .* y)`` — ``src/apply_adjoint_operator.m:1-23``."""
    ctx = G.ctx
    _require(y, torch.complex64, ctx.device, 'apply_adjoint_operator: y')
    if y.numel() != G.M:
        raise ValueError(f'apply_adjoint_operator: y has {y.numel()} entries, G has {G.M} rows')
    lam = None
    if Lambda is not None:
        lam = _require(Lambda, torch.float32, ctx.device, 'apply_adjoint_operator: Lambda')
        if lam.numel() != G.M:
            raise ValueError('apply_adjoint_operator: Lambda must have one entry per row of G')
    Fx = torch.empty(G.ncols, dtype=torch.complex64, device=ctx.device)
    ctx.check(lib().fhs_apply_adjoint_operator(ctx.h, G.h, _ptr(y), _ptr(lam), _ptr(Fx)),
              'apply_adjoint_operator')
    return Fx


def apply_scaled_fourier_transform(x, A, No):
    """``Fx(:, l) = A(x(:, :, l))`` — ``src/apply_scaled_fourier_transform.m:21-24``.

    x: (L, Nx, Ny) float32; A: ``NufftPlan``; returns (L, No) complex64."""
    if A.No != No:
        raise ValueError(f'apply_scaled_fourier_transform: plan grid {A.No} != No {No}')
    return A.A(x)


def proj_l2ball(x, eps, y):
    """``p = y + (x - y) * min(eps/||x - y||, 1)`` — ``src/proj_l2ball.m:20-22``.

    x, y: complex64 tensors of equal shape on the GPU (norm in FP64)."""
    if x.shape != y.shape:
        raise ValueError('proj_l2ball: x and y must have the same shape')
    from .core import default_context
    ctx = default_context(x.device.index or 0)
    _require(x, torch.complex64, ctx.device, 'proj_l2ball: x')
    _require(y, torch.complex64, ctx.device, 'proj_l2ball: y')
    p = torch.empty_like(x)
    ctx.check(lib().fhs_proj_l2ball(ctx.h, _ptr(x), float(eps), _ptr(y), x.numel(), _ptr(p)),
              'proj_l2ball')
    return p


def fhs_update_primal(xsol, g, *more_g):
    """``[xsol, xhat, rel_x, norm_x] = fhs_update_primal(xsol, g)`` —
    ``src/fhs/fhs_update_primal.m:33-38``.  ``g`` may be given as up to three
    summands (the g0/g1/g2 parts of the iteration).  Updates ``xsol`` in
    place and returns (xsol, xhat, rel_x, norm_x)."""
    from .core import default_context
    ctx = default_context(xsol.device.index or 0)
    _require(xsol, torch.float32, ctx.device, 'update_primal: xsol')
    gs = [g] + list(more_g)
    if len(gs) > 3:
        raise ValueError('update_primal: at most three g summands')
    for gi in gs:
        _require(gi, torch.float32, ctx.device, 'update_primal: g')
        if gi.shape != xsol.shape:
            raise ValueError('update_primal: g must match xsol')
    gs += [None] * (3 - len(gs))
    xhat = torch.empty_like(xsol)
    rel = C.c_double()
    nrm = C.c_double()
    ctx.check(lib().fhs_update_primal(ctx.h, _ptr(xsol), _ptr(gs[0]), _ptr(gs[1]), _ptr(gs[2]),
                                      _ptr(xhat), xsol.numel(), C.byref(rel), C.byref(nrm)),
              'update_primal')
    return xsol, xhat, rel.value, nrm.value


def update_epsilon(epsilon, t, t_block, norm_res, adapt_eps_tol_in, adapt_eps_tol_out,
                   adapt_eps_steps, adapt_eps_change_percentage):
    """``[epsilon, t_block] = update_epsilon(...)`` — ``src/update_epsilon.m:1-73``.

    Cells ``{L}{nblocks}`` as nested lists (or flat arrays); returns updated
    copies in the same nesting."""
    import numpy as np
    nested = isinstance(epsilon, (list, tuple)) and len(epsilon) and isinstance(epsilon[0], (list, tuple, np.ndarray))
    shape = [len(e) for e in epsilon] if nested else None
    flat = lambda c: np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=float) for e in c]) if nested
                                          else np.asarray(c, dtype=float))  # noqa: E731
    eps = flat(epsilon)
    tb = np.ascontiguousarray(flat(t_block).astype(np.int32))
    nr = flat(norm_res)
    n_upd = C.c_int()
    st = lib().fhs_update_epsilon(eps.ctypes.data_as(C.POINTER(C.c_double)), int(t),
                                  tb.ctypes.data_as(C.POINTER(C.c_int)), nr.ctypes.data_as(C.POINTER(C.c_double)),
                                  len(eps), float(adapt_eps_tol_in), float(adapt_eps_tol_out), int(adapt_eps_steps),
                                  float(adapt_eps_change_percentage), C.byref(n_upd))
    if st != 0:
        raise ValueError('update_epsilon: bad argument')
    if not nested:
        return eps, tb
    out_e, out_t, i = [], [], 0
    for n in shape:
        out_e.append(list(eps[i:i + n])); out_t.append(list(tb[i:i + n])); i += n
    return out_e, out_t
