"""SARA / HyperSARA / Faceted HyperSARA solvers over the GPU PDFB engine.

The per-iteration work runs in libfhs (``fhs_pdfb_iterate``: primal,
low-rank, sparsity and data steps, one hipGraph replay); this module keeps
the reference's outer logic on the host: parameter wiring
(``imaging/imaging.m:829-860``), step sizes (``src/fhs/facetHyperSARA.m:582-603``,
``src/sara/sara.m:501-510``), stopping rules (``src/fhs/facetHyperSARA.m:877-881``,
``src/hs/hyperSARA.m``) and reweighting
(``src/fhs/facetHyperSARA.m:952-1137``, ``src/hs/hs_update_weights_*.m``).
"""
import ctypes as C
import time
from math import gamma as gamma_fn, sqrt

import numpy as np
import torch

from ._lib import lib
from .core import GMatrix, _ptr
from .data_fidelity import DataOperator
from .lowrank import LowRankPrior
from .nufft import Nufft2D, uniform_weights, uniform_weights_gpu
from .sara import SaraDictionary
from .synth import add_noise

DEFAULT_PARAM = {
    # imaging/default_parameters.json:5-27
    'reweighting_min_iter': 1, 'reweighting_max_iter': 20, 'reweighting_rel_var': 1e-4,
    'reweighting_alpha': 1.0,
    'pdfb_min_iter': 10, 'pdfb_max_iter': 2000, 'pdfb_rel_var': 5e-5,
    'pdfb_fidelity_tolerance': 1.01, 'pdfb_rel_var_low': 5e-6,
    'elipse_proj_min_iter': 1, 'elipse_proj_max_iter': 20, 'elipse_proj_eps': 1e-8,
    # imaging/imaging.m:213-214 (reg_gam, reg_gam_bar)
    'gam': 1.0, 'gam_bar': 1.0,
}


def _cube(x):
    """(Ny, Nx, L) numpy -> (L, Nx, Ny) float32 tensor on the GPU."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x).transpose(2, 1, 0)).astype(np.float32)).cuda()


def _uncube(t):
    return t.detach().cpu().numpy().astype(np.float64).transpose(2, 1, 0)


class Measurement:
    """GPU-resident measurement operator and whitened data of a ``synth.Problem``.

    If the problem carries no data yet, noiseless visibilities are computed
    with the GPU operator and noise is added (``synth.add_noise``)."""

    def __init__(self, ctx, problem, precond=True, sort_rows=True):
        self.ctx, self.problem = ctx, problem
        P = problem
        self.L, self.Ny, self.Nx = P.L, P.Ny, P.Nx
        self.nufft = Nufft2D(P.Ny, P.Nx)
        self.plan = self.nufft.plan(ctx)
        Kty, Ktx = self.nufft.Kty, self.nufft.Ktx
        # interpolation coefficients, computed once per channel on the GPU
        # (anchor, cy, cx): min-max rows, or w-projected rows with the w-term
        if getattr(P, 'w', None):
            coef = [self.nufft.separable_wproj_gpu(ctx, P.u[l], P.v[l], P.w[l], P.fov) for l in range(P.L)]
        else:
            coef = [self.nufft.separable_gpu(ctx, P.u[l], P.v[l]) for l in range(P.L)]
        if not P.y:
            Graw = [GMatrix.from_separable(ctx, a, cy, cx, Kty, Ktx) for a, cy, cx in coef]
            Draw = DataOperator(ctx, self.plan, Graw, P.blocks)
            y0 = self._split(Draw, self._forward(Draw, _cube(P.X0)).cpu().numpy().astype(np.complex128))
            del Draw, Graw
            add_noise(P, y0)
        # Visibilities are reordered inside each block by the 32 x 32 grid
        # tile of their interpolation window's anchor, then by the anchor
        # inside the tile: the rows of a wavefront then gather one small grid
        # patch (forward) and the rows a tile's adjoint reads sit together.
        # Per-block results are unchanged.
        ty, tx = min(32, Kty), min(32, Ktx)
        self.order = []
        for l in range(P.L):
            if sort_rows:
                a = coef[l][0].astype(np.int64)
                ay, ax = a % Kty, a // Kty
                key = ((ax // tx) * (Kty // ty) + ay // ty) * (ty * tx) + (ax % tx) * ty + ay % ty
                b = np.concatenate([[0], np.cumsum(P.blocks[l])])
                idx = np.concatenate([b[i] + np.argsort(key[b[i]:b[i + 1]], kind='stable')
                                      for i in range(len(P.blocks[l]))])
            else:
                idx = np.arange(len(P.u[l]))
            self.order.append(idx)
        self.u = [P.u[l][self.order[l]] for l in range(P.L)]
        self.v = [P.v[l][self.order[l]] for l in range(P.L)]
        self.y_host = [np.asarray(P.y[l])[self.order[l]] for l in range(P.L)]
        self.G = []
        self.anchor = [coef[l][0][self.order[l]] for l in range(P.L)]
        for l, (a, cy, cx) in enumerate(coef):
            o = self.order[l]
            self.G.append(GMatrix.from_separable(ctx, a[o], cy[o], cx[o] * np.float32(1.0 / P.sigma_noise[l]),
                                                 Kty, Ktx))
        del coef
        self.D = DataOperator(ctx, self.plan, self.G, P.blocks)
        self.y = torch.from_numpy(np.concatenate(self.y_host).astype(np.complex64)).cuda()
        if precond:
            # device weights from the GPU counter; the host copy (FP64, same
            # cells) serves host-side consumers such as residual checks
            self.pU = torch.cat([uniform_weights_gpu(ctx, self.u[l], self.v[l], self.nufft.Kty, self.nufft.Ktx)
                                 for l in range(P.L)])
            self.pU_host = [uniform_weights(self.u[l], self.v[l], self.nufft.Kty, self.nufft.Ktx) for l in range(P.L)]
        else:
            self.pU_host = [np.ones(len(u)) for u in self.u]
            self.pU = None
        self.epsilon = np.concatenate([np.asarray(e, dtype=np.float64) for e in P.epsilon])

    @staticmethod
    def _split(D, y):
        return [y[D.row_offset[l]:D.row_offset[l + 1]] for l in range(D.L)]

    def _forward(self, D, x, lam=None):
        out = torch.empty(D.Mtot, dtype=torch.complex64, device=self.ctx.device)
        self.ctx.check(lib().fhs_data_forward(self.ctx.h, D.h, _ptr(x), _ptr(lam), _ptr(out)), 'data_forward')
        return out

    def forward(self, x, lam=None):
        """Stacked G_l A x_l (whitened operator)."""
        return self._forward(self.D, x, lam)

    def adjoint(self, y, lam=None, scale=None):
        out = torch.empty((self.L, self.Nx, self.Ny), dtype=torch.float32, device=self.ctx.device)
        sc = None if scale is None else (C.c_float * self.L)(*[float(s) for s in scale])
        self.ctx.check(lib().fhs_data_adjoint(self.ctx.h, self.D.h, _ptr(y), _ptr(lam), sc, _ptr(out)),
                       'data_adjoint')
        return out

    def op_norm(self, precond, tol=1e-8, max_iter=200, seed=0):
        """Per-channel ||Phi_l||^2 (or ||sqrt(pU) Phi_l||^2) by power iteration
        (``lib/utils/util_operator_norm.m:43-61``), all channels at once."""
        g = torch.Generator(device='cpu').manual_seed(seed)
        x = torch.randn((self.L, self.Nx, self.Ny), generator=g, dtype=torch.float64).float().cuda()
        x /= torch.linalg.vector_norm(x.double(), dim=(1, 2), keepdim=True).float()
        val = np.ones(self.L)
        lam = self.pU if precond else None
        for _ in range(max_iter):
            xn = self.adjoint(self.forward(x), lam=lam)
            nv = torch.linalg.vector_norm(xn.double(), dim=(1, 2)).cpu().numpy()
            rel = np.abs(nv - val) / nv
            x = (xn / torch.from_numpy(nv).float().cuda()[:, None, None])
            val = nv
            if np.all(rel < tol):
                break
        return val

    def residual_images(self, x):
        """``src/compute_residual_images.m:39-103`` (non-DR): real(At(G'(y - G A x)))
        normalised per channel by the PSF peak (Dirac at (Ny/2, Nx/2))."""
        r = self.y - self.forward(x)
        res = self.adjoint(r)
        dirac = torch.zeros_like(res)
        dirac[:, self.Nx // 2, self.Ny // 2] = 1.0
        psf = self.adjoint(self.forward(dirac))
        peak = psf.amax(dim=(1, 2), keepdim=True)
        return res / peak


def noise_levels_hs(Ny, Nx, L, std_noise, squared_operator_norm):
    """``src/heuristics/compute_noise_level.m:49-56`` ('hs')."""
    std = np.asarray(std_noise, dtype=np.float64)
    son = np.asarray(squared_operator_norm, dtype=np.float64)
    sig_sara = sqrt(np.mean(std ** 2 / son))
    mu_chi = sqrt(2) * gamma_fn((L + 1) / 2) / gamma_fn(L / 2)
    sig_chi = sqrt(L - mu_chi ** 2)
    sig = sig_sara * (mu_chi + sig_chi)
    sig_bar = sqrt(Nx * Ny * np.sum(std ** 2 / son) / min(Nx * Ny, L))
    return sig, sig_bar


def noise_levels_fhs(L, std_noise, squared_operator_norm, facet_pixels):
    """``src/heuristics/compute_noise_level.m:57-88`` ('fhs'): per-facet sig_bar."""
    std = np.asarray(std_noise, dtype=np.float64)
    son = np.asarray(squared_operator_norm, dtype=np.float64)
    sig_sara = sqrt(np.mean(std ** 2 / son))
    mu_chi = sqrt(2) * gamma_fn((L + 1) / 2) / gamma_fn(L / 2)
    sig_chi = sqrt(L - mu_chi ** 2)
    sig = sig_sara * (mu_chi + sig_chi)
    sig_bar = np.array([sqrt(P * np.sum(std ** 2 / son) / min(P, L)) for P in facet_pixels])
    return sig, sig_bar


class PdfbEngine:
    """``fhs_pdfb`` with torch-owned buffers."""

    def __init__(self, meas: Measurement, sara: SaraDictionary, lowrank: LowRankPrior = None, x_init=None):
        ctx = meas.ctx
        self.ctx, self.meas, self.sara, self.lowrank = ctx, meas, sara, lowrank
        L, Ny, Nx = meas.L, meas.Ny, meas.Nx
        dev = ctx.device
        f32 = dict(dtype=torch.float32, device=dev)
        self.x = torch.zeros((L, Nx, Ny), **f32) if x_init is None else x_init.clone()
        self.xhat = torch.zeros_like(self.x)
        self.g0 = torch.zeros_like(self.x) if lowrank else None
        self.g1 = torch.zeros_like(self.x)
        self.g2 = torch.zeros_like(self.x)
        self.v0 = lowrank.zeros_v0() if lowrank else None
        self.v1 = torch.zeros((L, sara.s), **f32)
        self.v2 = meas.D.zeros()
        self.proj = meas.D.zeros()
        self.dr = bool(getattr(meas, 'dr', False))
        self.Lambda = getattr(meas, 'Lambda', None) if self.dr else None
        self.gx_old = meas.D.init_fx_old(self.x, self.Lambda, self.dr)
        self.weights1 = torch.ones(sara.s, **f32)
        b = _Buffers()
        b.Ny, b.Nx = Ny, Nx
        for name in ('x', 'xhat', 'g0', 'g1', 'g2', 'v0', 'v1', 'v2', 'proj'):
            setattr(b, name, _vp(getattr(self, name)))
        b.Gx_old = _vp(self.gx_old)
        b.y = _vp(meas.y)
        b.pU = _vp(meas.pU)
        b.weights1 = _vp(self.weights1)
        b.Lambda = _vp(self.Lambda)
        self._b = b
        h = C.c_void_p()
        ctx.check(lib().fhs_pdfb_create(ctx.h, meas.D.h, sara.h, lowrank.h if lowrank else None,
                                        C.byref(b), C.byref(h)), 'pdfb_create')
        self.h = h
        torch.cuda.synchronize()

    def set_params(self, thresh0, beta1, sigma00, sigma11, sigma22, epsilon, fb_max, fb_min, fb_eps, dr=None):
        L = self.meas.L
        th = None
        if self.lowrank is not None:
            th_arr = np.ascontiguousarray(np.asarray(thresh0, dtype=np.float64).reshape(self.lowrank.Q, L))
            self._th = th_arr
            th = th_arr.ctypes.data_as(C.POINTER(C.c_double))
        s22 = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma22, dtype=np.float32), (L,)))
        eps = np.ascontiguousarray(np.asarray(epsilon, dtype=np.float64))
        dr = self.dr if dr is None else dr
        # kept for the drop-in form (dropin_iterate), which passes them per call
        self._prm = dict(beta1=float(beta1), sigma00=float(sigma00), sigma11=float(sigma11), sigma22=s22,
                         epsilon=eps, fb=(int(fb_max), int(fb_min), float(fb_eps)), dr=bool(dr))
        self.ctx.check(lib().fhs_pdfb_set_params(
            self.h, th, float(beta1), float(sigma00), float(sigma11), s22.ctypes.data_as(C.POINTER(C.c_float)),
            eps.ctypes.data_as(C.POINTER(C.c_double)), int(fb_max), int(fb_min), float(fb_eps), int(dr)),
            'pdfb_set_params')

    def set_graph(self, on):
        self.ctx.check(lib().fhs_pdfb_set_graph(self.h, int(bool(on))), 'pdfb_set_graph')

    def set_concurrent(self, on):
        self.ctx.check(lib().fhs_pdfb_set_concurrent(self.h, int(bool(on))), 'pdfb_set_concurrent')

    def iterate(self, n=1):
        out = np.zeros(4 * n)
        self.ctx.check(lib().fhs_pdfb_iterate(self.h, int(n), out.ctypes.data_as(C.POINTER(C.c_double))),
                       'pdfb_iterate')
        return out.reshape(n, 4)

    def run(self, rule, k0, prm, n_max, err2_stop=-1.0, seconds_max=0.0):
        """Iterations until the PDFB stopping rule (``fhs_pdfb_run``): returns
        (scalars (n, 6): four sums, ||X0 - x||^2, seconds since the call;
        converged)."""
        out = np.zeros(6 * max(1, int(n_max)))
        n, conv = C.c_int(), C.c_int()
        self.ctx.check(lib().fhs_pdfb_run(
            self.h, int(rule), int(k0), int(prm['pdfb_min_iter']), int(prm['pdfb_max_iter']),
            float(prm['pdfb_rel_var']), float(prm['pdfb_fidelity_tolerance']), float(prm['pdfb_rel_var_low']),
            int(n_max), float(err2_stop), float(seconds_max), out.ctypes.data_as(C.POINTER(C.c_double)),
            C.byref(n), C.byref(conv)), 'pdfb_run')
        return out[:6 * n.value].reshape(n.value, 6), bool(conv.value)

    def dropin_iterate(self):
        """One iteration in the drop-in form: the reference-named entry points
        called one after the other on the engine's buffers, eagerly on the
        context stream, each returning its host scalars -- the sequence a
        MATLAB solver makes through the MEX drop-ins
        (``src/hs/hyperSARA.m:574-613``: primal update, low-rank and sparsity
        dual steps, ``update_dual_data_fidelity``).  Same kernels and the
        same arithmetic as ``iterate``; no graph, no concurrent streams, no
        fused reductions.  Returns the four sums of ``iterate``."""
        ctx, P, m = self.ctx, self._prm, self.meas
        L, n = m.L, m.L * m.Ny * m.Nx
        rel, nrm = C.c_double(), C.c_double()
        ctx.check(lib().fhs_update_primal(ctx.h, _ptr(self.x), _ptr(self.g1), _ptr(self.g2), _ptr(self.g0),
                                          _ptr(self.xhat), n, C.byref(rel), C.byref(nrm)), 'update_primal')
        if self.lowrank is not None:
            th = self._th
            ctx.check(lib().fhs_update_dual_lowrankness(ctx.h, self.lowrank.h, _ptr(self.v0), _ptr(self.xhat),
                                                        th.ctypes.data_as(C.POINTER(C.c_double)), 1.0,
                                                        P['sigma00'], _ptr(self.g0)), 'update_dual_lowrankness')
        ctx.check(lib().fhs_update_dual_sparsity(ctx.h, self.sara.h, _ptr(self.v1), _ptr(self.xhat), L,
                                                 _ptr(self.weights1), P['beta1'], P['sigma11'], _ptr(self.g1)),
                  'update_dual_sparsity')
        nr = np.zeros(m.D.nblocks)
        gnr, neps = C.c_double(), C.c_double()
        fb_max, fb_min, fb_eps = P['fb']
        ctx.check(lib().fhs_update_dual_data_fidelity(
            ctx.h, m.D.h, _ptr(self.v2), _ptr(m.y), _ptr(self.x), _ptr(self.gx_old), _ptr(self.proj),
            _ptr(m.pU), P['epsilon'].ctypes.data_as(C.POINTER(C.c_double)), fb_max, fb_min, fb_eps,
            P['sigma22'].ctypes.data_as(C.POINTER(C.c_float)), int(P['dr']), _ptr(self.Lambda), _ptr(self.g2),
            nr.ctypes.data_as(C.POINTER(C.c_double)), C.byref(gnr), C.byref(neps)), 'update_dual_data_fidelity')
        return np.array([rel.value, nrm.value, gnr.value, neps.value])

    def set_truth(self, X0):
        """Reference cube (L, Nx, Ny) float32 on the GPU, or None: the primal
        update then also sums ||X0 - x||^2 (see ``truth_error``)."""
        if X0 is not None:
            X0 = X0.detach().to(device=self.ctx.device, dtype=torch.float32).reshape(
                self.meas.L, self.meas.Nx, self.meas.Ny).contiguous()
        self._truth = X0          # the engine keeps the pointer
        self.ctx.check(lib().fhs_pdfb_set_truth(self.h, None if X0 is None else _ptr(X0)), 'pdfb_set_truth')

    def truth_error(self):
        """||X0 - x||^2 after the last iteration (FP64, no GPU call)."""
        out = C.c_double()
        self.ctx.check(lib().fhs_pdfb_truth_error(self.h, C.byref(out)), 'pdfb_truth_error')
        return out.value

    def norm_res(self):
        out = np.zeros(self.meas.D.nblocks)
        self.ctx.check(lib().fhs_pdfb_norm_res(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), 'pdfb_norm_res')
        return out

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None:
            try:
                lib().fhs_pdfb_destroy(h)
            except Exception:
                pass
            self.h = None


class _Buffers(C.Structure):
    _fields_ = [('Ny', C.c_int), ('Nx', C.c_int)] + [
        (n, C.c_void_p) for n in ('x', 'xhat', 'g0', 'g1', 'g2', 'v0', 'v1', 'v2', 'proj', 'Gx_old', 'y', 'pU',
                                  'weights1', 'Lambda')]


def _vp(t):
    return None if t is None else t.data_ptr()


def snr_db(X0, x):
    """20 log10(||X0|| / ||X0 - X||) (``src/fhs/facetHyperSARA.m:864``)."""
    X0 = np.asarray(X0, dtype=np.float64)
    return 20 * np.log10(np.linalg.norm(X0) / np.linalg.norm(X0 - x))


def _solve(kind, meas, param, x_init=None, X0=None, facets=None, callback=None, graph=True,
           max_total_iter=None, time_limit=None, concurrent=True, weights_from_init=True, stop_snr=None,
           form='engine', gram='fp64', rows=32):
    """Shared PDFB + reweighting driver.  kind: 'sara' | 'hs' | 'fhs'.

    With a nonzero ``x_init`` the reference derives the first weights from it
    (``weights_from_init=True``, ``src/fhs/fhs_initialize_dual_and_weights.m``);
    ``weights_from_init=False`` keeps the zero start's unit weights, so only
    the starting point of the same sequence of problems changes.

    Without a ``callback`` the iterations of a reweighting cycle run inside
    the library (``fhs_pdfb_run``: the same stopping rule, no host round trip
    per iteration); ``stop_snr`` (with X0) ends the solve at the first
    iteration whose SNR reaches it.  ``form='dropin'`` runs every iteration
    as the sequence of reference-named entry points instead
    (``PdfbEngine.dropin_iterate``; SNR by a separate reduction).
    ``gram`` selects the nuclear-norm Gram precision
    (``LowRankPrior.set_gram_precision``: 'fp64', 'bf16x3', 'bf16x2');
    ``rows`` the storage precision of the G rows the data kernels read
    (``DataOperator.set_row_precision``: 32 or 16).  It stays set on the
    measurement after the call (the returned engine keeps using it); every
    call sets it again."""
    prm = dict(DEFAULT_PARAM)
    prm.update(param or {})
    ctx = meas.ctx
    L, Ny, Nx = meas.L, meas.Ny, meas.Nx
    if getattr(meas.D, 'row_bits', 32) != int(rows):
        meas.D.set_row_precision(int(rows))
    sara = SaraDictionary(ctx, Ny, Nx)
    lowrank = None
    if kind == 'hs':
        lowrank = LowRankPrior(ctx, Ny, Nx, L)
    elif kind == 'fhs':
        lowrank = LowRankPrior(ctx, Ny, Nx, L, facets=facets)
    if lowrank is not None and gram != 'fp64':
        lowrank.set_gram_precision(gram)
    if x_init is None:
        xi = None
    elif isinstance(x_init, torch.Tensor):          # (L, Nx, Ny) cube already on the GPU
        xi = x_init.detach().to(device=ctx.device, dtype=torch.float32).reshape(L, Nx, Ny).contiguous()
    else:
        xi = _cube(x_init)
    eng = PdfbEngine(meas, sara, lowrank, xi)
    eng.set_graph(graph)
    eng.set_concurrent(concurrent)
    # operator norms and regularisation (imaging/imaging.m:728,825-860)
    son = prm.get('squared_operator_norm')
    if son is None:
        son = meas.op_norm(False)
    nu2 = prm.get('nu2')
    if nu2 is None:
        nu2 = meas.op_norm(meas.pU is not None)
    son = np.asarray(son, dtype=np.float64)
    nu2 = np.asarray(nu2, dtype=np.float64)
    # whitened data: unit noise std (the DR form measures it on its gridded noise)
    std = getattr(meas, 'noise_std', None)
    std = np.ones(L) if std is None else np.asarray(std, dtype=np.float64)
    alpha = prm['reweighting_alpha']
    if kind == 'sara':
        sig = sqrt(std[0] ** 2 / son[0])
        sig_bar = None
        sigma1, sigma2 = 1.0, 1.0 / nu2
        tau = 0.99 / (sigma1 * 1.0 + sigma2[0] * nu2[0])            # src/sara/sara.m:505
    else:
        if kind == 'hs':
            sig, sb = noise_levels_hs(Ny, Nx, L, std, son)
            sig_bar = np.array([sb])
        else:
            pix = [facets['py'][q] * facets['px'][q] for q in range(lowrank.Q)]
            sig, sig_bar = noise_levels_fhs(L, std, son, pix)
        sigma1, sigma2 = 1.0, 1.0 / nu2
        tau = 0.99 / 3                                               # src/fhs/facetHyperSARA.m:587
    sigma00, sigma11 = tau * 1.0, tau * sigma1
    sigma22 = tau * sigma2
    gamma = prm['gam'] * sig
    beta1 = gamma / sigma1
    beta0 = None if sig_bar is None else prm['gam_bar'] * sig_bar / 1.0      # gamma0 / sigma0, per facet
    # initial weights (hs_initialize_*, fhs_initialize_dual_and_weights.m:88-112)
    Q = lowrank.Q if lowrank else 0
    w0 = np.ones((Q, L))
    if weights_from_init and xi is not None and bool(torch.any(xi != 0)):
        _reweight(eng, alpha, sig, sig_bar)
        w0 = eng._w0
    thresh0 = None if lowrank is None else (beta0[:, None] * w0)
    eng.set_params(thresh0, beta1, sigma00, sigma11, sigma22, meas.epsilon, prm['elipse_proj_max_iter'],
                   prm['elipse_proj_min_iter'], prm['elipse_proj_eps'])
    max_iter = (prm['reweighting_max_iter'] + 1) * prm['pdfb_max_iter']
    if max_total_iter is not None:
        max_iter = min(max_iter, max_total_iter)
    x_last = eng.x.clone()
    rw_count, last_step = 0, 0
    info = {'rel_val': [], 'norm_res': [], 'norm_eps': [], 'snr': [], 'time': [], 'reweights': [],
            'tau': tau, 'sigma22': sigma22, 'sig': sig, 'sig_bar': sig_bar, 'nu2': nu2, 'son': son,
            'converged': False}
    track = X0 is not None and (callback is not None or prm.get('track_snr', False))
    if track:
        # the primal kernel sums ||X0 - x||^2 with its own reductions
        X0d = X0 if isinstance(X0, torch.Tensor) else _cube(X0)
        eng.set_truth(X0d)
        nX0 = float(torch.linalg.vector_norm(eng._truth.double()))
    t0 = time.perf_counter()
    t = 0

    def end_of_cycle(t):
        """Reweighting at the end of a PDFB cycle; True when the solve is done."""
        nonlocal x_last, rw_count, last_step
        nl = torch.linalg.vector_norm(x_last.double()).item()
        rel_rw = (torch.linalg.vector_norm((x_last - eng.x).double()).item() / nl) if nl > 0 else float('inf')
        x_last = eng.x.clone()
        if (rw_count >= prm['reweighting_min_iter'] and
                (rw_count >= prm['reweighting_max_iter'] or rel_rw <= prm['reweighting_rel_var'])):
            info['converged'] = True
            return True
        _reweight(eng, alpha, sig, sig_bar)
        if lowrank is not None:
            eng.set_params(beta0[:, None] * eng._w0, beta1, sigma00, sigma11, sigma22, meas.epsilon,
                           prm['elipse_proj_max_iter'], prm['elipse_proj_min_iter'], prm['elipse_proj_eps'])
        rw_count += 1
        last_step = t
        info['reweights'].append(t)
        return False

    # the PDFB stopping rule: only Faceted HyperSARA gates the relative
    # variation on the data fidelity (src/fhs/facetHyperSARA.m:877-881);
    # SARA (src/sara/sara.m:666-670) and HyperSARA (src/hs/hyperSARA.m:682-686) do not
    rule = 1 if kind == 'fhs' else 0
    if callback is None and form == 'dropin':
        t = 0
        while t < max_iter:
            sc = eng.dropin_iterate()
            t += 1
            tc = time.perf_counter() - t0
            rel_val = sqrt(sc[0]) / sqrt(sc[1]) if sc[1] > 0 else float('nan')
            nres, neps = sqrt(sc[2]), sqrt(sc[3])
            info['rel_val'].append(rel_val)
            info['norm_res'].append(nres)
            info['norm_eps'].append(neps)
            if track:
                e2 = torch.linalg.vector_norm((eng._truth - eng.x).double()).item() ** 2
                info['snr'].append(20 * np.log10(nX0 / sqrt(e2)) if e2 > 0 else float('inf'))
            info['time'].append(time.perf_counter() - t0)
            if not np.isfinite(nres):
                raise RuntimeError('dropin: non-finite data residual')
            if pdfb_converged(rule, t - last_step, rel_val, nres, neps, prm):
                if end_of_cycle(t):
                    break
                continue
            if track and stop_snr is not None and info['snr'][-1] >= stop_snr:
                break
            if time_limit is not None and tc > time_limit:
                break
        info['iterations'] = t
        info['wall'] = time.perf_counter() - t0
        return eng, info
    if callback is None:
        # whole cycles inside the library
        err2_stop = -1.0
        if track and stop_snr is not None:
            err2_stop = (nX0 / 10 ** (stop_snr / 20)) ** 2
        while t < max_iter:
            tc = time.perf_counter() - t0
            left = 0.0 if time_limit is None else max(time_limit - tc, 1e-9)
            out, conv = eng.run(rule, t - last_step, prm, max_iter - t, err2_stop, left)
            for sc in out:
                info['rel_val'].append(sqrt(sc[0]) / sqrt(sc[1]) if sc[1] > 0 else float('nan'))
                info['norm_res'].append(sqrt(sc[2]))
                info['norm_eps'].append(sqrt(sc[3]))
                info['time'].append(tc + sc[5])
                if track:
                    info['snr'].append(20 * np.log10(nX0 / sqrt(sc[4])) if sc[4] > 0 else float('inf'))
            t += len(out)
            if conv:
                if end_of_cycle(t):
                    break
                continue
            break   # n_max, the SNR target or the time limit
        info['iterations'] = t
        info['wall'] = time.perf_counter() - t0
        return eng, info
    for t in range(1, max_iter + 1):
        s = eng.iterate(1)[0]
        rel_val = sqrt(s[0]) / sqrt(s[1]) if s[1] > 0 else float('nan')
        nres, neps = sqrt(s[2]), sqrt(s[3])
        info['rel_val'].append(rel_val)
        info['norm_res'].append(nres)
        info['norm_eps'].append(neps)
        info['time'].append(time.perf_counter() - t0)
        if track:
            e2 = eng.truth_error()
            info['snr'].append(20 * np.log10(nX0 / sqrt(e2)) if e2 > 0 else float('inf'))
        if callback is not None and callback(t, eng, info) is False:
            break
        if pdfb_converged(rule, t - last_step, rel_val, nres, neps, prm):
            if end_of_cycle(t):
                break
            continue
        if track and stop_snr is not None and info['snr'][-1] >= stop_snr:
            break
        if time_limit is not None and info['time'][-1] > time_limit:
            break
    info['iterations'] = t
    info['wall'] = time.perf_counter() - t0
    return eng, info


def pdfb_converged(rule, k, rel_val, norm_res, norm_eps, prm):
    """The library's PDFB stopping rule (``fhs_pdfb_stopping_rule``; rule 0:
    ``src/sara/sara.m:666-670``, ``src/hs/hyperSARA.m:682-686``; rule 1:
    ``src/fhs/facetHyperSARA.m:877-881``) -- the same code ``fhs_pdfb_run``
    evaluates per iteration."""
    return bool(lib().fhs_pdfb_stopping_rule(
        int(rule), int(k), float(rel_val), float(norm_res), float(norm_eps), int(prm['pdfb_min_iter']),
        int(prm['pdfb_max_iter']), float(prm['pdfb_rel_var']), float(prm['pdfb_fidelity_tolerance']),
        float(prm['pdfb_rel_var_low'])))


def _reweight(eng, alpha, sig, sig_bar):
    """Weights from the current x: w1 = sig a/(sig a + ||Psit x||_row),
    w0 = sig_bar a/(sig_bar a + s(x)) (``src/hs/hs_update_weights_*.m``,
    ``src/fhs/fhs_update_weights.m:73-92``)."""
    ctx, S = eng.ctx, eng.sara
    ctx.check(lib().fhs_sara_update_weights(ctx.h, S.h, _ptr(eng.x), eng.meas.L, float(sig * alpha),
                                            _ptr(eng.weights1)), 'sara_update_weights')
    if eng.lowrank is not None:
        sv = eng.lowrank.singular_values(eng.x)
        ub = np.asarray(sig_bar)[:, None] * alpha
        eng._w0 = ub / (ub + np.abs(sv))
    torch.cuda.synchronize()


def hyper_sara(meas, param=None, x_init=None, X0=None, **kw):
    """HyperSARA (``src/hs/hyperSARA.m``): returns (xsol (Ny, Nx, L), info)."""
    eng, info = _solve('hs', meas, param, x_init, X0, **kw)
    return _uncube(eng.x), info


def facet_hyper_sara(meas, facets, param=None, x_init=None, X0=None, **kw):
    """Faceted HyperSARA (``src/fhs/facetHyperSARA.m``) with overlapping
    nuclear-norm facets ``facets`` (geometry dict, see ``LowRankPrior``)."""
    eng, info = _solve('fhs', meas, param, x_init, X0, facets=facets, **kw)
    return _uncube(eng.x), info


def sara(meas, param=None, x_init=None, X0=None, **kw):
    """SARA (``src/sara/sara.m``), single channel: returns (xsol (Ny, Nx), info)."""
    if meas.L != 1:
        raise ValueError('sara: single-channel problems only')
    eng, info = _solve('sara', meas, param, x_init, X0, **kw)
    return _uncube(eng.x)[:, :, 0], info
