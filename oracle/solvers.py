"""Oracle: PDFB solvers (HyperSARA, SARA), float64.

Test infrastructure only (see ``oracle/__init__.py``).

``problem`` is a ``OracleProblem``: per-channel NUFFT operators, block row
ranges, visibilities ``y``, preconditioning weights ``pU`` and l2 bounds.
"""
import numpy as np

from .data_fidelity import update_dual_data_fidelity
from .nufft import NufftOracle
from .prox import nuclear_prox_dual, soft_l21, soft_l1_dual, update_primal
from .wavelets import SaraDictionary


class OracleProblem:
    """Data of one synthetic problem (whitened), float64.

    u, v: per-channel uv arrays; y: per-channel visibilities; pU: per-channel
    preconditioning weights; blocks: per-channel list of (start, stop) row
    ranges; epsilon: per-channel list of block bounds.
    """

    def __init__(self, Ny, Nx, u, v, y, pU, blocks, epsilon, g_scale=None, nufft_kwargs=None):
        self.Ny, self.Nx = Ny, Nx
        self.L = len(u)
        kw = nufft_kwargs or {}
        self.ops = [NufftOracle(u[l], v[l], Ny, Nx, **kw) for l in range(self.L)]
        if g_scale is not None:     # whitening G <- G / sigma (util_gen_measurements_snr_new.m:78)
            for l, op in enumerate(self.ops):
                op.G = op.G * g_scale[l]
        self.No = self.ops[0].No
        self.A = [op.A for op in self.ops]
        self.At = [op.At for op in self.ops]
        self.G = [[op.G[a:b] for (a, b) in blocks[l]] for l, op in enumerate(self.ops)]
        self.W = [[np.arange(self.No) for _ in blocks[l]] for l in range(self.L)]
        self.y = [[np.asarray(y[l][a:b]) for (a, b) in blocks[l]] for l in range(self.L)]
        self.pU = [[np.asarray(pU[l][a:b]) for (a, b) in blocks[l]] for l in range(self.L)]
        self.epsilon = [list(e) for e in epsilon]
        self.blocks = blocks
        self.dr, self.Lambda = False, None

    def holographic(self, noise):
        """The DR form (``imaging/imaging.m:619-631``): per channel one block
        holding H, W, Lambda, the gridded data and the l2 bound of the
        gridded ``noise[l]`` (``data_fidelity.holographic_operator``); no
        preconditioning.  Returns a new problem."""
        from .data_fidelity import holographic_operator
        import copy
        P = copy.copy(self)
        P.G, P.W, P.y, P.pU, P.epsilon, P.Lambda = [], [], [], [], [], []
        for l, op in enumerate(self.ops):
            y = np.concatenate(self.y[l])
            H, W, lam, yg, l2, _ = holographic_operator(op.G, y, noise[l])
            P.G.append([H])
            P.W.append([W])
            P.y.append([yg])
            P.pU.append([np.ones(len(W))])
            P.epsilon.append([l2])
            P.Lambda.append([lam])
        P.dr = True
        return P


def hs_initial_weights(x, sara, alpha, sig, sig_bar):
    """hs_initialize_dual_lowrankness_serial.m:25-29 and
    hs_initialize_dual_sparsity_distributed.m:24-35 (and the per-reweight
    hs_update_weights_*.m, identical formulas)."""
    Ny, Nx, L = x.shape
    X = x.reshape((-1, L), order='F')
    d0 = np.abs(np.linalg.svd(X, compute_uv=False))
    ub = sig_bar * alpha
    w0 = ub / (ub + d0)
    r = np.stack([sara.psit(x[:, :, l]) for l in range(L)], axis=1)
    d1 = np.sqrt(np.sum(r ** 2, axis=1))
    up = sig * alpha
    w1 = up / (up + d1)
    return w0, w1


def snr(x0, x):
    """``src/hs/hyperSARA.m`` SNR: 20 log10(||X0|| / ||X0 - X||)."""
    return 20 * np.log10(np.linalg.norm(x0.ravel()) / np.linalg.norm((x0 - x).ravel()))


def hypersara(problem, param, x_init=None, X0=None, max_iter_total=None, callback=None):
    """PDFB + reweighting of ``src/hs/hyperSARA.m`` (no checkpointing, no noise
    adjustment: ``adjust_flag_noise`` false as in
    ``imaging/default_parameters.json:70``).

    param keys: nu2 (per channel), gamma, gamma0, sig, sig_bar, alpha,
    pdfb_min_iter, pdfb_max_iter, pdfb_rel_var, pdfb_rel_var_low,
    reweighting_min_iter, reweighting_max_iter, reweighting_rel_var,
    elipse_proj_max_iter, elipse_proj_min_iter, elipse_proj_eps.
    Returns (xsol, info).
    """
    P = problem
    Ny, Nx, L = P.Ny, P.Nx, P.L
    N = Ny * Nx
    sara = SaraDictionary(Ny, Nx)
    xsol = np.zeros((Ny, Nx, L)) if x_init is None else np.array(x_init, dtype=np.float64)
    g = np.zeros_like(xsol)
    xlast = xsol.copy()
    w0, w1 = hs_initial_weights(xsol, sara, param['alpha'], param['sig'], param['sig_bar'])
    if not np.any(xsol):
        w0 = np.ones(min(N, L)); w1 = np.ones(sara.s)
    v0 = np.zeros((N, L))
    v1 = np.zeros((sara.s, L))
    v2 = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[l]] for l in range(L)]
    proj = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[l]] for l in range(L)]
    Fx_old = np.stack([P.A[l](xsol[:, :, l]) for l in range(L)], axis=1)
    tau = 0.99 / 3
    sigma00 = tau * 1.0
    sigma11 = tau * 1.0
    sigma22 = tau / np.asarray(param['nu2'], dtype=np.float64)
    beta0 = param['gamma0'] / 1.0
    beta1 = param['gamma'] / 1.0
    max_iter = (param['reweighting_max_iter'] + 1) * param['pdfb_max_iter']
    if max_iter_total is not None:
        max_iter = min(max_iter, max_iter_total)
    rw_count = 0
    last_step = 0
    info = {'rel_val': [], 'snr': [], 'norm_res': [], 'norm_eps': [], 'reweights': []}
    t = 0
    for t in range(1, max_iter + 1):
        prev = xsol
        xsol, xhat, rel_x, norm_x = update_primal(xsol, g)
        # nuclear prox (hs_update_dual_lowrankness_serial.m:5-10)
        Z = v0 + xhat.reshape((N, L), order='F')
        v0 = nuclear_prox_dual(Z, beta0 * w0)
        g0 = sigma00 * v0.reshape((Ny, Nx, L), order='F')
        # l21 prox (hs_update_dual_sparsity_distributed.m:4-17)
        r1 = v1 + np.stack([sara.psit(xhat[:, :, l]) for l in range(L)], axis=1)
        v1 = soft_l21(r1, beta1 * w1)
        g1 = np.stack([sigma11 * sara.psi(v1[:, l]) for l in range(L)], axis=2)
        # data fidelity
        v2, g2, Fx_old, proj, norm_res, gnr, neps = update_dual_data_fidelity(
            v2, P.y, xsol, Fx_old, proj, P.A, P.At, P.G, P.W, P.pU, P.epsilon,
            param['elipse_proj_max_iter'], param['elipse_proj_min_iter'],
            param['elipse_proj_eps'], sigma22, flag_dimensionality_reduction=P.dr, Lambda=P.Lambda)
        g = g0 + g1 + g2
        nx = np.linalg.norm(xsol.ravel())
        rel_val = np.linalg.norm((prev - xsol).ravel()) / nx if nx > 0 else np.nan
        info['rel_val'].append(rel_val)
        info['norm_res'].append(np.sqrt(gnr))
        info['norm_eps'].append(np.sqrt(neps))
        if X0 is not None:
            info['snr'].append(snr(X0, xsol))
        if callback is not None:
            callback(t, xsol, info)
        k = t - last_step
        pdfb_converged = (k >= param['pdfb_min_iter']) and (
            k >= param['pdfb_max_iter'] or rel_val <= param['pdfb_rel_var']
            or rel_val <= param['pdfb_rel_var_low'])
        if pdfb_converged:
            rel_rw = np.linalg.norm((xlast - xsol).ravel()) / np.linalg.norm(xlast.ravel()) \
                if np.any(xlast) else np.inf
            xlast = xsol.copy()
            converged = (rw_count >= param['reweighting_min_iter'] and
                         (rw_count >= param['reweighting_max_iter'] or
                          rel_rw <= param['reweighting_rel_var']))
            if converged:
                break
            w0, w1 = hs_initial_weights(xsol, sara, param['alpha'], param['sig'], param['sig_bar'])
            rw_count += 1
            last_step = t
            info['reweights'].append(t)
    info['iterations'] = t
    info['v0'], info['v1'], info['v2'], info['g'] = v0, v1, v2, g
    return xsol, info


def sara_solver(problem, param, x_init=None, X0=None, max_iter_total=None):
    """Single-channel PDFB of ``src/sara/sara.m:521-633`` (FP64, no reweighting
    when ``reweighting_max_iter`` is 0): per-basis l1 dual steps
    (``run_par_waverec``, ``src/sara/sara.m:921-932``) and the data term with
    step ``tau = 0.99/(sigma1 nu1 + sigma2 nu2)`` (``src/sara/sara.m:505``)."""
    P = problem
    Ny, Nx = P.Ny, P.Nx
    sara = SaraDictionary(Ny, Nx)
    xsol = np.zeros((Ny, Nx, 1)) if x_init is None else np.array(x_init, dtype=np.float64).reshape(Ny, Nx, 1)
    g = np.zeros_like(xsol)
    v1 = np.zeros(sara.s)
    w1 = np.ones(sara.s)
    v2 = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[0]]]
    proj = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[0]]]
    Fx_old = P.A[0](xsol[:, :, 0])[:, None]
    nu2 = float(np.atleast_1d(param['nu2'])[0])
    sigma1, sigma2 = 1.0, 1.0 / nu2
    tau = 0.99 / (sigma1 * 1.0 + sigma2 * nu2)
    sigma11, sigma22 = tau * sigma1, tau * sigma2
    beta1 = param['gamma'] / sigma1
    n_iter = param['pdfb_max_iter'] if max_iter_total is None else max_iter_total
    info = {'rel_val': [], 'snr': []}
    for t in range(1, n_iter + 1):
        prev = xsol
        xsol, xhat, _, _ = update_primal(xsol, g)
        r1 = v1 + sara.psit(xhat[:, :, 0])
        v1 = soft_l1_dual(r1, beta1 * w1)
        g1 = sara.psi(v1)
        v2, Ftx, Fx_old, proj, _, _, _ = update_dual_data_fidelity(
            v2, P.y, xsol, Fx_old, proj, P.A, P.At, P.G, P.W, P.pU, P.epsilon,
            param['elipse_proj_max_iter'], param['elipse_proj_min_iter'],
            param['elipse_proj_eps'], [1.0])
        g = sigma11 * g1[:, :, None] + sigma22 * Ftx
        nx = np.linalg.norm(xsol.ravel())
        info['rel_val'].append(np.linalg.norm((prev - xsol).ravel()) / nx if nx > 0 else np.nan)
        if X0 is not None:
            info['snr'].append(snr(X0.reshape(Ny, Nx, 1), xsol))
    return xsol[:, :, 0], info


def facet_hypersara(problem, param, geom, x_init=None, X0=None, max_iter_total=None):
    """PDFB of ``src/fhs/facetHyperSARA.m:701-889`` on one process: faceted
    weighted nuclear norm (overlapping facets ``geom`` with apodisation
    windows, per-facet beta0_q = gamma0_q / sigma0, ``:598-601``) and the
    SARA l21 term whose facet coefficient sets tile the global transform, so
    the global l21 step is the union of the facet steps
    (``src/fhs/fhs_update_dual_sparsity.m:67-93``).  Stopping rule with the
    data-fidelity check (``:877-881``).  param as ``hypersara`` with
    ``sig_bar`` per facet."""
    from .facets import faceted_lowrank_step, facet_singular_values
    P = problem
    Ny, Nx, L = P.Ny, P.Nx, P.L
    Q = len(geom['oy'])
    sara = SaraDictionary(Ny, Nx)
    xsol = np.zeros((Ny, Nx, L)) if x_init is None else np.array(x_init, dtype=np.float64)
    g = np.zeros_like(xsol)
    sig_bar = np.asarray(param['sig_bar'], dtype=np.float64).reshape(Q)
    w0 = np.ones((Q, L))
    w1 = np.ones(sara.s)
    if np.any(xsol):
        ub = sig_bar[:, None] * param['alpha']
        w0 = ub / (ub + facet_singular_values(xsol, geom))
        _, w1 = hs_initial_weights(xsol, sara, param['alpha'], param['sig'], 1.0)
    v0s = [np.zeros((geom['py'][q] * geom['px'][q], L)) for q in range(Q)]
    v1 = np.zeros((sara.s, L))
    v2 = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[l]] for l in range(L)]
    proj = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[l]] for l in range(L)]
    Fx_old = np.stack([P.A[l](xsol[:, :, l]) for l in range(L)], axis=1)
    tau = 0.99 / 3
    sigma00, sigma11 = tau, tau
    sigma22 = tau / np.asarray(param['nu2'], dtype=np.float64)
    beta0 = np.asarray(param['gamma0'], dtype=np.float64).reshape(Q)
    beta1 = param['gamma']
    max_iter = (param['reweighting_max_iter'] + 1) * param['pdfb_max_iter']
    if max_iter_total is not None:
        max_iter = min(max_iter, max_iter_total)
    info = {'rel_val': [], 'snr': [], 'norm_res': [], 'norm_eps': [], 'reweights': []}
    rw_count, last_step = 0, 0
    xlast = xsol.copy()
    t = 0
    for t in range(1, max_iter + 1):
        prev = xsol
        xsol, xhat, _, _ = update_primal(xsol, g)
        v0s, g0 = faceted_lowrank_step(v0s, xhat, geom, beta0[:, None] * w0, 1.0, sigma00)
        r1 = v1 + np.stack([sara.psit(xhat[:, :, l]) for l in range(L)], axis=1)
        v1 = soft_l21(r1, beta1 * w1)
        g1 = np.stack([sigma11 * sara.psi(v1[:, l]) for l in range(L)], axis=2)
        v2, g2, Fx_old, proj, norm_res, gnr, neps = update_dual_data_fidelity(
            v2, P.y, xsol, Fx_old, proj, P.A, P.At, P.G, P.W, P.pU, P.epsilon,
            param['elipse_proj_max_iter'], param['elipse_proj_min_iter'], param['elipse_proj_eps'], sigma22)
        g = g0 + g1 + g2
        nx = np.linalg.norm(xsol.ravel())
        rel_val = np.linalg.norm((prev - xsol).ravel()) / nx if nx > 0 else np.nan
        info['rel_val'].append(rel_val)
        info['norm_res'].append(np.sqrt(gnr))
        info['norm_eps'].append(np.sqrt(neps))
        if X0 is not None:
            info['snr'].append(snr(X0, xsol))
        k = t - last_step
        conv = (k >= param['pdfb_min_iter']) and (
            k >= param['pdfb_max_iter'] or
            (rel_val <= param['pdfb_rel_var'] and np.sqrt(gnr) <= param['pdfb_fidelity_tolerance'] * np.sqrt(neps))
            or rel_val <= param['pdfb_rel_var_low'])
        if conv:
            rel_rw = np.linalg.norm((xlast - xsol).ravel()) / np.linalg.norm(xlast.ravel()) \
                if np.any(xlast) else np.inf
            xlast = xsol.copy()
            if (rw_count >= param['reweighting_min_iter'] and
                    (rw_count >= param['reweighting_max_iter'] or rel_rw <= param['reweighting_rel_var'])):
                break
            ub = sig_bar[:, None] * param['alpha']
            w0 = ub / (ub + facet_singular_values(xsol, geom))
            _, w1 = hs_initial_weights(xsol, sara, param['alpha'], param['sig'], 1.0)
            rw_count += 1
            last_step = t
            info['reweights'].append(t)
    info['iterations'] = t
    return xsol, info
