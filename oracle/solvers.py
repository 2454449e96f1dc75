"""Oracle: PDFB solvers (SARA, HyperSARA, Faceted HyperSARA), float64.

Test infrastructure only (see ``oracle/__init__.py``).

``problem`` is an ``OracleProblem``: per-channel NUFFT operators, block row
ranges, visibilities ``y``, preconditioning weights ``pU`` and l2 bounds.

One loop, ``pdfb``, restates the three reference solvers; where they differ
(step sizes, prior steps, the stopping rule) the branch cites the file:line
it follows.  Left out, as in the reference's default run: noise-level
adjustment (``adjust_flag_noise`` false, ``imaging/default_parameters.json:70``),
checkpointing and logging.

How the loop is pinned (``tests/test_oracle_solvers.py``): the stopping and
reweighting rules by hand-built sequences that drive every branch of every
rule, checked against the ``.m`` text; the whole loop by the optimality
conditions of the problem it solves, checked at a converged iterate (primal
stationarity, the dual variables in the subdifferentials of the priors and
the normal cone of the l2 ball).  The MATLAB iterates themselves cannot be
produced here (no MATLAB): parity unpinned at that level.
"""
import numpy as np

from .data_fidelity import update_dual_data_fidelity
from .facets import faceted_lowrank_step, facet_singular_values
from .nufft import NufftOracle
from .prox import nuclear_prox_dual, soft_l21, soft_l1_dual, update_primal
from .wavelets import SaraDictionary, DEFAULT_BASES, DEFAULT_NLEVEL


class OracleProblem:
    """Data of one synthetic problem (whitened), float64.

    u, v: per-channel uv arrays; y: per-channel visibilities; pU: per-channel
    preconditioning weights; blocks: per-channel list of (start, stop) row
    ranges; epsilon: per-channel list of block bounds.
    """

    def __init__(self, Ny, Nx, u, v, y, pU, blocks, epsilon, g_scale=None, nufft_kwargs=None, threads=1):
        self.Ny, self.Nx = Ny, Nx
        self.L = len(u)
        kw = nufft_kwargs or {}

        def build(l):
            return NufftOracle(u[l], v[l], Ny, Nx, **kw)
        if threads > 1:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(threads) as ex:
                self.ops = list(ex.map(build, range(self.L)))
        else:
            self.ops = [build(l) for l in range(self.L)]
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

    def holographic(self, noise, datasets=None):
        """The DR form (``imaging/imaging.m:619-631``): per channel one block
        holding H, W, Lambda, the gridded data and the l2 bound of the
        gridded ``noise[l]`` (``data_fidelity.holographic_operator_sets``);
        no preconditioning.  ``datasets[l]``: the row ranges (a, b) of the
        channel's data sets, summed as the reference sums its data sets
        (default: all rows, one set).  Returns a new problem."""
        from .data_fidelity import holographic_operator_sets
        import copy
        P = copy.copy(self)
        P.G, P.W, P.y, P.pU, P.epsilon, P.Lambda = [], [], [], [], [], []
        for l, op in enumerate(self.ops):
            y = np.concatenate(self.y[l])
            sets = [(0, op.G.shape[0])] if datasets is None else datasets[l]
            H, W, lam, yg, l2, _ = holographic_operator_sets(
                [op.G[a:b] for a, b in sets], [y[a:b] for a, b in sets], [noise[l][a:b] for a, b in sets])
            P.G.append([H])
            P.W.append([W])
            P.y.append([yg])
            P.pU.append([np.ones(len(W))])
            P.epsilon.append([l2])
            P.Lambda.append([lam])
        P.dr = True
        return P

    def forward(self, x):
        """Per channel the stacked blocks of G_l A x_l (whitened measurement
        operator, ``lib/operators/HS_forward_operator_G.m``)."""
        return [np.concatenate([G @ self.A[l](x[:, :, l]) for G in self.G[l]]) for l in range(self.L)]

    def adjoint_blocks(self, v, l):
        """real(At(sum_j G_j' v_j)) of channel l (``HS_adjoint_operator_G.m``)."""
        g = np.zeros(self.No, dtype=np.complex128)
        for G, vj in zip(self.G[l], v):
            g += G.conj().T @ vj
        return np.real(self.At[l](g))


def snr(x0, x):
    """``src/hs/hyperSARA.m`` SNR: 20 log10(||X0|| / ||X0 - X||)."""
    return 20 * np.log10(np.linalg.norm(x0.ravel()) / np.linalg.norm((x0 - x).ravel()))


# --------------------------------------------------------------- rules --

def pdfb_converged(kind, k, rel_val, norm_res, norm_eps, param):
    """Inner (PDFB) stopping rule at iteration k counted from the last
    reweighting (``t - redefine_min_task_last_step_iter``).

    'sara' (``src/sara/sara.m:666-670``) and 'hs' (``src/hs/hyperSARA.m:682-686``):
        k >= min_iter && (k >= max_iter || rel_val <= rel_var || rel_val <= rel_var_low)
    'fhs' (``src/fhs/facetHyperSARA.m:877-881``): the rel_var branch also needs
        the data fidelity, norm_res <= fidelity_tolerance * norm_eps.
    NaN compares false, as in MATLAB."""
    if not k >= param['pdfb_min_iter']:
        return False
    if k >= param['pdfb_max_iter']:
        return True
    small = rel_val <= param['pdfb_rel_var']
    if kind == 'fhs':
        small = small and norm_res <= param['pdfb_fidelity_tolerance'] * norm_eps
    return bool(small or rel_val <= param['pdfb_rel_var_low'])


def reweighting_converged(rw_count, rel_x_reweighting, param, reweighting_flag=True):
    """Outer (reweighting) stopping rule, evaluated when the PDFB converged:
    ``src/sara/sara.m:713-717``, ``src/hs/hyperSARA.m:748-752``,
    ``src/fhs/facetHyperSARA.m:965-969`` (identical)."""
    return (not reweighting_flag) or (
        rw_count >= param['reweighting_min_iter'] and
        (rw_count >= param['reweighting_max_iter'] or rel_x_reweighting <= param['reweighting_rel_var']))


def rel_x_reweighting(xlast, xsol):
    """||xlast - xsol|| / ||xlast|| (``src/hs/hyperSARA.m:744``; the facet sums of
    ``src/fhs/facetHyperSARA.m:950-963`` are the same quantity).  MATLAB's
    x/0 is Inf (or NaN for 0/0), which never passes ``<= rel_var``."""
    d = np.linalg.norm((xlast - xsol).ravel())
    n = np.linalg.norm(xlast.ravel())
    if n > 0:
        return d / n
    return np.inf if d > 0 else np.nan


# ------------------------------------------------------------- weights --

def sparsity_weights(x, sara, upsilon, psit_all=None):
    """l21 weights upsilon ./ (upsilon + ||Psit(x)||_row):
    ``src/hs/hs_update_weights_sparsity_distributed.m``,
    ``src/fhs/fhs_update_weights.m:82-92``; for L = 1 the SARA weights
    ``upsilon ./ (upsilon + abs(Psit(x)))`` (``src/sara/sara.m:764-769``)."""
    L = x.shape[2]
    if psit_all is not None:
        r = psit_all(x)
    else:
        r = np.stack([sara.psit(x[:, :, l]) for l in range(L)], axis=1)
    d = np.sqrt(np.sum(r ** 2, axis=1))
    return upsilon / (upsilon + d)


def lowrank_weights(x, upsilon_bar):
    """``src/hs/hs_update_weights_lowrankness_serial.m``:
    upsilon_bar ./ (upsilon_bar + svd(reshape(x, [], L)))."""
    L = x.shape[2]
    d = np.abs(np.linalg.svd(x.reshape((-1, L), order='F'), compute_uv=False))
    return upsilon_bar / (upsilon_bar + d)


def hs_initial_weights(x, sara, alpha, sig, sig_bar):
    """``hs_initialize_dual_lowrankness_serial.m:25-29`` and
    ``hs_initialize_dual_sparsity_distributed.m:24-35`` (the same formulas
    as the per-reweighting ``hs_update_weights_*.m``): (w0, w1)."""
    return lowrank_weights(x, sig_bar * alpha), sparsity_weights(x, sara, sig * alpha)


# ---------------------------------------------------------------- loop --

def pdfb(kind, problem, param, geom=None, x_init=None, X0=None, max_iter_total=None, callback=None,
         threads=1):
    """PDFB + reweighting of one of the reference solvers, kind 'sara'
    (``src/sara/sara.m:521-840``), 'hs' (``src/hs/hyperSARA.m:565-790``) or
    'fhs' (``src/fhs/facetHyperSARA.m:701-1140``, one process; ``geom`` the
    overlapping nuclear-norm facets).

    param keys: nu2 (per channel), gamma, sig, alpha (reweighting_alpha),
    pdfb_* and reweighting_* of ``imaging/default_parameters.json:5-21``,
    elipse_proj_*; for 'hs'/'fhs' also gamma0 and sig_bar (per facet for
    'fhs'); optional 'bases'/'nlevel' (default db1-db8 + Dirac, 4 levels).
    ``threads`` > 1 visits the independent channels (wavelet transforms, the
    data step) from a thread pool; the arithmetic is unchanged.
    Returns (xsol, info) with the per-iteration rel_val, norm_res, norm_eps,
    snr, the reweighting iterations and the final duals and weights."""
    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(threads) as ex:
            return _pdfb(kind, problem, param, geom, x_init, X0, max_iter_total, callback, ex.map)
    return _pdfb(kind, problem, param, geom, x_init, X0, max_iter_total, callback, map)


def _pdfb(kind, problem, param, geom, x_init, X0, max_iter_total, callback, mapper):
    P = problem
    Ny, Nx, L = P.Ny, P.Nx, P.L
    N = Ny * Nx
    if kind == 'sara' and L != 1:
        raise ValueError('sara: one channel')
    sara = SaraDictionary(Ny, Nx, param.get('bases', DEFAULT_BASES), param.get('nlevel', DEFAULT_NLEVEL))
    xsol = np.zeros((Ny, Nx, L)) if x_init is None else np.array(x_init, dtype=np.float64).reshape(Ny, Nx, L)
    g = np.zeros_like(xsol)
    alpha = param.get('alpha', param.get('reweighting_alpha', 1.0))
    nu2 = np.broadcast_to(np.asarray(param['nu2'], dtype=np.float64), (L,))
    # step sizes and thresholds
    if kind == 'sara':
        # src/sara/sara.m:500-510: sigma1 = 1/nu1, sigma2 = 1/nu2, tau = 0.99/(sigma1 nu1 + sigma2 nu2)
        sigma1, sigma2 = 1.0, 1.0 / nu2
        tau = 0.99 / (sigma1 * 1.0 + sigma2[0] * nu2[0])
        sigma00 = None
    else:
        # src/hs/hyperSARA.m:471-486, src/fhs/facetHyperSARA.m:582-603 (nu0 = nu1 = 1)
        sigma0, sigma1, sigma2 = 1.0, 1.0, 1.0 / nu2
        tau = 0.99 / 3
        sigma00 = tau * sigma0
    sigma11 = tau * sigma1
    sigma22 = tau * sigma2
    beta1 = param['gamma'] / sigma1
    sig = param['sig']
    Q = 0
    if kind == 'hs':
        beta0 = np.atleast_1d(np.asarray(param['gamma0'], dtype=np.float64))[:1] / sigma0
        sig_bar = np.atleast_1d(np.asarray(param['sig_bar'], dtype=np.float64))[:1]
    elif kind == 'fhs':
        Q = len(geom['oy'])
        beta0 = np.asarray(param['gamma0'], dtype=np.float64).reshape(Q) / sigma0
        sig_bar = np.asarray(param['sig_bar'], dtype=np.float64).reshape(Q)

    def psit_all(x):
        return np.stack(list(mapper(lambda l: sara.psit(x[:, :, l]), range(L))), axis=1)

    def psi_all(v, scale):
        return np.stack(list(mapper(lambda l: scale * sara.psi(v[:, l]), range(L))), axis=2)

    def weights(x):
        w1 = sparsity_weights(x, sara, sig * alpha, psit_all)
        if kind == 'hs':
            w0 = lowrank_weights(x, sig_bar[0] * alpha)[None, :]
        elif kind == 'fhs':
            w0 = (sig_bar[:, None] * alpha) / (sig_bar[:, None] * alpha + facet_singular_values(x, geom))
        else:
            w0 = None
        return w0, w1

    # initial weights from x (all ones at x = 0): src/sara/sara.m:355-363,
    # hs_initialize_*, src/fhs/fhs_initialize_dual_and_weights.m:88-112
    w0, w1 = weights(xsol)
    if kind == 'hs':
        v0 = [np.zeros((N, L))]
    elif kind == 'fhs':
        v0 = [np.zeros((geom['py'][q] * geom['px'][q], L)) for q in range(Q)]
    else:
        v0 = None
    v1 = np.zeros((sara.s, L))
    v2 = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[l]] for l in range(L)]
    proj = [[np.zeros(len(b), dtype=np.complex128) for b in P.y[l]] for l in range(L)]
    Fx_old = np.stack([P.A[l](xsol[:, :, l]) for l in range(L)], axis=1)
    max_iter = (param['reweighting_max_iter'] + 1) * param['pdfb_max_iter']   # sara.m:449
    if max_iter_total is not None:
        max_iter = min(max_iter, max_iter_total)
    rw_count, last_step = 0, 0
    xlast = xsol.copy()
    info = {'rel_val': [], 'snr': [], 'norm_res': [], 'norm_eps': [], 'reweights': [], 'converged': False,
            'tau': tau, 'sigma00': sigma00, 'sigma11': sigma11, 'sigma22': sigma22, 'beta1': beta1}
    g0 = g1 = g2 = None
    t = 0
    for t in range(1, max_iter + 1):
        prev = xsol
        # src/fhs/fhs_update_primal.m:33-38 (sara.m:529-531, hyperSARA.m:574-576)
        xsol, xhat, rel_x, norm_x = update_primal(xsol, g)
        # low-rank prior
        if kind == 'hs':
            # src/hs/hs_update_dual_lowrankness_serial.m:5-10
            Z = v0[0] + xhat.reshape((N, L), order='F')
            v0 = [nuclear_prox_dual(Z, beta0[0] * w0[0])]
            g0 = sigma00 * v0[0].reshape((Ny, Nx, L), order='F')
        elif kind == 'fhs':
            # src/fhs/fhs_update_dual_lowrankness.m:40-45, overlap-add :746-754
            v0, g0 = faceted_lowrank_step(v0, xhat, geom, beta0[:, None] * w0, 1.0, sigma00)
        # sparsity prior
        r1 = v1 + psit_all(xhat)
        if kind == 'sara':
            # src/sara/sara.m:921-926 (per-basis l1 dual step)
            v1 = soft_l1_dual(r1[:, 0], beta1 * w1)[:, None]
        else:
            # src/hs/hs_update_dual_sparsity_distributed.m:4-17,
            # src/fhs/fhs_update_dual_sparsity.m:83-93
            v1 = soft_l21(r1, beta1 * w1)
        g1 = psi_all(v1, sigma11)
        # data fidelity: src/update_dual_data_fidelity.m (Ftx already times sigma22)
        v2, g2, Fx_old, proj, norm_res, gnr, neps = update_dual_data_fidelity(
            v2, P.y, xsol, Fx_old, proj, P.A, P.At, P.G, P.W, P.pU, P.epsilon,
            param['elipse_proj_max_iter'], param['elipse_proj_min_iter'], param['elipse_proj_eps'],
            sigma22, flag_dimensionality_reduction=P.dr, Lambda=P.Lambda, mapper=mapper)
        # src/sara/sara.m:615 g = sigma11 g1 + sigma22 Ftx; hyperSARA.m:585-613 g = g0 + g1 + g2
        g = g1 + g2 if g0 is None else g0 + g1 + g2
        # relative variation: sara.m:535, hyperSARA.m:620-622, facetHyperSARA.m:793
        rel_val = np.sqrt(rel_x) / np.sqrt(norm_x) if norm_x > 0 else np.nan
        nres, neps_ = np.sqrt(gnr), np.sqrt(neps)
        info['rel_val'].append(rel_val)
        info['norm_res'].append(nres)
        info['norm_eps'].append(neps_)
        if X0 is not None:
            info['snr'].append(snr(X0.reshape(xsol.shape), xsol))
        if callback is not None:
            callback(t, xsol, info)
        if pdfb_converged(kind, t - last_step, rel_val, nres, neps_, param):
            rel_rw = rel_x_reweighting(xlast, xsol)
            xlast = xsol.copy()
            if reweighting_converged(rw_count, rel_rw, param, param.get('reweighting_flag', True)):
                info['converged'] = True
                break
            # reweighting: sara.m:764-769, hyperSARA.m (hs_update_weights_*),
            # facetHyperSARA.m:990-1000 (fhs_update_weights)
            w0, w1 = weights(xsol)
            rw_count += 1
            last_step = t
            info['reweights'].append(t)
    info['iterations'] = t
    info['v0'], info['v1'], info['v2'], info['proj'] = v0, v1, v2, proj
    info['g'], info['g0'], info['g1'], info['g2'] = g, g0, g1, g2
    info['w0'], info['w1'] = w0, w1
    info['beta0'] = None if kind == 'sara' else beta0
    return xsol, info


def hypersara(problem, param, x_init=None, X0=None, max_iter_total=None, callback=None):
    """HyperSARA (``src/hs/hyperSARA.m``): ``pdfb('hs', ...)``."""
    return pdfb('hs', problem, param, x_init=x_init, X0=X0, max_iter_total=max_iter_total, callback=callback)


def sara_solver(problem, param, x_init=None, X0=None, max_iter_total=None, callback=None):
    """SARA (``src/sara/sara.m``): ``pdfb('sara', ...)``; returns (x (Ny, Nx), info).
    param needs 'sig' (the reweighting noise level) besides 'gamma'."""
    p = dict(param)
    p.setdefault('sig', p['gamma'])
    x, info = pdfb('sara', problem, p, x_init=x_init, X0=X0, max_iter_total=max_iter_total, callback=callback)
    return x[:, :, 0], info


def facet_hypersara(problem, param, geom, x_init=None, X0=None, max_iter_total=None, callback=None):
    """Faceted HyperSARA (``src/fhs/facetHyperSARA.m``) on one process:
    ``pdfb('fhs', ...)`` with the overlapping facets ``geom``.  The SARA l21
    term's facet coefficient sets tile the global transform, so the global
    l21 step is the union of the facet steps
    (``src/fhs/fhs_update_dual_sparsity.m:67-93``)."""
    return pdfb('fhs', problem, param, geom=geom, x_init=x_init, X0=X0, max_iter_total=max_iter_total,
                callback=callback)
