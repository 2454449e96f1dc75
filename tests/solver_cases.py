"""Seeded solver problems shared by the oracle tests, the golden-fixture
script (tests/golden/make_solver_fixtures.py) and the GPU solver tests.

A case is built the same way on every machine: the sky, uv coverage and
noise come from ``fhs.synth`` (numpy, seeded) and the noiseless
visibilities are computed in FP64 on the host (``synth.add_host_data``), so
the GPU run and the oracle run see identical data.
"""
import numpy as np

from fhs import synth
from fhs.nufft import uniform_weights
from oracle.facets import facet_geometry
from oracle.heuristics import compute_noise_level_hs, compute_noise_level_sara, op_norm
from oracle.solvers import OracleProblem

# imaging/default_parameters.json:5-27
DEFAULT_PARAM = {
    'reweighting_min_iter': 1, 'reweighting_max_iter': 20, 'reweighting_rel_var': 1e-4,
    'reweighting_alpha': 1.0, 'alpha': 1.0,
    'pdfb_min_iter': 10, 'pdfb_max_iter': 2000, 'pdfb_rel_var': 5e-5,
    'pdfb_fidelity_tolerance': 1.01, 'pdfb_rel_var_low': 5e-6,
    'elipse_proj_min_iter': 1, 'elipse_proj_max_iter': 20, 'elipse_proj_eps': 1e-8,
}


def block_ranges(prob):
    out = []
    for l in range(prob.L):
        b = np.concatenate([[0], np.cumsum(prob.blocks[l])])
        out.append([(int(b[i]), int(b[i + 1])) for i in range(len(prob.blocks[l]))])
    return out


def make_case(Ny, L, M, nblocks, seed, n_sources, precond):
    """(problem with host data, oracle problem in the original row order)."""
    prob = synth.make_problem(Ny, Ny, L, M, nblocks, seed=seed, n_sources=n_sources)
    synth.add_host_data(prob)
    if precond:
        pU = [uniform_weights(prob.u[l], prob.v[l], 2 * Ny, 2 * Ny) for l in range(L)]
    else:
        pU = [np.ones(len(u)) for u in prob.u]
    op = OracleProblem(Ny, Ny, prob.u, prob.v, prob.y, pU, block_ranges(prob), prob.epsilon,
                       g_scale=[1.0 / s for s in prob.sigma_noise])
    return prob, op


def oracle_op_norms(op, tol=1e-10, max_iter=500):
    """Per channel ||Phi_l||^2 and ||sqrt(pU) Phi_l||^2 by FP64 power iteration
    (``lib/utils/util_operator_norm.m:43-61``)."""
    son, nu2 = [], []
    for l in range(op.L):
        G = op.ops[l].G
        sq = np.sqrt(np.concatenate(op.pU[l]))

        def F(x, l=l):
            return G @ op.A[l](x)

        def Ft(y, l=l):
            return np.real(op.At[l](G.conj().T @ y))

        def F2(x, l=l):
            return sq * (G @ op.A[l](x))

        def Ft2(y, l=l):
            return np.real(op.At[l](G.conj().T @ (sq * y)))
        son.append(op_norm(F, Ft, (op.Ny, op.Nx), tol, max_iter)[0])
        nu2.append(op_norm(F2, Ft2, (op.Ny, op.Nx), tol, max_iter)[0])
    return np.array(son), np.array(nu2)


def solver_param(kind, Ny, L, son, nu2, geom=None, **over):
    """Parameter wiring of ``imaging/imaging.m:825-860`` for whitened data
    (unit noise std): gamma = gam sig, gamma0 = gam_bar sig_bar."""
    p = dict(DEFAULT_PARAM)
    p['nu2'] = np.asarray(nu2, dtype=np.float64)
    p['squared_operator_norm'] = np.asarray(son, dtype=np.float64)
    std = np.ones(L)
    if kind == 'sara':
        sig = compute_noise_level_sara(std[0], son[0])
        p.update(gamma=sig, sig=sig)
    elif kind == 'hs':
        sig, sig_bar, _, _, _ = compute_noise_level_hs(Ny, Ny, L, std, son)
        p.update(gamma=sig, sig=sig, gamma0=sig_bar, sig_bar=sig_bar)
    else:
        sig, _, _, _, _ = compute_noise_level_hs(Ny, Ny, L, std, son)
        pix = [geom['py'][q] * geom['px'][q] for q in range(len(geom['oy']))]
        # src/heuristics/compute_noise_level.m:57-88 ('fhs'): per-facet sig_bar
        sig_bar = np.array([np.sqrt(P * np.sum(std ** 2 / son) / min(P, L)) for P in pix])
        p.update(gamma=sig, sig=sig, gamma0=sig_bar, sig_bar=sig_bar)
    p.update(over)
    return p


# the solver cases with golden fixtures (tests/golden/solver_<name>.npz):
# name -> (kind, Ny, L, M per channel, blocks, seed, sources, preconditioned,
#          facets (Qy, Qx, overlap) or None, parameter overrides)
SOLVER_CASES = {
    'sara32': ('sara', 32, 1, 2000, 2, 5, 15, True, None, {'reweighting_max_iter': 3}),
    'hs32x4': ('hs', 32, 4, 1500, 2, 3, 20, True, None, {'reweighting_max_iter': 3}),
    'fhs32x4': ('fhs', 32, 4, 1500, 2, 13, 20, True, (2, 2, 0.25), {'reweighting_max_iter': 3}),
}


def case_geometry(name):
    kind, Ny, L, M, nb, seed, ns, pre, fac, over = SOLVER_CASES[name]
    if fac is None:
        return None
    return facet_geometry(Ny, Ny, fac[0], fac[1], (fac[2], fac[2]))
