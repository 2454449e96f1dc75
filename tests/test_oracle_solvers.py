"""Pins of the oracle's solver loop (oracle/solvers.py), CPU only.

1. The stopping and reweighting rules, driven through every branch by
   hand-built values and checked against the text of the reference:
     sara  src/sara/sara.m:666-670           (no data-fidelity term)
     hs    src/hs/hyperSARA.m:682-686        (no data-fidelity term)
     fhs   src/fhs/facetHyperSARA.m:877-881  (rel_var branch gated on ||res|| <= 1.01 eps)
     reweighting: sara.m:713-717, hyperSARA.m:748-752, facetHyperSARA.m:965-969.
   The library's host rule (fhs_pdfb_stopping_rule, what fhs_pdfb_run
   evaluates) is checked against the oracle's on the same grid.
2. The whole loop, by the optimality conditions of the problem the solvers
   solve, at a converged iterate of a tiny problem:
     min  gamma0 sum_q ||w_q . x_q||_{*,omega0} + gamma ||Psi' x||_{2,1,omega1}
     s.t. ||Phi_b x - y_b|| <= eps_b,  x >= 0.
   The checks assemble the gradient from the returned duals with the tests'
   own operator calls and use only the definitions of the subdifferentials
   (Fenchel equalities, shared singular vectors, the normal cone of a
   ball), not the oracle's prox code.  A planted mistake in the oracle
   (a wrong threshold scale, a flipped primal step) fails them.
"""
import ctypes
import math

import numpy as np
import pytest

import solver_cases as sc
from oracle import solvers as osol
from oracle.facets import facet_geometry
from oracle.wavelets import SaraDictionary

P = dict(sc.DEFAULT_PARAM)
NAN = float('nan')


# ------------------------------------------------------------- 1. rules --

@pytest.mark.parametrize('kind', ['sara', 'hs', 'fhs'])
def test_pdfb_rule_branches(kind):
    """(k, rel_val, norm_res, norm_eps) -> converged, from the .m text."""
    fid = kind == 'fhs'
    cases = [
        # below min_iter: never, whatever the rest says
        ((9, 0.0, 0.0, 1.0), False),
        ((9, NAN, 0.0, 1.0), False),
        # max_iter reached
        ((2000, 1.0, 9.0, 1.0), True),
        ((2001, NAN, NAN, 1.0), True),
        # relative variation small, data fidelity met
        ((10, 4e-5, 1.0, 1.0), True),
        ((10, 5e-5, 1.01, 1.0), True),            # both at their bounds (<=)
        # relative variation small, data fidelity NOT met: only FHS refuses
        ((10, 4e-5, 1.02, 1.0), not fid),
        ((500, 1e-5, 5.0, 1.0), not fid),
        # relative variation really small: converged whatever the fidelity
        ((10, 5e-6, 5.0, 1.0), True),
        ((10, 4e-6, 5.0, 1.0), True),
        # relative variation too large
        ((10, 6e-5, 1.0, 1.0), False),
        ((1999, 1e-3, 0.5, 1.0), False),
        # NaN relative variation (x = 0) compares false
        ((10, NAN, 1.0, 1.0), False),
    ]
    for (k, rel, res, eps), want in cases:
        got = osol.pdfb_converged(kind, k, rel, res, eps, P)
        assert got == want, (kind, k, rel, res, eps, got)


def test_reweighting_rule_branches():
    p = dict(P)
    assert not osol.reweighting_converged(0, 0.0, p)          # min 1 reweighting
    assert osol.reweighting_converged(1, 1e-4, p)             # rel <= 1e-4
    assert not osol.reweighting_converged(1, 1.1e-4, p)
    assert osol.reweighting_converged(20, 1.0, p)             # max count reached
    assert not osol.reweighting_converged(19, 1.0, p)
    assert not osol.reweighting_converged(5, math.inf, p)
    assert not osol.reweighting_converged(5, NAN, p)
    assert osol.reweighting_converged(0, 1.0, p, reweighting_flag=False)   # ~reweighting_flag
    p0 = dict(P, reweighting_min_iter=0, reweighting_max_iter=0)
    assert osol.reweighting_converged(0, math.inf, p0)


def test_rel_x_reweighting_special_values():
    z = np.zeros((4, 4, 2))
    x = np.ones((4, 4, 2))
    assert osol.rel_x_reweighting(z, x) == math.inf           # norm(x)/0 in MATLAB
    assert math.isnan(osol.rel_x_reweighting(z, z))           # 0/0
    assert osol.rel_x_reweighting(x, 1.5 * x) == pytest.approx(0.5)


def test_library_rule_matches_oracle():
    """fhs_pdfb_stopping_rule (host code, no GPU) against the oracle."""
    from fhs import _lib
    lib = _lib.lib()
    ks = [0, 9, 10, 11, 1999, 2000, 2500]
    rels = [NAN, 0.0, 1e-6, 5e-6, 5.0001e-6, 4e-5, 5e-5, 5.1e-5, 1e-2]
    ratios = [NAN, 0.5, 1.0, 1.01, 1.0100001, 2.0]
    for rule, kind in ((0, 'sara'), (0, 'hs'), (1, 'fhs')):
        for k in ks:
            for rel in rels:
                for r in ratios:
                    want = osol.pdfb_converged(kind, k, rel, r, 1.0, P)
                    got = bool(lib.fhs_pdfb_stopping_rule(
                        rule, k, rel, r, 1.0, P['pdfb_min_iter'], P['pdfb_max_iter'], P['pdfb_rel_var'],
                        P['pdfb_fidelity_tolerance'], P['pdfb_rel_var_low']))
                    assert got == want, (rule, k, rel, r)


def test_product_driver_rule_per_solver():
    """The GPU driver evaluates SARA and HyperSARA with rule 0 and Faceted
    HyperSARA with rule 1 (the fidelity-gated one)."""
    import inspect
    from fhs import solvers
    src = inspect.getsource(solvers._solve)
    assert "rule = 1 if kind == 'fhs' else 0" in src
    assert solvers.pdfb_converged(0, 10, 4e-5, 2.0, 1.0, solvers.DEFAULT_PARAM)
    assert not solvers.pdfb_converged(1, 10, 4e-5, 2.0, 1.0, solvers.DEFAULT_PARAM)


# ---------------------------------------------- 2. optimality conditions --

BASES = ('db1', 'db2', 'self')
NLEVEL = 2


def _converged(kind, iters, reweights=0, L=2):
    L = 1 if kind == 'sara' else L
    prob, op = sc.make_case(16, L, 300, 1, 4, 6, False)
    son, nu2 = sc.oracle_op_norms(op)
    geom = facet_geometry(16, 16, 2, 2, (0.25, 0.25)) if kind == 'fhs' else None
    p = sc.solver_param(kind, 16, L, son, nu2, geom, bases=BASES, nlevel=NLEVEL, pdfb_rel_var=0.0,
                        pdfb_rel_var_low=0.0, pdfb_max_iter=iters, reweighting_max_iter=reweights,
                        reweighting_min_iter=reweights)
    x, info = osol.pdfb(kind, op, p, geom=geom)
    assert info['converged'] and len(info['reweights']) == reweights
    return x, info, p, op, geom


def _kkt(kind, x, info, p, op, geom):
    """Residuals of the optimality conditions (see the module docstring)."""
    L = x.shape[2]
    out = {}
    sara = SaraDictionary(16, 16, BASES, NLEVEL)
    # sparsity prior: v1 in the subdifferential of ||.||_{2,1} weighted by
    # t = gamma w1 at R = Psi' x:  ||v1_r|| <= t_r and <v1, R> = sum t_r ||R_r||
    t1 = p['gamma'] * info['w1']
    R = np.stack([sara.psit(x[:, :, l]) for l in range(L)], axis=1)
    v1 = info['v1']
    nr, nv = np.linalg.norm(R, axis=1), np.linalg.norm(v1, axis=1)
    out['l21_ball'] = (nv / t1).max() - 1.0
    den = np.sum(t1 * nr)
    out['l21_fenchel'] = (den - np.sum(v1 * R)) / den if den > 0 else 0.0
    grad = np.stack([sara.psi(v1[:, l]) for l in range(L)], axis=2)
    scale = np.abs(grad).max()
    scale = scale if scale > 0 else 1.0      # a wrecked iterate (planted mistakes) can zero v1
    # low-rank prior: v0_q shares the singular vectors of X_q = w_q . x_q,
    # equal to the threshold on its active singular values, below it elsewhere
    nuc = []
    if kind in ('hs', 'fhs'):
        Q = 1 if kind == 'hs' else len(geom['oy'])
        for q in range(Q):
            if kind == 'hs':
                y0, x0, ny, nx, w = 0, 0, 16, 16, np.ones((16, 16))
            else:
                y0, x0, ny, nx = geom['oy'][q], geom['ox'][q], geom['py'][q], geom['px'][q]
                w = geom['windows'][q]
            X = (w[:, :, None] * x[y0:y0 + ny, x0:x0 + nx, :]).reshape((-1, L), order='F')
            U, S, Vt = np.linalg.svd(X, full_matrices=False)
            t0 = np.atleast_1d(p['gamma0'])[q] * info['w0'][q]
            D = U.T @ info['v0'][q] @ Vt.T
            act = S > 1e-6 * S[0]
            nuc.append(np.abs(D - np.diag(np.diag(D))).max() / t0.max())
            if act.any():
                nuc.append(np.abs(np.diag(D)[act] - t0[act]).max() / t0.max())
            if (~act).any():
                nuc.append(max(0.0, (np.diag(D)[~act] / t0[~act]).max() - 1.0))
            grad[y0:y0 + ny, x0:x0 + nx, :] += w[:, :, None] * info['v0'][q].reshape((ny, nx, L), order='F')
        out['nuclear'] = max(nuc)
    # data term: v2_b in the normal cone of the ball B(y_b, eps_b) at Phi_b x:
    # feasible, and Re<v2, Phi x> = max_{p in B} Re<v2, p> = Re<v2, y> + eps ||v2||
    feas, cone = [], []
    for l in range(L):
        gd = np.zeros(op.No, dtype=np.complex128)
        for G, y, v, e in zip(op.G[l], op.y[l], info['v2'][l], op.epsilon[l]):
            Fx = G @ op.A[l](x[:, :, l])
            feas.append(np.linalg.norm(Fx - y) / e - 1.0)
            cone.append(abs(np.real(np.vdot(v, y)) + e * np.linalg.norm(v) - np.real(np.vdot(v, Fx)))
                        / (e * np.linalg.norm(v)))
            gd += G.conj().T @ v
        # any positive multiple of v2 lies in the cone: the step ratio of the loop
        grad[:, :, l] += info['sigma22'][l] / info['tau'] * np.real(op.At[l](gd))
    out['feasibility'] = max(feas)
    out['normal_cone'] = max(cone)
    # primal: -grad in the normal cone of x >= 0 (zero on the support, >= 0 off it)
    sup = x > 0
    out['stationarity'] = np.abs(grad[sup]).max() / scale if sup.any() else 0.0
    out['positivity'] = max(0.0, -grad[~sup].min() / scale) if (~sup).any() else 0.0
    # the loop's own gradient is the one assembled here (times tau)
    out['gradient'] = np.abs(info['g'] - info['tau'] * grad).max() / scale
    return out


TOL = {'l21_ball': 1e-9, 'l21_fenchel': 1e-6, 'nuclear': 1e-6, 'feasibility': 1e-6, 'normal_cone': 1e-6,
       'stationarity': 1e-3, 'positivity': 1e-6, 'gradient': 1e-12}


def _assert_kkt(r, tol):
    bad = {k: v for k, v in r.items() if not v <= tol[k]}
    assert not bad, (bad, r)


def test_kkt_hypersara():
    x, info, p, op, geom = _converged('hs', 2000)
    _assert_kkt(_kkt('hs', x, info, p, op, geom), TOL)


def test_kkt_sara():
    x, info, p, op, geom = _converged('sara', 2500)
    _assert_kkt(_kkt('sara', x, info, p, op, geom), TOL)


def test_kkt_facet_hypersara_after_a_reweighting():
    """Two PDFB cycles: the second solves the reweighted problem (weights
    from the first cycle's end), so the thresholds are not uniform."""
    x, info, p, op, geom = _converged('fhs', 2500, reweights=1)
    assert np.ptp(info['w1']) > 0.1 and np.ptp(info['w0']) > 0.1
    tol = dict(TOL, l21_fenchel=1e-4, nuclear=1e-5, feasibility=1e-4, normal_cone=1e-4, stationarity=5e-3)
    _assert_kkt(_kkt('fhs', x, info, p, op, geom), tol)


def test_kkt_detects_planted_mistakes(monkeypatch):
    """The same checks fail on an oracle with a planted mistake."""
    real_soft = osol.soft_l21
    monkeypatch.setattr(osol, 'soft_l21', lambda w, t: real_soft(w, 0.5 * t))      # threshold halved
    x, info, p, op, geom = _converged('hs', 800)
    r = _kkt('hs', x, info, p, op, geom)
    assert r['l21_fenchel'] > 1e-2 or r['stationarity'] > 1e-2, r
    monkeypatch.setattr(osol, 'soft_l21', real_soft)
    real_primal = osol.update_primal

    def flipped(xsol, g):     # x <- max(x + g, 0): a sign error of the primal step
        return real_primal(xsol, -g)
    monkeypatch.setattr(osol, 'update_primal', flipped)
    x, info, p, op, geom = _converged('hs', 300)
    r = _kkt('hs', x, info, p, op, geom)
    assert r['stationarity'] > 1e-2 or r['feasibility'] > 1e-2, r
