"""Whole solves on the GPU against the FP64 oracle, through the reweighting
steps to convergence (golden fixtures of tests/golden/make_solver_fixtures.py:
the oracle's full PDFB + reweighting loop on the same seeded data).

Bar: the same decisions -- every reweighting at the same iteration and the
same total count -- and the final reconstruction SNR within 0.1 dB of the
oracle's, for SARA (stopping rule of src/sara/sara.m:666-670), HyperSARA
(src/hs/hyperSARA.m:682-686) and Faceted HyperSARA (the fidelity-gated rule
of src/fhs/facetHyperSARA.m:877-881).

Also here: the residual images at x != 0 (src/compute_residual_images.m:
39-103, run at every reweighting) against the oracle, and the engine's
report of a timed-out ellipsoid-projection exchange (forced from the host).
"""
import os

import numpy as np
import pytest
import torch

import solver_cases as sc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _fixture(name):
    return np.load(os.path.join(GOLDEN, f'solver_{name}.npz'))


@pytest.mark.parametrize('name', list(sc.SOLVER_CASES))
def test_solve_to_convergence_matches_oracle(ctx, name):
    from fhs import solvers, synth
    from fhs.lowrank import facet_geometry
    kind, Ny, L, M, nb, seed, ns, pre, fac, over = sc.SOLVER_CASES[name]
    fx = _fixture(name)
    prob = synth.make_problem(Ny, Ny, L, M, nb, seed=seed, n_sources=ns)
    synth.add_host_data(prob)
    meas = solvers.Measurement(ctx, prob, precond=pre)
    param = dict(over, squared_operator_norm=fx['son'], nu2=fx['nu2'], track_snr=True)
    print(f'[{name}] solving', flush=True)
    if kind == 'sara':
        x, info = solvers.sara(meas, param, X0=prob.X0)
        x = x[:, :, None]
    elif kind == 'hs':
        x, info = solvers.hyper_sara(meas, param, X0=prob.X0)
    else:
        geom = facet_geometry(Ny, Ny, fac[0], fac[1], (fac[2], fac[2]))
        for k in ('oy', 'ox', 'py', 'px'):      # the product's facets are the oracle's
            assert list(geom[k]) == list(fx['geom_' + k]), k
        x, info = solvers.facet_hyper_sara(meas, geom, param, X0=prob.X0)
    ref_rw = [int(t) for t in fx['reweights']]
    print(f'[{name}] GPU: {info["iterations"]} iterations, reweights {info["reweights"]}, '
          f'SNR {info["snr"][-1]:.4f}; oracle: {int(fx["iterations"])}, {ref_rw}, {fx["snr"][-1]:.4f}', flush=True)
    assert info['converged'] and bool(fx['converged'])
    assert len(ref_rw) >= 2
    assert info['reweights'] == ref_rw
    assert info['iterations'] == int(fx['iterations'])
    assert abs(info['snr'][-1] - fx['snr'][-1]) <= 0.1
    rel = np.linalg.norm(x - fx['x']) / np.linalg.norm(fx['x'])
    assert rel < 1e-2, rel
    # the per-iteration relative variation follows the oracle's
    n = len(fx['rel_val'])
    assert np.allclose(info['rel_val'][5:n], fx['rel_val'][5:], rtol=2e-2, atol=1e-7)


def test_residual_images_match_oracle(ctx):
    """compute_residual_images at x != 0 (the reweighting-step residual)."""
    from fhs import solvers, synth
    from oracle.data_fidelity import compute_residual_images
    prob, op = sc.make_case(64, 3, 4000, 2, 29, 25, False)
    meas = solvers.Measurement(ctx, prob, precond=False)
    rng = np.random.default_rng(1)
    x = prob.X0 * (1.0 + 0.2 * rng.standard_normal(prob.X0.shape))
    got = solvers._uncube(meas.residual_images(solvers._cube(x)))
    ref = compute_residual_images(x, op.y, op.A, op.At, op.G, op.W)
    for l in range(prob.L):
        e = np.linalg.norm(got[:, :, l] - ref[:, :, l]) / np.linalg.norm(ref[:, :, l])
        assert e < 1e-4, (l, e)


def test_fb_exchange_timeout_is_reported(ctx):
    """A timed-out workgroup exchange of the ellipsoid projection (flag set
    from the host, as a real time-out sets it) makes the next iteration
    fail loudly instead of reading as 'not converged', and the following
    call runs normally again (the exchange counters were reset)."""
    from fhs import solvers, synth
    from fhs._lib import FhsError
    from fhs.lowrank import LowRankPrior
    from fhs.sara import SaraDictionary
    prob = synth.make_problem(32, 32, 2, 3000, 2, seed=31, n_sources=10)
    synth.add_host_data(prob)
    meas = solvers.Measurement(ctx, prob, precond=True)
    eng = solvers.PdfbEngine(meas, SaraDictionary(ctx, 32, 32), LowRankPrior(ctx, 32, 32, 2))
    eng.set_params(np.ones((1, 2)), 1e-3, 0.33, 0.33, np.full(2, 1e-4), meas.epsilon, 20, 1, 1e-8)
    ok = eng.iterate(3)
    assert np.all(np.isfinite(ok))
    eng.ctx.check(solvers.lib().fhs_pdfb_inject_fb_timeout(eng.h), 'inject')
    with pytest.raises(FhsError, match='timed out'):
        eng.iterate(1)
    again = eng.iterate(2)
    assert np.all(np.isfinite(again))
    # the same through fhs_pdfb_run (the whole-cycle call)
    eng.ctx.check(solvers.lib().fhs_pdfb_inject_fb_timeout(eng.h), 'inject')
    with pytest.raises(FhsError, match='timed out'):
        eng.run(0, 0, solvers.DEFAULT_PARAM, 5)
    out, _ = eng.run(0, 0, solvers.DEFAULT_PARAM, 3)
    assert np.all(np.isfinite(out))


@pytest.mark.parametrize('kind', ['hs', 'fhs'])
def test_dropin_form_matches_engine(ctx, kind):
    """The drop-in form (the reference-named entry points called one after
    the other every iteration) takes the engine's decisions and iterates."""
    from fhs import solvers, synth
    from fhs.lowrank import facet_geometry
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=37, n_sources=30)
    synth.add_host_data(prob)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    p = {'squared_operator_norm': son, 'nu2': nu2, 'track_snr': True, 'reweighting_max_iter': 2}
    facets = facet_geometry(64, 64, 2, 2, (0.25, 0.25)) if kind == 'fhs' else None
    _, a = solvers._solve(kind, meas, p, X0=prob.X0, facets=facets)
    eb, b = solvers._solve(kind, meas, p, X0=prob.X0, facets=facets, form='dropin')
    assert a['reweights'] and a['reweights'] == b['reweights'] and a['iterations'] == b['iterations']
    assert np.allclose(a['snr'], b['snr'], rtol=0, atol=1e-5)
    assert np.allclose(np.array(a['rel_val']), np.array(b['rel_val']), rtol=1e-4, equal_nan=True)


@pytest.mark.parametrize('gram,rows', [('fp64', 32), ('bf16x3', 32), ('bf16x3', 16)])
def test_hs512x16_solve_matches_fp64_reference(ctx, gram, rows):
    """The benchmark workload (HyperSARA 512x512x16, 3 reweighting steps)
    solved on the GPU ends within 0.1 dB of the FP64 oracle's converged SNR
    on the same data, with the oracle's reweighting iterations (fixture of
    tests/golden/make_hs512_reference.py); also with the mixed-precision
    (BF16 split, matrix-core) Gram, and with it the FP16 G rows, of the
    chosen form."""
    path = os.path.join(GOLDEN, 'hs512x16_reference.npz')
    if not os.path.exists(path):
        pytest.skip('no FP64 reference fixture for the 512x512x16 workload')
    from fhs import solvers, synth
    fx = np.load(path)
    prob = synth.make_problem(512, 512, 16, 200000, 4, seed=0)
    synth.add_host_data(prob)
    meas = solvers.Measurement(ctx, prob, precond=True)
    p = {'squared_operator_norm': fx['son'], 'nu2': fx['nu2'], 'track_snr': True, 'reweighting_max_iter': 3}
    _, info = solvers.hyper_sara(meas, p, X0=prob.X0, gram=gram, rows=rows)
    ref_rw = [int(t) for t in fx['reweights']]
    print(f'GPU ({gram} Gram, FP{rows} rows) {info["iterations"]} iterations, reweights {info["reweights"]}, SNR {info["snr"][-1]:.4f}; '
          f'oracle {int(fx["iterations"])}, {ref_rw}, {fx["snr"][-1]:.4f}', flush=True)
    assert info['converged'] and bool(fx['converged'])
    assert abs(info['snr'][-1] - float(fx['snr'][-1])) <= 0.1
    assert len(info['reweights']) == len(ref_rw)
    assert all(abs(a - b) <= 2 for a, b in zip(info['reweights'], ref_rw))
