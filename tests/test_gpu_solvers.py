"""End-to-end parity: GPU HyperSARA / SARA / Faceted HyperSARA iterations
against the FP64 oracle solver on the same seeded problem.

Bar: after K iterations the GPU iterate matches the oracle's to 1e-3
relative (FP32 operators, error grows slowly with iterations) and the
reconstruction SNR to 0.1 dB; both solvers use the same operator norms."""
import numpy as np
import pytest
import torch

from oracle.solvers import OracleProblem, hypersara as hs_ref, sara_solver as sara_ref, snr

pytestmark = pytest.mark.gpu


def oracle_problem(prob, meas):
    """The oracle problem on the measurement's own (block-sorted) row order."""
    blocks = []
    for l in range(prob.L):
        b = np.concatenate([[0], np.cumsum(prob.blocks[l])])
        blocks.append([(int(b[i]), int(b[i + 1])) for i in range(len(prob.blocks[l]))])
    pU = [np.asarray(p, dtype=np.float32).astype(np.float64) for p in meas.pU_host]
    y = [np.asarray(yl).astype(np.complex64).astype(np.complex128) for yl in meas.y_host]
    return OracleProblem(prob.Ny, prob.Nx, meas.u, meas.v, y, pU, blocks, prob.epsilon,
                         g_scale=[1.0 / s for s in prob.sigma_noise])


# 40 x 48 has an 80 x 96 grid (16 5 and 32 3): the mixed-radix FFT path and
# the untiled G layouts
@pytest.mark.parametrize('precond,Ny,Nx', [(True, 32, 32), (False, 32, 32), (True, 40, 48)])
def test_hypersara_matches_oracle(ctx, precond, Ny, Nx):
    import fhs
    from fhs import solvers, synth
    prob = synth.make_problem(Ny, Nx, 4, 1500, 2, seed=3, n_sources=20)
    meas = solvers.Measurement(ctx, prob, precond=precond)
    son = meas.op_norm(False)
    nu2 = meas.op_norm(precond)
    K = 40
    param = {'squared_operator_norm': son, 'nu2': nu2}
    x_gpu, info = solvers.hyper_sara(meas, param, X0=prob.X0, max_total_iter=K)
    op = oracle_problem(prob, meas)
    sig, sig_bar = solvers.noise_levels_hs(Ny, Nx, 4, np.ones(4), son)
    pr = dict(solvers.DEFAULT_PARAM)
    pr.update({'nu2': nu2, 'gamma': sig, 'gamma0': sig_bar, 'sig': sig, 'sig_bar': sig_bar,
               'alpha': 1.0})
    x_ref, rinfo = hs_ref(op, pr, max_iter_total=K)
    rel = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    assert rel < 1e-3, rel
    assert abs(snr(prob.X0, x_gpu) - snr(prob.X0, x_ref)) < 0.1
    assert np.allclose(info['rel_val'][5:], rinfo['rel_val'][5:], rtol=1e-2)


def test_sara_matches_oracle(ctx):
    from fhs import solvers, synth
    prob = synth.make_problem(32, 32, 1, 2000, 2, seed=5, n_sources=15)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son = meas.op_norm(False)
    nu2 = meas.op_norm(True)
    K = 40
    x_gpu, info = solvers.sara(meas, {'squared_operator_norm': son, 'nu2': nu2}, max_total_iter=K)
    op = oracle_problem(prob, meas)
    sig = np.sqrt(1.0 / son[0])
    pr = dict(solvers.DEFAULT_PARAM)
    pr.update({'nu2': nu2, 'gamma': sig, 'pdfb_max_iter': K})
    x_ref, _ = sara_ref(op, pr, max_iter_total=K)
    rel = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    assert rel < 1e-3, rel


def test_hipgraph_replay_matches_eager(ctx):
    """The captured iteration replays to the same iterates as eager launches."""
    from fhs import solvers, synth
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=11, n_sources=30)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    p = {'squared_operator_norm': son, 'nu2': nu2}
    xg, ig = solvers.hyper_sara(meas, p, max_total_iter=25, graph=True)
    xe, ie = solvers.hyper_sara(meas, p, max_total_iter=25, graph=False)
    assert np.array_equal(xg, xe)
    assert np.array_equal(np.array(ig['rel_val']), np.array(ie['rel_val']), equal_nan=True)


def test_facet_hypersara_matches_oracle(ctx):
    from fhs import solvers, synth
    from oracle.facets import facet_geometry
    from oracle.solvers import facet_hypersara as fhs_ref
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=13, n_sources=30)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    geom = facet_geometry(64, 64, 2, 2, (0.25, 0.25))
    K = 30
    x_gpu, info = solvers.facet_hyper_sara(meas, geom, {'squared_operator_norm': son, 'nu2': nu2},
                                           max_total_iter=K)
    op = oracle_problem(prob, meas)
    pix = [geom['py'][q] * geom['px'][q] for q in range(4)]
    sig, sig_bar = solvers.noise_levels_fhs(4, np.ones(4), son, pix)
    pr = dict(solvers.DEFAULT_PARAM)
    pr.update({'nu2': nu2, 'gamma': sig, 'gamma0': sig_bar, 'sig': sig, 'sig_bar': sig_bar, 'alpha': 1.0})
    x_ref, _ = fhs_ref(op, pr, geom, max_iter_total=K)
    rel = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    assert rel < 1e-3, rel
    assert abs(snr(prob.X0, x_gpu) - snr(prob.X0, x_ref)) < 0.1


def test_facet_hypersara_iterations_at_1024_geometry(ctx):
    """Faceted HyperSARA iterations at the 1024x1024 configuration's facet
    geometry (4x4 facets, 25% overlap, uneven facet sizes) against the
    oracle's: three channels, three iterations, all the steps (primal,
    faceted low-rank, sparsity, preconditioned data step)."""
    from fhs import solvers, synth
    from oracle.facets import facet_geometry
    from oracle.solvers import facet_hypersara as fhs_ref
    N, L = 1024, 3
    prob = synth.make_problem(N, N, L, 40000, 2, seed=23, n_sources=60)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    geom = facet_geometry(N, N, 4, 4, (0.25, 0.25))
    K = 3
    x_gpu, info = solvers.facet_hyper_sara(meas, geom, {'squared_operator_norm': son, 'nu2': nu2},
                                           max_total_iter=K, graph=False)
    op = oracle_problem(prob, meas)
    pix = [geom['py'][q] * geom['px'][q] for q in range(16)]
    sig, sig_bar = solvers.noise_levels_fhs(L, np.ones(L), son, pix)
    pr = dict(solvers.DEFAULT_PARAM)
    pr.update({'nu2': nu2, 'gamma': sig, 'gamma0': sig_bar, 'sig': sig, 'sig_bar': sig_bar, 'alpha': 1.0})
    x_ref, rinfo = fhs_ref(op, pr, geom, max_iter_total=K)
    rel = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    print(f'FHS 1024x1024x{L}, 4x4 facets, {K} iterations: rel {rel:.2e}')
    assert rel < 1e-4, rel
    assert np.allclose(info['rel_val'][1:], rinfo['rel_val'][1:], rtol=1e-3)


def test_concurrent_streams_match_sequential(ctx):
    """Priors forked onto the side stream give the same iterates as in order."""
    from fhs import solvers, synth
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=17, n_sources=30)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    p = {'squared_operator_norm': son, 'nu2': nu2}
    xa, _ = solvers.hyper_sara(meas, p, max_total_iter=20, concurrent=True)
    xb, _ = solvers.hyper_sara(meas, p, max_total_iter=20, concurrent=False)
    assert np.array_equal(xa, xb)


def test_fused_snr_tracking(ctx):
    """The SNR summed inside the primal update equals the SNR of the cube,
    computed in FP64 on the host, at every iteration (graph replay included),
    and following it does not change the iterates."""
    from fhs import solvers, synth
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=19, n_sources=30)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    p = {'squared_operator_norm': son, 'nu2': nu2, 'track_snr': True}
    host = []

    def cb(t, eng, info):
        host.append(snr(prob.X0, solvers._uncube(eng.x)))

    x_t, info = solvers.hyper_sara(meas, p, X0=prob.X0, max_total_iter=15, callback=cb)
    assert len(info['snr']) == len(host) == 15
    # X0 is rounded to FP32 once; the sums are FP64
    assert np.allclose(info['snr'], host, rtol=0, atol=1e-4), (info['snr'], host)
    x_u, _ = solvers.hyper_sara(meas, {'squared_operator_norm': son, 'nu2': nu2}, max_total_iter=15)
    assert np.array_equal(x_t, x_u)


def test_library_cycles_match_per_iteration_loop(ctx):
    """Cycles run inside the library (fhs_pdfb_run) take the same decisions
    and iterates as the per-iteration loop (forced by a callback), including
    the reweighting steps, and stop_snr ends both at the same iteration."""
    from fhs import solvers, synth
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=23, n_sources=30)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    p = {'squared_operator_norm': son, 'nu2': nu2, 'track_snr': True, 'pdfb_max_iter': 40,
         'reweighting_max_iter': 2}
    x_a, ia = solvers.hyper_sara(meas, p, X0=prob.X0)
    x_b, ib = solvers.hyper_sara(meas, p, X0=prob.X0, callback=lambda t, eng, info: None)
    assert ia['iterations'] == ib['iterations'] and ia['reweights'] == ib['reweights']
    assert ia['reweights'], 'the case should reweight'
    assert np.array_equal(x_a, x_b)
    assert np.array_equal(np.array(ia['rel_val']), np.array(ib['rel_val']), equal_nan=True)
    assert np.allclose(ia['snr'], ib['snr'], rtol=0, atol=1e-9)
    target = ia['snr'][len(ia['snr']) // 2]
    _, sa = solvers._solve('hs', meas, p, X0=prob.X0, stop_snr=target)
    _, sb = solvers._solve('hs', meas, p, X0=prob.X0, stop_snr=target, callback=lambda t, eng, info: None)
    assert sa['iterations'] == sb['iterations'] and sa['snr'][-1] >= target
