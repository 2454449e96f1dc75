"""The MEX shims on the GPU, called as the reference's MATLAB solvers call
the functions they shadow, through the MEX API stand-in (tests/mex_mock/),
against the FP64 oracle: the data-fidelity step from a nonzero start with
the reference's own Fx_old = A(x_old) array (src/fhs/facetHyperSARA.m:574-576,
src/hs/hyperSARA.m:446) and then from the state it returns, the scaled
Fourier transform, the facet primal step and the facet low-rank step."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), 'mex_mock'))
import harness  # noqa: E402

pytestmark = pytest.mark.gpu

NY, NX, L = 32, 64, 2
BLOCKS = (300, 260)


@pytest.fixture(scope='module')
def mex(ctx):
    m = harness.Mex()
    yield m
    m.close()


@pytest.fixture(scope='module')
def problem(mex):
    from oracle.nufft import NufftOracle
    rng = np.random.default_rng(7)
    ops, Gh, y, pU, eps = [], [], [], [], []
    for l in range(L):
        M = sum(BLOCKS)
        u, v = rng.uniform(-np.pi, np.pi, M), rng.uniform(-np.pi, np.pi, M)
        op = NufftOracle(u, v, NY, NX)
        ops.append(op)
        Gh.append(mex.call('fhs_gmatrix', 1, op.G))
        b = np.concatenate([[0], np.cumsum(BLOCKS)])
        y.append([rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in BLOCKS])
        pU.append([rng.uniform(0.5, 2.0, n) for n in BLOCKS])
        eps.append([float(np.sqrt(n)) for n in BLOCKS])
        op.blocks = [(int(b[i]), int(b[i + 1])) for i in range(len(BLOCKS))]
    D = mex.call('fhs_data_state', 1, ops[0].sy, ops[0].sx, float(ops[0].Kty), float(ops[0].Ktx), Gh,
                 [np.array(BLOCKS, dtype=float) for _ in range(L)])
    return {'ops': ops, 'D': D, 'y': y, 'pU': pU, 'eps': eps, 'rng': rng}


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_data_state_row_precision_argument(mex, problem):
    """fhs_data_state's optional row_bits reaches fhs_data_set_row_precision:
    32 is accepted; 16 is refused for these CSR G (FP16 rows need tile-gridded
    separable G) with a MATLAB error, not a silent FP32 state."""
    op = problem['ops'][0]
    Gh = [mex.call('fhs_gmatrix', 1, o.G) for o in problem['ops']]
    blocks = [np.array(BLOCKS, dtype=float) for _ in range(L)]
    mex.call('fhs_data_state', 1, op.sy, op.sx, float(op.Kty), float(op.Ktx), Gh, blocks, 32.0)
    with pytest.raises(harness.MexError, match='row_precision|FP16'):
        mex.call('fhs_data_state', 1, op.sy, op.sx, float(op.Kty), float(op.Ktx), Gh, blocks, 16.0)


def test_apply_scaled_fourier_transform(mex, problem):
    op = problem['ops'][0]
    x = problem['rng'].random((NY, NX, L))
    F = mex.call('apply_scaled_fourier_transform', 1, x, problem['D'], float(op.No))
    assert F.shape == (op.No, L)
    for l in range(L):
        assert _rel(F[:, l], op.A(x[:, :, l])) < 1e-5


def _oracle_step(problem, x, Fx_old, v2, proj, sigma22):
    from oracle.data_fidelity import update_dual_data_fidelity
    ops = problem['ops']
    G = [[op.G[a:b] for (a, b) in op.blocks] for op in ops]
    W = [[np.arange(op.No) for _ in op.blocks] for op in ops]
    return update_dual_data_fidelity(v2, problem['y'], x, Fx_old, proj, [op.A for op in ops],
                                     [op.At for op in ops], G, W, problem['pU'], problem['eps'], 20, 1, 1e-8,
                                     sigma22)


def test_update_dual_data_fidelity_from_a_nonzero_start(mex, problem):
    """Two consecutive steps: the first with the reference's Fx_old = A(x_old)
    (a nonzero start, as after a warm start or a checkpoint resume), the
    second with the state the first returned, both against the oracle."""
    rng, ops = problem['rng'], problem['ops']
    x_old = rng.random((NY, NX, L))
    Fx_old = np.stack([op.A(x_old[:, :, l]) for l, op in enumerate(ops)], axis=1)
    x1 = x_old + 0.1 * rng.random((NY, NX, L))
    x2 = x1 + 0.1 * rng.random((NY, NX, L))
    v2 = [[0.01 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)) for n in BLOCKS] for _ in range(L)]
    proj = [[np.zeros(n, dtype=complex) for n in BLOCKS] for _ in range(L)]
    sigma22 = np.array([0.3, 0.2])
    D = problem['D']
    args = lambda v, x, fo, p: (v, problem['y'], x, fo, p, np.zeros((0, 0)), np.zeros((0, 0)), D, np.zeros((0, 0)),  # noqa: E731
                                problem['pU'], problem['eps'], 20.0, 1.0, 1e-8, sigma22, 0.0, np.zeros((0, 0)))
    # step 1: the reference-form Fx_old
    out = mex.call('update_dual_data_fidelity', 7, *args(v2, x1, Fx_old, proj))
    ref = _oracle_step(problem, x1, Fx_old, v2, proj, sigma22)
    _check(out, ref)
    # the shim forms G Fx_old on the device: a zero Fx_old would differ
    bad = mex.call('update_dual_data_fidelity', 7, *args(v2, x1, np.zeros((0, 0)), proj))
    assert _rel(np.concatenate([np.ravel(b) for b in bad[0][0]]),
                np.concatenate([np.ravel(b) for b in ref[0][0]])) > 1e-2
    # step 2: the state returned by step 1
    out2 = mex.call('update_dual_data_fidelity', 7, *args(out[0], x2, out[2], out[3]))
    ref2 = _oracle_step(problem, x2, ref[2], ref[0], ref[3], sigma22)
    _check(out2, ref2)


def _check(out, ref):
    v2, Ftx, state, proj, nres, gres, neps = out
    for l in range(L):
        for b in range(len(BLOCKS)):
            assert _rel(np.ravel(v2[l][b]), ref[0][l][b]) < 2e-4, (l, b)
            assert _rel(np.ravel(proj[l][b]), ref[3][l][b]) < 2e-4, (l, b)
            assert abs(float(np.ravel(nres[l][b])[0]) - ref[4][l][b]) <= 1e-4 * ref[4][l][b]
    assert _rel(Ftx, ref[1]) < 2e-4
    assert abs(gres - ref[5]) <= 1e-4 * ref[5]
    assert abs(neps - ref[6]) <= 1e-9 * ref[6]
    assert len(state) == L and all(np.size(s) == sum(BLOCKS) for s in state)


def test_update_dual_data_fidelity_rejects_mismatched_cells(mex, problem):
    rng = problem['rng']
    v2 = [[np.zeros(n, dtype=complex) for n in BLOCKS] for _ in range(L)]
    proj = [[np.zeros(n, dtype=complex) for n in BLOCKS] for _ in range(L)]
    pU_short = [[np.ones(n - 1) for n in BLOCKS] for _ in range(L)]
    x = rng.random((NY, NX, L))
    e = np.zeros((0, 0))
    base = [v2, problem['y'], x, e, proj, e, e, problem['D'], e, problem['pU'], problem['eps'], 20.0, 1.0, 1e-8,
            np.array([0.3, 0.2]), 0.0, e]
    for i, bad, msg in [(9, pU_short, 'pU'), (3, np.ones((5, L)), 'Fx_old'),
                        (3, [np.zeros(sum(BLOCKS) - 1, dtype=complex)] * L, 'Fx_old'),
                        (2, rng.random((NY, NX + 1, L)), 'x must')]:
        a = list(base)
        a[i] = bad
        with pytest.raises(harness.MexError, match=msg):
            mex.call('update_dual_data_fidelity', 7, *a)


def test_fhs_update_primal(mex):
    from oracle.prox import update_primal
    rng = np.random.default_rng(3)
    x = rng.standard_normal((20, 24, 3))
    g = rng.standard_normal((20, 24, 3))
    xs, xh, rel, nrm = mex.call('fhs_update_primal', 4, x, g)
    ref = update_primal(x, g)
    assert np.abs(xs - ref[0]).max() < 1e-6 and np.abs(xh - ref[1]).max() < 1e-5
    assert abs(rel - ref[2]) <= 1e-5 * ref[2] and abs(nrm - ref[3]) <= 1e-5 * ref[3]


def test_fhs_update_dual_lowrankness(mex):
    """One facet with an apodisation window against the reference's
    formula (src/fhs/fhs_update_dual_lowrankness.m:40-45) in FP64."""
    from oracle.prox import nuclear_prox_dual
    rng = np.random.default_rng(5)
    M, N, c = 20, 24, 6
    base = rng.random((M * N, 2)) @ rng.random((2, c))
    f32 = lambda a: a.astype(np.float32).astype(np.float64)      # noqa: E731  (the device holds FP32)
    xhat = f32((base + 0.05 * rng.random((M * N, c))).reshape((M, N, c), order='F'))
    v0 = f32(0.01 * rng.standard_normal((M * N, c)))
    w = f32(np.outer(np.linspace(0.2, 1.0, M), np.linspace(1.0, 0.3, N)))
    weights0 = np.linspace(1.0, 2.0, c)
    beta0 = 0.8
    v0n, g0 = mex.call('fhs_update_dual_lowrankness', 2, v0, xhat, w, weights0, beta0)
    Z = v0 + (xhat * w[:, :, None]).reshape((M * N, c), order='F')
    v_ref = nuclear_prox_dual(Z, beta0 * weights0)
    g_ref = w[:, :, None] * v_ref.reshape((M, N, c), order='F')
    assert _rel(v0n, v_ref) < 1e-5 and _rel(g0, g_ref) < 1e-5
    # the threshold is active on some singular values and not on others
    s = np.linalg.svd(Z, compute_uv=False)
    assert np.any(s > beta0 * weights0) and np.any(s < beta0 * weights0)


def test_faceted_wavelet_shims(mex):
    """sdwt2_setup, sdwt2_sara_faceting and isdwt2_sara_faceting (the
    faceted-wavelet submodule's functions, absent from the reference tree)
    used exactly as src/fhs/fhs_update_dual_sparsity.m:65-93 uses them: the
    facet's extended image built from dims_overlap_ref / offsetL / offsetR,
    its coefficients against the oracle's faceted transform, their count
    against the solver's own formula on Ncoefs, and the facets'
    contributions over dims_overlap adding up to the whole-image synthesis."""
    from oracle.wavelets import SaraFacets
    rng = np.random.default_rng(3)
    Ny, Nx, L = 64, 48, 2
    ys, xs = [0, 20, 41, 64], [0, 25, 48]
    Qy, Qx = len(ys) - 1, len(xs) - 1
    Q = Qy * Qx
    I = np.array([[ys[q % Qy], xs[q // Qy]] for q in range(Q)], dtype=float)
    dims = np.array([[ys[q % Qy + 1] - ys[q % Qy], xs[q // Qy + 1] - xs[q // Qy]] for q in range(Q)], dtype=float)
    wavelet = ['db1', 'db2', 'db3', 'db4', 'db5', 'db6', 'db7', 'db8', 'self']
    flen = np.array([2, 4, 6, 8, 10, 12, 14, 16, 0], dtype=float)
    nlevel = 4
    (I_ref, dims_ref, I_ov, dims_ov, status, offset, offL, offR, Ncoefs, temL,
     temR) = mex.call('sdwt2_setup', 11, np.array([[Ny, Nx]], dtype=float), I, dims, float(nlevel), wavelet, flen)
    ref = SaraFacets(Ny, Nx, ys, xs)
    x = rng.random((Ny, Nx, L))
    c = rng.standard_normal((ref.glob.s, L))
    total = np.zeros((Ny, Nx, L))
    for q in range(Q):
        oy, ox = int(I_ref[q, 0]), int(I_ref[q, 1])
        ny, nx = int(dims_ref[q, 0]), int(dims_ref[q, 1])
        assert np.array_equal(np.asarray(dims_ov[q]).ravel(), [ny, nx])
        p = np.prod(np.asarray(Ncoefs[q]), axis=1)
        sz = 3 * p.sum() - 2 * p[nlevel::nlevel + 1].sum() - 2 * p[-1]    # fhs_initialize_dual_and_weights.m:80-81
        for l in range(L):
            x_facet = x[oy:oy + ny, ox:ox + nx, l]
            x_curr = np.zeros((ny + int(offL[q, 0] + offR[q, 0]), nx + int(offL[q, 1] + offR[q, 1])))
            x_curr[int(offL[q, 0]):int(offL[q, 0]) + ny, int(offL[q, 1]):int(offL[q, 1]) + nx] = x_facet
            z = mex.call('sdwt2_sara_faceting', 1, x_curr, I[q], offset, float(status[q, 0]), float(nlevel), wavelet,
                         Ncoefs[q])
            zr = ref.analysis(q, ref.extended(q, x[:, :, l]))
            assert len(np.ravel(z)) == sz == len(zr)
            assert _rel(np.ravel(z), zr) < 1e-5, (q, l)
            g = mex.call('isdwt2_sara_faceting', 1, c[ref.facets[q]['global'], l], I[q], dims[q], I_ov[q], dims_ov[q],
                         Ncoefs[q], float(nlevel), wavelet, temL[q], temR[q])
            assert g.shape == (ny, nx)
            total[oy:oy + ny, ox:ox + nx, l] += g
    for l in range(L):
        assert _rel(total[:, :, l], ref.glob.psi(c[:, l])) < 1e-5, l
