"""FP16 storage of the interpolation rows in a data state
(fhs_data_set_row_precision): the tiled data forward and adjoint with FP16
rows equal, to FP32 arithmetic, the FP64 operator built from the same
coefficients rounded to FP16 (numpy's round-to-nearest-even, as the
device's conversion), for the 8-slot min-max rows and the 16-slot
w-projected rows; they stay within the FP16 rounding of the FP32-row
operator; repeats are bit-identical and going back to 32 bits restores
the FP32 results exactly; a solve with FP16 rows ends within 0.05 dB of
the FP32-row solve; untiled G and other bit counts are refused."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu
FOV, S = 0.08, 4


def _sparse(a, cy, cx, Kty, Ktx):
    M, Ky = cy.shape
    Kx = cx.shape[1]
    ky = (a % Kty)[:, None] + np.arange(Ky)[None, :]
    kx = (a // Kty)[:, None] + np.arange(Kx)[None, :]
    cols = ((ky % Kty)[:, :, None] + Kty * (kx % Ktx)[:, None, :]).reshape(M, -1)
    vals = (cy.astype(np.complex128)[:, :, None] * cx.astype(np.complex128)[:, None, :]).reshape(M, -1)
    return sp.csr_matrix((vals.ravel(), (np.repeat(np.arange(M), Ky * Kx), cols.ravel())), shape=(M, Kty * Ktx))


def _f16(c):
    """complex64 coefficients with real and imaginary parts rounded to FP16."""
    return (c.real.astype(np.float16).astype(np.float64) + 1j * c.imag.astype(np.float16).astype(np.float64))


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize('rows', ['minmax', 'wproj'])
def test_fp16_rows_match_rounded_operator(ctx, rows):
    from fhs import solvers
    from fhs.core import GMatrix
    from fhs.data_fidelity import DataOperator
    from fhs.nufft import Nufft2D
    from oracle.nufft import NufftOracle
    N, L = 64, 2
    nf = Nufft2D(N, N)
    plan = nf.plan(ctx)
    G32, G16, Gd, blocks = [], [], [], []
    for l in range(L):
        rng = np.random.default_rng(40 + l)
        M = 3000 + 700 * l
        u, v = rng.uniform(-np.pi, np.pi, M), rng.uniform(-np.pi, np.pi, M)
        if rows == 'wproj':
            a, cy, cx = nf.separable_wproj_gpu(ctx, u, v, rng.uniform(0, 200, M), FOV, S=S)
        else:
            a, cy, cx = nf.separable_gpu(ctx, u, v)
        Gd.append(GMatrix.from_separable(ctx, a, cy, cx, nf.Kty, nf.Ktx))
        G32.append(_sparse(a, cy, cx, nf.Kty, nf.Ktx))
        G16.append(_sparse(a, _f16(cy), _f16(cx), nf.Kty, nf.Ktx))
        blocks.append([M // 2, M - M // 2])
    D = DataOperator(ctx, plan, Gd, blocks)
    lib = solvers.lib()
    op = NufftOracle(np.zeros(1), np.zeros(1), N, N)
    x = torch.rand((L, N, N), dtype=torch.float32, device='cuda')
    xs = x.cpu().numpy().astype(np.float64)
    yv = (np.random.default_rng(7).standard_normal(D.Mtot)
          + 1j * np.random.default_rng(8).standard_normal(D.Mtot)).astype(np.complex64)
    yd = torch.from_numpy(yv).cuda()

    def fwd():
        out = torch.empty(D.Mtot, dtype=torch.complex64, device='cuda')
        ctx.check(lib.fhs_data_forward(ctx.h, D.h, x.data_ptr(), None, out.data_ptr()), 'data_forward')
        return out.cpu().numpy()

    def adj():
        back = torch.empty((L, N, N), dtype=torch.float32, device='cuda')
        ctx.check(lib.fhs_data_adjoint(ctx.h, D.h, yd.data_ptr(), None, None, back.data_ptr()), 'data_adjoint')
        return back.cpu().numpy()

    f32, a32 = fwd(), adj()
    D.set_row_precision(16)
    f16, a16 = fwd(), adj()
    assert np.array_equal(f16, fwd()) and np.array_equal(a16, adj())   # deterministic
    off = np.concatenate([[0], np.cumsum([g.shape[0] for g in G32])])
    y64 = yv.astype(np.complex128)
    for l in range(L):
        Fx = op.A(xs[l].T)
        # FP16 rows: the FP64 operator of the rounded coefficients, to FP32 arithmetic
        assert _rel(f16[off[l]:off[l + 1]], G16[l] @ Fx) < 1e-5
        r16 = np.real(op.At(G16[l].conj().T @ y64[off[l]:off[l + 1]]))
        assert _rel(a16[l].T, r16) < 1e-5
        # and within the FP16 rounding of the unrounded operator
        assert _rel(f16[off[l]:off[l + 1]], G32[l] @ Fx) < 2e-3
        r32 = np.real(op.At(G32[l].conj().T @ y64[off[l]:off[l + 1]]))
        assert _rel(a16[l].T, r32) < 2e-3
    D.set_row_precision(32)
    assert np.array_equal(fwd(), f32) and np.array_equal(adj(), a32)


def test_fp16_rows_solve(ctx):
    """HyperSARA on FP16 rows against the same solve on FP32 rows."""
    from fhs import solvers, synth
    from oracle.solvers import snr
    prob = synth.make_problem(64, 64, 4, 4000, 2, seed=29, n_sources=30)
    meas = solvers.Measurement(ctx, prob, precond=True)
    son, nu2 = meas.op_norm(False), meas.op_norm(True)
    p = {'squared_operator_norm': son, 'nu2': nu2}
    x32, _ = solvers.hyper_sara(meas, p, max_total_iter=80, rows=32)
    x16, _ = solvers.hyper_sara(meas, p, max_total_iter=80, rows=16)
    assert meas.D.row_bits == 16
    x32b, _ = solvers.hyper_sara(meas, p, max_total_iter=80, rows=32)
    assert np.array_equal(x32, x32b)
    assert _rel(x16, x32) < 5e-3
    assert abs(snr(prob.X0, x16) - snr(prob.X0, x32)) < 0.05


def test_fp16_rows_refused(ctx):
    import fhs
    from fhs.core import GMatrix
    from fhs.data_fidelity import DataOperator
    from fhs.nufft import Nufft2D
    from oracle.nufft import NufftOracle
    N = 32
    nf = Nufft2D(N, N)
    plan = nf.plan(ctx)
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-np.pi, np.pi, 500), rng.uniform(-np.pi, np.pi, 500)
    op = NufftOracle(u, v, N, N)
    D = DataOperator(ctx, plan, [GMatrix.from_scipy(ctx, op.G)], [[500]])
    with pytest.raises(fhs.FhsError):
        D.set_row_precision(16)    # CSR G: no FP16 rows
    with pytest.raises(fhs.FhsError):
        D.set_row_precision(8)
    D.set_row_precision(32)
    # a tiled separable channel next to a CSR one: refused, and the state unchanged
    from fhs import solvers
    a, cy, cx = nf.separable_gpu(ctx, u, v)
    D2 = DataOperator(ctx, plan, [GMatrix.from_separable(ctx, a, cy, cx, nf.Kty, nf.Ktx),
                                  GMatrix.from_scipy(ctx, op.G)], [[500], [500]])
    x = torch.rand((2, N, N), dtype=torch.float32, device='cuda')

    def fwd():
        out = torch.empty(D2.Mtot, dtype=torch.complex64, device='cuda')
        ctx.check(solvers.lib().fhs_data_forward(ctx.h, D2.h, x.data_ptr(), None, out.data_ptr()), 'data_forward')
        return out.cpu().numpy()
    before = fwd()
    with pytest.raises(fhs.FhsError):
        D2.set_row_precision(16)
    assert D2.row_bits == 32 and np.array_equal(fwd(), before)
