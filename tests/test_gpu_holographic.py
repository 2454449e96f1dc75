"""GPU parity of the holographic (dimensionality-reduced) data builder,
lib/operators/util_load_data_gen_dr_measurement_operator.m:160-225, against
the FP64 oracle, and of the DR dual step (src/update_dual_data_fidelity.m:
85-125) on the data it builds.  Tolerance 1e-5 relative (FP32 storage, FP64
sums)."""
import time

import numpy as np
import pytest
import torch

from oracle.data_fidelity import holographic_operator
from oracle.data_fidelity import update_dual_data_fidelity as dual_ref
from oracle.nufft import NufftOracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b, scale=0.0):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), scale, 1e-300)


def uv(rng, M, scale=1.0):
    return (np.clip(rng.normal(0, np.pi / 4, M), -3.1, 3.1) * scale,
            np.clip(rng.normal(0, np.pi / 4, M), -3.1, 3.1) * scale)


def cplx(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).cuda()


# 16 x 16 has a sparse, ragged coverage of its 32 x 32 grid (inactive modes,
# rows with few entries); 64 x 64 with 30,000 visibilities covers most of it
@pytest.mark.parametrize('Ny,Nx,M,K', [(16, 16, 300, (7, 7)), (32, 64, 4000, (7, 7)), (64, 64, 30000, (7, 7)),
                                       (32, 32, 3000, (5, 5)), (32, 32, 3000, (6, 4))])
def test_holographic_matches_oracle(ctx, Ny, Nx, M, K):
    import fhs
    from fhs.holographic import Holographic
    rng = np.random.default_rng(M)
    u, v = uv(rng, M)
    op = NufftOracle(u, v, Ny, Nx, Ky=K[0], Kx=K[1])
    G = fhs.GMatrix.from_separable(ctx, op.anchor, op.cy_sep, op.cx_sep, op.Kty, op.Ktx)
    Gs = op.G.tocsr()
    y, n = cplx(rng, M), cplx(rng, M) / np.sqrt(2)
    H, W, lam, yg, l2, _ = holographic_operator(Gs, y, n)
    hf = Holographic(ctx, G)
    No = op.No
    off = np.setdiff1d(np.arange(No), W)
    assert hf.n_modes == len(W) and np.array_equal(hf.W, W)
    if Ny == 16:
        assert len(off) > 0
    lam_d = hf.Lambda.cpu().numpy()
    assert rel(lam_d[W], lam) < TOL and np.all(lam_d[off] == 0)
    assert hf.H.nnz == H.nnz
    # H (ELL rows) and H' (CSC columns) against the oracle's on random vectors
    f = np.zeros(No, complex)
    f[W] = cplx(rng, len(W))
    Hf = fhs.apply_direct_operator(to_dev(f), hf.H).cpu().numpy()
    Htf = fhs.apply_adjoint_operator(to_dev(f), hf.H).cpu().numpy()
    assert rel(Hf[W], H @ f[W]) < TOL and np.all(Hf[off] == 0)
    assert rel(Htf[W], H.conj().T @ f[W]) < TOL and np.all(Htf[off] == 0)
    # gridded data and the noise l2 bound
    ygd = hf.grid(to_dev(y)).cpu().numpy()
    assert rel(ygd[W], yg) < TOL and np.all(ygd[off] == 0)
    gn = hf.grid(to_dev(n)).cpu().numpy()
    assert abs(np.linalg.norm(gn) - l2) < TOL * l2
    assert np.array_equal(hf.compact(torch.from_numpy(ygd)).numpy(), ygd[W])


def test_holographic_build_is_deterministic(ctx):
    import fhs
    from fhs.holographic import Holographic
    rng = np.random.default_rng(5)
    u, v = uv(rng, 20000)
    op = NufftOracle(u, v, 64, 64)
    G = fhs.GMatrix.from_separable(ctx, op.anchor, op.cy_sep, op.cx_sep, op.Kty, op.Ktx)
    f = to_dev(cplx(rng, op.No))
    a, b = Holographic(ctx, G), Holographic(ctx, G)
    assert torch.equal(a.Lambda, b.Lambda)
    assert torch.equal(fhs.apply_direct_operator(f, a.H), fhs.apply_direct_operator(f, b.H))
    assert torch.equal(fhs.apply_adjoint_operator(f, a.H), fhs.apply_adjoint_operator(f, b.H))


@pytest.mark.parametrize('L,K', [(1, (7, 7)), (2, (7, 7)), (2, (6, 4))])
def test_dr_dual_step_on_built_data(ctx, L, K):
    """update_dual_data_fidelity's DR branch on the builder's H, Lambda and
    gridded data against the oracle's DR branch on the oracle's (compact |W|
    vectors), over consecutive iterations; eps from the gridded noise."""
    import fhs
    from fhs.holographic import Holographic
    Ny = Nx = 32
    M = 3000
    rng = np.random.default_rng(40 + L)
    ops, forms, Hs, Ws, lams, ygs, eps, yds = [], [], [], [], [], [], [], []
    for l in range(L):
        u, v = uv(rng, M, 0.8 + 0.1 * l)
        op = NufftOracle(u, v, Ny, Nx, Ky=K[0], Kx=K[1])
        x0 = rng.random((Ny, Nx))
        y = op.G @ op.A(x0) + 0.05 * cplx(rng, M)
        n = cplx(rng, M) / np.sqrt(2)
        H, W, lam, yg, l2, _ = holographic_operator(op.G.tocsr(), y, n)
        G = fhs.GMatrix.from_separable(ctx, op.anchor, op.cy_sep, op.cx_sep, op.Kty, op.Ktx)
        hf = Holographic(ctx, G)
        assert np.array_equal(hf.W, W)
        ops.append(op)
        forms.append(hf)
        Hs.append(H)
        Ws.append(W)
        lams.append(lam)
        ygs.append(hf.grid(to_dev(y)).cpu().numpy()[W].astype(np.complex128))
        yds.append(hf.grid(to_dev(y)))
        eps.append((1.0 + 0.5 * l) * l2)
    No = ops[0].No
    plan = fhs.NufftPlan(ctx, Ny, Nx, ops[0].Kty, ops[0].Ktx, ops[0].sy, ops[0].sx)
    D = fhs.DataOperator(ctx, plan, [f.H for f in forms], [[No]] * L)
    lamd = torch.cat([f.Lambda for f in forms])
    yd = torch.cat(yds)
    v2 = [[np.zeros(len(Ws[l]), complex)] for l in range(L)]
    proj = [[np.zeros(len(Ws[l]), complex)] for l in range(L)]
    x = np.zeros((Ny, Nx, L))
    Fx_old = np.zeros((No, L), complex)
    sigma22 = np.array([0.3 + 0.1 * l for l in range(L)], dtype=np.float32)
    v2d, projd = D.zeros(), D.zeros()
    gx_old = D.init_fx_old(torch.zeros((L, Nx, Ny), dtype=torch.float32, device='cuda'), lamd, True)
    yw = [[ygs[l]] for l in range(L)]
    for it in range(3):
        x = np.maximum(x + rng.normal(size=x.shape) * 0.5, 0).astype(np.float32).astype(np.float64)
        xd = torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0)).astype(np.float32)).cuda()
        v2, Ftx, Fx_old, proj, nr, gnr, neps = dual_ref(
            v2, yw, x, Fx_old, proj, [op.A for op in ops], [op.At for op in ops], [[H] for H in Hs],
            [[W] for W in Ws], None, [[e] for e in eps], 20, 1, 1e-8, sigma22.astype(np.float64),
            flag_dimensionality_reduction=True, Lambda=[[lm] for lm in lams])
        v2d, Ftxd, gx_old, projd, nrd, gnrd, nepsd = fhs.update_dual_data_fidelity(
            v2d, yd, xd, gx_old, projd, plan, plan, D, None, None, np.asarray(eps), 20, 1, 1e-8, sigma22,
            flag_dimensionality_reduction=True, Lambda=lamd)
        v2h = v2d.cpu().numpy().reshape(L, No)
        for l in range(L):
            off = np.setdiff1d(np.arange(No), Ws[l])
            assert rel(v2h[l][Ws[l]], v2[l][0], np.linalg.norm(ygs[l])) < TOL, (it, l)
            assert np.all(v2h[l][off] == 0)
        assert rel(nrd, [nr[l][0] for l in range(L)]) < TOL, it
        assert rel(Ftxd.cpu().numpy().transpose(2, 1, 0), Ftx) < TOL, it
        assert abs(gnrd - gnr) / gnr < TOL and abs(nepsd - neps) / neps < 1e-12


def test_holographic_full_size_is_gram_of_g(ctx):
    """512 x 512 image (1024 x 1024 grid), 160,000 visibilities: on random
    grid vectors Lambda (H + H') Lambda equals Lambda G'G Lambda on the active
    modes, the active modes are the grid points under some window, and
    Lambda^-2 is the column energy of G (size-independent properties)."""
    import fhs
    from fhs.holographic import Holographic
    from fhs.nufft import Nufft2D
    Ny = Nx = 512
    M = 160000
    rng = np.random.default_rng(2)
    u, v = uv(rng, M)
    nf = Nufft2D(Ny, Nx)
    a, cy, cx = nf.separable_gpu(ctx, u, v)
    G = fhs.GMatrix.from_separable(ctx, a, cy, cx, nf.Kty, nf.Ktx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hf = Holographic(ctx, G)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    No = nf.Kty * nf.Ktx
    print(f'holographic build 512^2 image, {M} visibilities: {dt * 1e3:.1f} ms, |W| = {hf.n_modes} '
          f'of {No}, nnz(H) = {hf.H.nnz}')
    Ky, Kx = cy.shape[1], cx.shape[1]
    ay, ax = a % nf.Kty, a // nf.Kty
    cover = np.zeros(No, bool)
    for jx in range(Kx):
        for jy in range(Ky):
            cover[(ay + jy) % nf.Kty + nf.Kty * ((ax + jx) % nf.Ktx)] = True
    W = hf.W
    assert np.array_equal(W, np.flatnonzero(cover))
    lam = hf.Lambda
    # diagonal at a few active modes p: <e_p, (H + H') e_p> = (G'G)_pp = Lambda_p^-2
    for p in W[:: max(1, len(W) // 7)][:8]:
        e = torch.zeros(No, dtype=torch.complex64, device='cuda')
        e[int(p)] = 1
        d = (fhs.apply_direct_operator(e, hf.H) + fhs.apply_adjoint_operator(e, hf.H))[int(p)].real.item()
        gp = fhs.apply_adjoint_operator(fhs.apply_direct_operator(e, G), G)[int(p)].real.item()
        assert abs(d - gp) < 1e-5 * gp and abs(lam[int(p)].item() ** -2 - gp) < 1e-5 * gp
    for s in range(2):
        f = torch.zeros(No, dtype=torch.complex64, device='cuda')
        Wt = torch.from_numpy(W).cuda()
        f[Wt] = to_dev(cplx(np.random.default_rng(10 + s), len(W))) * lam[Wt]
        lhs = (fhs.apply_direct_operator(f, hf.H) + fhs.apply_adjoint_operator(f, hf.H)) * lam
        rhs = fhs.apply_adjoint_operator(fhs.apply_direct_operator(f, G), G) * lam
        assert rel(lhs.cpu().numpy(), rhs.cpu().numpy()) < TOL


def _oracle_problem(prob, meas):
    """The oracle problem on the measurement's own (block-sorted) row order."""
    from oracle.solvers import OracleProblem
    blocks = []
    for l in range(prob.L):
        b = np.concatenate([[0], np.cumsum(prob.blocks[l])])
        blocks.append([(int(b[i]), int(b[i + 1])) for i in range(len(prob.blocks[l]))])
    pU = [np.ones(len(u)) for u in meas.u]
    y = [np.asarray(yl).astype(np.complex64).astype(np.complex128) for yl in meas.y_host]
    return OracleProblem(prob.Ny, prob.Nx, meas.u, meas.v, y, pU, blocks, prob.epsilon,
                         g_scale=[1.0 / s for s in prob.sigma_noise])


def _dataset_ranges(prob):
    """Row ranges of every channel's data sets (rows are set after set)."""
    out = []
    for d in prob.datasets:
        b = np.concatenate([[0], np.cumsum(d)])
        out.append([(int(b[i]), int(b[i + 1])) for i in range(len(d))])
    return out


# n_datasets = 3: each channel's DR data sum three independent coverages, as
# the reference sums its data sets (util_load_data_gen_dr_measurement_operator.m:
# 112-196); the oracle sums per set, the GPU builds from the stacked rows
@pytest.mark.parametrize('n_datasets', [1, 3])
def test_dr_operator_and_adjoint(ctx, n_datasets):
    """The DR measurement operator Lambda (H + H') A (HS_forward_operator_G.m,
    istril branch) against the oracle's on the oracle's H, and its adjoint by
    the identity Re<Phi x, y> = <x, Phi' y>; the gridded data and the l2
    bounds of the gridded noise."""
    from fhs import solvers, synth
    from fhs.holographic import HolographicMeasurement
    prob = synth.make_problem(32, 32, 3, 1500 // n_datasets, 2, seed=4, n_sources=20, n_datasets=n_datasets)
    meas = solvers.Measurement(ctx, prob, precond=False)
    dm = HolographicMeasurement(meas, noise_seed=2)
    P = _oracle_problem(prob, meas).holographic(dm.noise_host, datasets=_dataset_ranges(prob))
    assert all(len(d) == n_datasets for d in prob.datasets)
    rng = np.random.default_rng(0)
    x = rng.random((prob.Ny, prob.Nx, prob.L))
    xd = torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0)).astype(np.float32)).cuda()
    yd = dm.forward(xd).cpu().numpy().reshape(prob.L, -1)
    for l in range(prob.L):
        H, W, lam = P.G[l][0], P.W[l][0], P.Lambda[l][0]
        f = P.A[l](x[:, :, l])[W]
        assert rel(yd[l][W], lam * (H @ f + H.conj().T @ f)) < TOL, l
        assert np.all(yd[l][np.setdiff1d(np.arange(P.No), W)] == 0)
        assert abs(dm.epsilon[l] - P.epsilon[l][0]) < 1e-5 * P.epsilon[l][0]
        yg = dm.y[l * P.No:(l + 1) * P.No].cpu().numpy()
        assert rel(yg[W], P.y[l][0]) < TOL, l
    z = np.zeros_like(yd)
    for l in range(prob.L):
        z[l][P.W[l][0]] = cplx(rng, len(P.W[l][0]))
    zd = to_dev(z.ravel())
    lhs = np.real(np.vdot(zd.cpu().numpy().astype(np.complex128), dm.forward(xd).cpu().numpy()))
    rhs = float(torch.sum(xd.double() * dm.adjoint(zd).double()).item())
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


@pytest.mark.parametrize('n_datasets', [1, 2])
def test_dr_hypersara_matches_oracle(ctx, n_datasets):
    """HyperSARA on the DR form (imaging/imaging.m:619-631: gridded data,
    Lambda, closed-form l2 dual) against the oracle's HyperSARA on its own
    DR problem, same operator norms: 1e-3 relative on the iterate, 0.1 dB;
    with two data sets per channel summed into the DR data."""
    from fhs import solvers, synth
    from fhs.holographic import HolographicMeasurement
    from oracle.solvers import hypersara as hs_ref, snr
    prob = synth.make_problem(32, 32, 4, 1500 // n_datasets, 2, seed=3, n_sources=20, n_datasets=n_datasets)
    meas = solvers.Measurement(ctx, prob, precond=False)
    dm = HolographicMeasurement(meas, noise_seed=1)
    son = dm.op_norm()
    K = 40
    x_gpu, info = solvers.hyper_sara(dm, {'squared_operator_norm': son, 'nu2': son}, X0=prob.X0,
                                     max_total_iter=K)
    P = _oracle_problem(prob, meas).holographic(dm.noise_host, datasets=_dataset_ranges(prob))
    sig, sig_bar = solvers.noise_levels_hs(32, 32, 4, dm.noise_std, son)
    pr = dict(solvers.DEFAULT_PARAM)
    pr.update({'nu2': son, 'gamma': sig, 'gamma0': sig_bar, 'sig': sig, 'sig_bar': sig_bar, 'alpha': 1.0})
    x_ref, rinfo = hs_ref(P, pr, max_iter_total=K)
    r = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    print(f'DR HyperSARA 32x32x4, {K} iterations: rel {r:.2e}, SNR gpu {snr(prob.X0, x_gpu):.3f} '
          f'oracle {snr(prob.X0, x_ref):.3f}')
    assert r < 1e-3, r
    assert abs(snr(prob.X0, x_gpu) - snr(prob.X0, x_ref)) < 0.1


def test_dr_dual_step_full_size(ctx):
    """512 x 512 x 2 with 160,000 visibilities per channel: one DR dual step
    from the zero state (grid-sized blocks of 1,048,576 rows, the dual split
    over 256 workgroups each) against its closed form evaluated in FP64 from
    the DR operator's own output: v2 = sc(2 gx - y, eps), ||gx - y||."""
    import fhs
    from fhs import solvers, synth
    from fhs.holographic import HolographicMeasurement
    prob = synth.make_problem(512, 512, 2, 160000, 1, seed=6)
    meas = solvers.Measurement(ctx, prob, precond=False)
    dm = HolographicMeasurement(meas, noise_seed=3)
    L, No = 2, dm.forms[0].No
    x = torch.from_numpy(np.random.default_rng(1).random((L, 512, 512)).astype(np.float32)).cuda()
    gx = dm.forward(x).to(torch.complex128)
    y = dm.y.to(torch.complex128)
    eps = np.array([0.5 * torch.linalg.vector_norm(2 * gx[l * No:(l + 1) * No] - y[l * No:(l + 1) * No]).item()
                    for l in range(L)])
    gx_old = dm.D.init_fx_old(torch.zeros_like(x), dm.Lambda, True)
    assert torch.count_nonzero(gx_old).item() == 0
    v2d, _, gx_old, _, nrd, _, _ = fhs.update_dual_data_fidelity(
        dm.D.zeros(), dm.y, x, gx_old, dm.D.zeros(), dm.plan, dm.plan, dm.D, None, None, eps, 20, 1, 1e-8,
        np.full(L, 0.5, np.float32), flag_dimensionality_reduction=True, Lambda=dm.Lambda)
    assert torch.equal(gx_old, dm.forward(x))
    for l in range(L):
        s = slice(l * No, (l + 1) * No)
        z = 2 * gx[s] - y[s]
        nz = torch.linalg.vector_norm(z).item()
        want = z * (1 - min(eps[l] / nz, 1))
        assert rel(v2d[s].cpu().numpy(), want.cpu().numpy()) < TOL, l
        assert abs(nrd[l] - torch.linalg.vector_norm(gx[s] - y[s]).item()) < TOL * nrd[l], l
