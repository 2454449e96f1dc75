"""GPU parity of the data-term operators against the FP64 oracle.

Tolerance: relative error in the 2-norm below 1e-5 (FP32 operators vs FP64
oracle on the same float32-rounded inputs)."""
import numpy as np
import pytest
import torch

from oracle.nufft import NufftOracle
from oracle.prox import proj_l2ball as proj_ref, update_primal as primal_ref

TOL = 1e-5
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def uv(rng, M, sigma=np.pi / 4):
    u = rng.normal(0, sigma, 3 * M); v = rng.normal(0, sigma, 3 * M)
    keep = (np.abs(u) < np.pi) & (np.abs(v) < np.pi)
    return u[keep][:M], v[keep][:M]


def crand(rng, n):
    return (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.cuda() if dtype is None else t.to(dtype).cuda()


@pytest.mark.parametrize('Ny,Nx,M', [(16, 16, 200), (32, 32, 500), (64, 48, 3000), (256, 256, 40000)])
@pytest.mark.parametrize('layout', ['csr', 'window', 'separable'])
def test_direct_and_adjoint(ctx, Ny, Nx, M, layout):
    import fhs
    rng = np.random.default_rng(Ny * 7 + M)
    u, v = uv(rng, M)
    op = NufftOracle(u, v, Ny, Nx)
    if layout == 'csr':
        G = fhs.GMatrix.from_scipy(ctx, op.G)
    elif layout == 'window':
        G = fhs.GMatrix.from_window(ctx, op.anchor, op.coef, op.Kty, op.Ktx, op.Ky, op.Kx)
    else:
        G = fhs.GMatrix.from_separable(ctx, op.anchor, op.cy_sep, op.cx_sep, op.Kty, op.Ktx)
    Gr = op.G.astype(np.complex64).astype(np.complex128)   # the matrix the GPU holds (+- FP32 products)
    Fx = crand(rng, op.No)
    y = fhs.apply_direct_operator(to_dev(Fx), G).cpu().numpy()
    assert rel(y, Gr @ Fx.astype(np.complex128)) < TOL
    lam = rng.uniform(0.5, 2.0, M).astype(np.float32)
    y2 = fhs.apply_direct_operator(to_dev(Fx), G, to_dev(lam)).cpu().numpy()
    assert rel(y2, lam * (Gr @ Fx.astype(np.complex128))) < TOL
    yy = crand(rng, M)
    g = fhs.apply_adjoint_operator(to_dev(yy), G).cpu().numpy()
    assert rel(g, Gr.conj().T @ yy.astype(np.complex128)) < TOL
    g2 = fhs.apply_adjoint_operator(to_dev(yy), G, to_dev(lam)).cpu().numpy()
    assert rel(g2, Gr.conj().T @ (lam * yy.astype(np.complex128))) < TOL
    # every output has a single owner and a fixed summation order: repeat calls are bit-identical
    g3 = fhs.apply_adjoint_operator(to_dev(yy), G, to_dev(lam)).cpu().numpy()
    assert np.array_equal(g2, g3)


# Ktx = 4096 with Nx <= 2048 takes the split x-passes (two 2048-point
# transforms); (16, 4096) without x oversampling takes the full 4096-point
# ones.  Grids that are not powers of two (2^a 3^b 5^c: 200 = 8 25,
# 120 = 8 3 5, 192 = 64 3, 160 = 32 5, 5120 = 1024 5 -- the oversampled
# 2560-pixel axis of the reference's real-data default) take the runtime
# mixed-radix path on the axis concerned.
@pytest.mark.parametrize('Ny,Nx,ox', [(8, 8, 2), (16, 32, 2), (64, 64, 2), (100, 60, 2), (96, 80, 2), (256, 256, 2),
                                      (512, 512, 2), (1024, 512, 2), (512, 1024, 2), (256, 2048, 2), (64, 2048, 2),
                                      (2048, 64, 2), (16, 4096, 1), (2560, 48, 2), (48, 2560, 2)])
def test_scaled_fourier(ctx, Ny, Nx, ox):
    import fhs
    rng = np.random.default_rng(Ny + Nx)
    op = NufftOracle(np.zeros(1), np.zeros(1), Ny, Nx, ox=ox)
    assert (op.Kty, op.Ktx) == (2 * Ny, ox * Nx)
    plan = fhs.NufftPlan(ctx, Ny, Nx, op.Kty, op.Ktx, op.sy, op.sx)
    L = 3
    x = rng.random((L, Nx, Ny)).astype(np.float32)
    Fx = plan.A(to_dev(x)).cpu().numpy()
    for l in range(L):
        ref = op.A(x[l].T.astype(np.float64))      # oracle takes (Ny, Nx)
        assert rel(Fx[l], ref) < TOL
    g = np.stack([crand(rng, op.No) for _ in range(L)])
    scale = [1.0, 0.5, 2.0]
    out = plan.At_real(to_dev(g), scale).cpu().numpy()
    for l in range(L):
        ref = scale[l] * np.real(op.At(g[l].astype(np.complex128)))
        assert rel(out[l].T, ref) < TOL


@pytest.mark.parametrize('n,inside', [(1000, False), (1 << 20, False), (5000, True), (1, False)])
def test_proj_l2ball(ctx, n, inside):
    import fhs
    rng = np.random.default_rng(n)
    x = crand(rng, n); y = crand(rng, n)
    d = np.linalg.norm((x - y).astype(np.complex128))
    eps = 2 * d if inside else 0.3 * d
    p = fhs.proj_l2ball(to_dev(x), eps, to_dev(y)).cpu().numpy()
    ref = proj_ref(x.astype(np.complex128), eps, y.astype(np.complex128))
    assert rel(p, ref) < TOL
    if inside:
        assert rel(p, x) < 1e-7


@pytest.mark.parametrize('n,parts', [(4096, 1), (1 << 20, 3), (12345, 2)])
def test_update_primal(ctx, n, parts):
    import fhs
    rng = np.random.default_rng(n)
    x = rng.normal(size=n).astype(np.float32)
    gs = [rng.normal(size=n).astype(np.float32) for _ in range(parts)]
    xd = to_dev(x)
    xd, xhat, rel_x, norm_x = fhs.fhs_update_primal(xd, *[to_dev(g) for g in gs])
    xr, xhr, rr, nr = primal_ref(x.astype(np.float64), sum(g.astype(np.float64) for g in gs))
    assert rel(xd.cpu().numpy(), xr) < TOL
    assert rel(xhat.cpu().numpy(), xhr) < TOL
    assert abs(rel_x - rr) / rr < 1e-6 and abs(norm_x - nr) / nr < 1e-6


@pytest.mark.parametrize('Ny,Nx', [(64, 64), (48, 80), (256, 256)])
def test_interp_coefficients_gpu(ctx, Ny, Nx):
    """GPU min-max coefficients (FP64 on the device) against the oracle's
    restatement and the host builder: identical anchors, coefficients equal
    to complex64 rounding, including on-grid and band-edge frequencies."""
    from fhs.nufft import Nufft2D
    rng = np.random.default_rng(Ny + Nx)
    nf = Nufft2D(Ny, Nx)
    u = rng.uniform(-np.pi, np.pi, 5000)
    v = rng.uniform(-np.pi, np.pi, 5000)
    # on-grid frequencies (exact Dirichlet zeros) and the band edges
    u[:8] = 2 * np.pi * np.arange(-4, 4) / nf.Ktx
    v[:8] = 2 * np.pi * np.arange(-4, 4) / nf.Kty
    u[8:12] = [-np.pi, np.pi - 1e-12, 0.0, -1e-15]
    v[8:12] = [np.pi - 1e-12, -np.pi, -1e-15, 0.0]
    scale = rng.uniform(0.5, 2.0, 5000)
    a, cy, cx = nf.separable_gpu(ctx, u, v, row_scale=scale)
    ah, cyh, cxh = nf.separable(u, v, row_scale=scale)
    assert np.array_equal(a, ah)
    for g, h in ((cy, cyh), (cx, cxh)):
        err = np.abs(g.astype(np.complex128) - h)
        assert err.max() <= 2e-7 * np.abs(h).max()
    # the product of the GPU factors against the oracle's G rows (unscaled)
    orc = NufftOracle(u[:200], v[:200], Ny, Nx)
    a0, cy0, cx0 = nf.separable_gpu(ctx, u[:200], v[:200])
    Kty = nf.Kty
    for m in range(0, 200, 37):
        row = np.zeros(nf.No, dtype=np.complex128)
        ky0, kx0 = a0[m] % Kty, a0[m] // Kty
        for jx in range(nf.Kx):
            for jy in range(nf.Ky):
                row[(ky0 + jy) % Kty + Kty * ((kx0 + jx) % nf.Ktx)] += cy0[m, jy] * cx0[m, jx]
        ref = orc.G.getrow(m).toarray().ravel()
        assert np.abs(row - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize('N,M', [(64, 5000), (512, 200000)])
def test_uniform_weights_gpu_match_oracle(ctx, N, M):
    """fhs_precond_uniform_weights against the oracle's count (exact: 1/count
    of integers) and the product's host rule, with points stacked in a few
    cells and on the grid's edge so that counts above one occur."""
    from fhs.nufft import uniform_weights, uniform_weights_gpu
    from oracle.nufft import uniform_precond_weights
    rng = np.random.default_rng(N)
    u = rng.normal(0, np.pi / 4, M).clip(-np.pi, np.pi)
    v = rng.normal(0, np.pi / 4, M).clip(-np.pi, np.pi)
    u[:50], v[:50] = 0.1, -0.2              # 50 points in one cell
    u[50:60], v[50:60] = np.pi, -np.pi      # the edge wraps to cell 0
    K = 2 * N
    got = uniform_weights_gpu(ctx, u, v, K, K).cpu().numpy()
    ref = uniform_precond_weights(u, v, K, K)
    assert np.array_equal(got, ref.astype(np.float32))
    assert np.array_equal(got, uniform_weights(u, v, K, K).astype(np.float32))
    assert round(1 / float(got[0])) >= 50 and np.all(got[:50] == got[0])    # the stacked cell
