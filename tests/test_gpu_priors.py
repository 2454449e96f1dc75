"""GPU parity of the prior kernels (SARA dictionary, l21 step, nuclear-norm
step, faceted overlap-add) against the FP64 oracle.  Tolerance: 1e-5
relative 2-norm error (FP32 kernels vs FP64 oracle on the same inputs)."""
import numpy as np
import pytest
import torch

from oracle.facets import facet_geometry, faceted_lowrank_step, facet_singular_values
from oracle.prox import nuclear_prox_dual, soft_l21
from oracle.wavelets import SaraDictionary as SaraRef

TOL = 1e-5
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def cube_to_dev(x):            # (Ny, Nx, L) -> (L, Nx, Ny) float32 on GPU
    return torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0)).astype(np.float32)).cuda()


def cube_from_dev(t):
    return t.cpu().numpy().astype(np.float64).transpose(2, 1, 0)


@pytest.mark.parametrize('Ny,Nx,L', [(8, 8, 1), (37, 29, 3), (64, 64, 2), (128, 96, 4), (512, 512, 2)])
def test_psit_psi(ctx, Ny, Nx, L):
    import fhs
    rng = np.random.default_rng(Ny * Nx + L)
    S = fhs.SaraDictionary(ctx, Ny, Nx)
    ref = SaraRef(Ny, Nx)
    assert S.s == ref.s
    x = rng.random((Ny, Nx, L))
    c = S.psit(cube_to_dev(x)).cpu().numpy()
    for l in range(L):
        assert rel(c[l], ref.psit(x[:, :, l].astype(np.float32))) < TOL
    cc = rng.normal(size=(L, ref.s)).astype(np.float32)
    img = cube_from_dev(S.psi(torch.from_numpy(cc).cuda(), scale=0.7))
    for l in range(L):
        assert rel(img[:, :, l], 0.7 * ref.psi(cc[l].astype(np.float64))) < TOL


@pytest.mark.parametrize('bases', [('db1',), ('db3', 'self'), ('self',), ('db8',)])
def test_dictionary_subsets(ctx, bases):
    import fhs
    rng = np.random.default_rng(len(bases))
    Ny, Nx = 45, 70
    S = fhs.SaraDictionary(ctx, Ny, Nx, bases=bases, nlevel=3)
    ref = SaraRef(Ny, Nx, bases=bases, nlevel=3)
    x = rng.random((Ny, Nx, 1))
    c = S.psit(cube_to_dev(x)).cpu().numpy()[0]
    assert rel(c, ref.psit(x[:, :, 0].astype(np.float32))) < TOL
    y = cube_from_dev(S.psi(torch.from_numpy(c[None].copy()).cuda()))
    assert rel(y[:, :, 0], ref.psi(c.astype(np.float64))) < TOL


@pytest.mark.parametrize('Ny,Nx,L', [(64, 48, 3), (256, 256, 16)])
def test_dual_sparsity_step(ctx, Ny, Nx, L):
    import fhs
    rng = np.random.default_rng(L)
    S = fhs.SaraDictionary(ctx, Ny, Nx)
    ref = SaraRef(Ny, Nx)
    xhat = rng.normal(size=(Ny, Nx, L)).astype(np.float32).astype(np.float64)
    v1 = (0.1 * rng.normal(size=(ref.s, L))).astype(np.float32).astype(np.float64)
    w = rng.uniform(0.2, 1.0, ref.s).astype(np.float32).astype(np.float64)
    beta, sigma = 0.8, 0.33
    r1 = v1 + np.stack([ref.psit(xhat[:, :, l]) for l in range(L)], axis=1)
    v1_ref = soft_l21(r1, beta * w)
    g1_ref = np.stack([sigma * ref.psi(v1_ref[:, l]) for l in range(L)], axis=2)
    v1d = torch.from_numpy(np.ascontiguousarray(v1.T).astype(np.float32)).cuda()
    v1d, g1 = fhs.hs_update_dual_sparsity_distributed(
        v1d, S, S, cube_to_dev(xhat), torch.from_numpy(w.astype(np.float32)).cuda(), beta, sigma)
    assert rel(v1d.cpu().numpy().T, v1_ref) < TOL
    assert rel(cube_from_dev(g1), g1_ref) < TOL
    # the prior value
    f = fhs.compute_sara_prior(cube_to_dev(xhat), S, ref.s)
    r = np.stack([ref.psit(xhat[:, :, l]) for l in range(L)], axis=1)
    fr = np.sum(np.sqrt(np.sum(r ** 2, axis=1)))
    assert abs(f - fr) / fr < TOL


@pytest.mark.parametrize('Ny,Nx,L', [(64, 64, 16), (128, 64, 7), (32, 32, 1), (512, 512, 16)])
def test_nuclear_norm(ctx, Ny, Nx, L):
    import fhs
    rng = np.random.default_rng(L * 3)
    x = rng.random((Ny, Nx, L)).astype(np.float32).astype(np.float64)
    f = fhs.nuclear_norm(cube_to_dev(x))
    fr = np.sum(np.linalg.svd(x.reshape((-1, L), order='F'), compute_uv=False))
    assert abs(f - fr) / fr < TOL


@pytest.mark.parametrize('Ny,Nx,L', [(64, 64, 16), (100, 60, 5), (512, 512, 16)])
def test_global_lowrank_step(ctx, Ny, Nx, L):
    """hs_update_dual_lowrankness_serial, three consecutive steps (warm start)."""
    import fhs
    rng = np.random.default_rng(Ny + L)
    N = Ny * Nx
    R = fhs.LowRankPrior(ctx, Ny, Nx, L)
    v0 = np.zeros((N, L))
    v0d = R.zeros_v0()
    base = rng.random((Ny, Nx, 3)) @ rng.random((3, L))          # near low rank
    for it in range(3):
        xhat = (base + 0.05 * rng.normal(size=(Ny, Nx, L))).astype(np.float32).astype(np.float64)
        Z = v0 + xhat.reshape((N, L), order='F')
        s = np.linalg.svd(Z, compute_uv=False)
        w0 = np.linspace(1.0, 0.3, L)
        beta = float(np.median(s))
        v0 = nuclear_prox_dual(Z, beta * w0)
        v0d, g0 = R.update_dual(v0d, cube_to_dev(xhat), w0, beta, 0.33)
        got = v0d.cpu().numpy().reshape(L, N).T
        assert rel(got, v0) < TOL, it
        assert rel(cube_from_dev(g0), 0.33 * v0.reshape((Ny, Nx, L), order='F')) < TOL


# the last two are the faceted configurations' geometries: 1024x1024 in 4x4
# facets with 25% overlap (spectral sub-cube of 16 channels) and 2048x2048 in
# 6x6 facets with 50% overlap
@pytest.mark.parametrize('Ny,Nx,Qy,Qx,frac,L', [(64, 96, 2, 3, 0.25, 5), (120, 120, 3, 3, 0.5, 16),
                                                (256, 256, 4, 4, 0.25, 16), (1024, 1024, 4, 4, 0.25, 16),
                                                (2048, 2048, 6, 6, 0.5, 16)])
def test_faceted_lowrank_step(ctx, Ny, Nx, Qy, Qx, frac, L):
    import fhs
    rng = np.random.default_rng(Qy * Qx + L)
    geom = facet_geometry(Ny, Nx, Qy, Qx, (frac, frac))
    R = fhs.LowRankPrior(ctx, Ny, Nx, L, facets=geom)
    Q = Qy * Qx
    v0s = [np.zeros((geom['py'][q] * geom['px'][q], L)) for q in range(Q)]
    v0d = R.zeros_v0()
    for it in range(2):
        xhat = rng.random((Ny, Nx, L)).astype(np.float32).astype(np.float64)
        sv = facet_singular_values(xhat, geom)
        w0 = np.tile(np.linspace(1.0, 0.5, L), (Q, 1))
        beta = float(np.median(sv))
        v0s, g0 = faceted_lowrank_step(v0s, xhat, geom, w0, beta, 0.5)
        v0d, g0d = R.update_dual(v0d, cube_to_dev(xhat), w0, beta, 0.5)
        flat = v0d.cpu().numpy()
        ref = np.concatenate([v.T.ravel() for v in v0s])
        assert rel(flat, ref) < TOL, it
        assert rel(cube_from_dev(g0d), g0) < TOL, it
    svd_gpu = R.singular_values(cube_to_dev(xhat))
    assert rel(svd_gpu, facet_singular_values(xhat, geom)) < TOL


@pytest.mark.parametrize('mode,tol_sv,tol_prox', [('bf16x3', 1e-6, 1e-3), ('bf16x2', 1e-4, 3e-2)])
@pytest.mark.parametrize('Ny,Nx,Qy,Qx,L', [(256, 256, 1, 1, 16), (192, 160, 2, 2, 11)])
def test_mixed_precision_gram(ctx, mode, tol_sv, tol_prox, Ny, Nx, Qy, Qx, L):
    """The BF16 split Gram on the matrix cores against the FP64 oracle: the
    singular values and the low-rank prox of a near-low-rank cube, with the
    errors printed (the mixed form is judged on the final SNR of a solve; the
    FP64 Gram meets the 1e-5 operator tolerance)."""
    import fhs
    rng = np.random.default_rng(L)
    geom = facet_geometry(Ny, Nx, Qy, Qx, (0.25, 0.25))
    Q = Qy * Qx
    base = rng.random((Ny, Nx, 3)) @ rng.random((3, L))
    xhat = (base + 0.05 * rng.normal(size=(Ny, Nx, L))).astype(np.float32).astype(np.float64)
    sv_ref = facet_singular_values(xhat, geom)
    w0 = np.tile(np.linspace(1.0, 0.5, L), (Q, 1))
    beta = float(np.median(sv_ref))
    v0s = [np.zeros((geom['py'][q] * geom['px'][q], L)) for q in range(Q)]
    v0s, g0 = faceted_lowrank_step(v0s, xhat, geom, w0, beta, 0.5)
    ref = np.concatenate([v.T.ravel() for v in v0s])
    errs = {}
    for m in ('fp64', mode):
        R = fhs.LowRankPrior(ctx, Ny, Nx, L, facets=geom)
        R.set_gram_precision(m)
        sv = R.singular_values(cube_to_dev(xhat))
        v0d, g0d = R.update_dual(R.zeros_v0(), cube_to_dev(xhat), w0, beta, 0.5)
        errs[m] = (rel(sv, sv_ref), rel(v0d.cpu().numpy(), ref), rel(cube_from_dev(g0d), g0))
    print(f'{mode} {Ny}x{Nx}x{L} Q={Q}: sv err {errs[mode][0]:.2e} prox err {errs[mode][1]:.2e} '
          f'(fp64 Gram: {errs["fp64"][0]:.2e} / {errs["fp64"][1]:.2e})')
    assert errs['fp64'][0] < TOL and errs['fp64'][1] < TOL
    assert errs[mode][0] < tol_sv, errs[mode]
    assert errs[mode][1] < tol_prox and errs[mode][2] < tol_prox, errs[mode]


@pytest.mark.parametrize('Ny,Nx,ys,xs,L', [(64, 48, [0, 20, 41, 64], [0, 25, 48], 3),
                                           (128, 96, [0, 64, 128], [0, 30, 61, 96], 4),
                                           (40, 40, [0, 40], [0, 40], 2)])
def test_faceted_sara_transform(ctx, Ny, Nx, ys, xs, L):
    """Per-facet SARA analysis, synthesis and dual step (the faceted wavelet
    transform of src/fhs/fhs_update_dual_sparsity.m:71-93) against the
    oracle's faceted transform; the facets' syntheses add up to the
    whole-image synthesis; the facet sizes follow the reference's coefficient
    count sz = 3 sum(p) - 2 sum(p(nlevel+1:nlevel+1:end)) - 2 p(end), p =
    prod(Ncoefs, 2) (src/fhs/fhs_initialize_dual_and_weights.m:80-81)."""
    import fhs
    from oracle.wavelets import SaraFacets as FacetsRef
    rng = np.random.default_rng(Ny + L)
    ref = FacetsRef(Ny, Nx, ys, xs)
    F = fhs.SaraFacets(ctx, Ny, Nx, ys, xs)
    S = fhs.SaraDictionary(ctx, Ny, Nx)
    x = rng.random((Ny, Nx, L)).astype(np.float32).astype(np.float64)
    xd = cube_to_dev(x)
    c = rng.normal(size=(S.s, L)).astype(np.float32)
    cd = torch.from_numpy(np.ascontiguousarray(c.T)).cuda()
    total = torch.zeros((L, Nx, Ny), dtype=torch.float32, device='cuda')
    for q, f in enumerate(ref.facets):
        assert F.ext(q) == (f['ey0'], f['ex0'], f['eny'], f['enx'])
        n = F.n_coefs(q)
        assert n == len(f['local'])
        p = np.prod(F.ncoefs[q].astype(np.int64), axis=1)
        sz = 3 * p.sum() - 2 * p[S.nlevel::S.nlevel + 1].sum() - 2 * p[-1]
        assert sz == n, (sz, n)
        xe = F.extended(q, xd)
        z = F.analysis(q, xe).cpu().numpy()
        zr = np.stack([ref.analysis(q, ref.extended(q, x[:, :, l])) for l in range(L)])
        assert rel(z, zr) < TOL, q
        # synthesis of this facet's share of a whole-image coefficient vector
        zq = cd[:, torch.from_numpy(f['global']).cuda()].contiguous()
        F.add_extended(q, F.synthesis(q, zq), total)
        # dual step
        v1 = rng.normal(size=(L, n)).astype(np.float32)
        w1 = rng.uniform(0.5, 1.0, n).astype(np.float32)
        beta = 0.3
        v1d, g = F.update_dual(q, torch.from_numpy(v1).cuda(), xe, torch.from_numpy(w1).cuda(), beta)
        vr = soft_l21((v1.astype(np.float64) + zr).T, beta * w1.astype(np.float64)).T
        gr = np.stack([ref.synthesis(q, vr[l]) for l in range(L)])
        assert rel(v1d.cpu().numpy(), vr) < TOL, q
        assert rel(g.cpu().numpy().transpose(0, 2, 1), gr) < TOL, q
    whole = S.psi(cd)
    assert rel(total.cpu().numpy(), whole.cpu().numpy()) < TOL
