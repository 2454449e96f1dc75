"""Pins of the FP64 oracle (CPU only).

The reference ships no golden vectors (MATLAB only; NUFFT and wavelet
arithmetic in un-vendored submodules), so every oracle function is checked
against what the mathematics fixes: published filter values, the exact
non-uniform DFT, closed forms, invariants and special cases."""
import numpy as np
import pytest
from scipy.stats import chi

from oracle import data_fidelity as odf
from oracle import facets as ofc
from oracle import heuristics as oh
from oracle import nufft as onu
from oracle import prox as opx
from oracle import wavelets as owl


# ------------------------------------------------------------------ NUFFT --
@pytest.mark.parametrize('Ny,Nx,M', [(16, 16, 300), (32, 24, 500), (64, 64, 800)])
def test_nufft_matches_exact_nudft(Ny, Nx, M):
    rng = np.random.default_rng(Ny + Nx)
    u = rng.uniform(-np.pi, np.pi, M)
    v = rng.uniform(-np.pi, np.pi, M)
    u[:4] = [-np.pi, np.pi - 1e-9, 0.0, 1e-7]
    op = onu.NufftOracle(u, v, Ny, Nx)
    x = rng.random((Ny, Nx))
    y = op.G @ op.A(x)
    ref = onu.nudft(x, u, v)
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) < 5e-5


def test_nufft_on_grid_frequencies_is_exact_when_sn_constant():
    """With constant scaling factors the min-max interpolator reproduces an
    on-grid frequency exactly (the neighbourhood contains it)."""
    N, K, J = 16, 32, 7
    alpha = np.array([1.0])
    w = 2 * np.pi / K * np.array([3.0, -5.0, 0.0])
    idx, coef = onu.minmax_coef_1d(w, N, J, K, alpha, N // 2)
    x = np.random.default_rng(0).random(N)
    X = np.fft.fft(x, K)
    approx = np.sum(coef * X[idx], axis=1)
    exact = np.array([np.sum(x * np.exp(-1j * wi * (np.arange(N) - N // 2))) for wi in w])
    assert np.max(np.abs(approx - exact)) < 1e-10


def test_nufft_adjoint_and_fft_convention():
    rng = np.random.default_rng(3)
    op = onu.NufftOracle(np.zeros(1), np.zeros(1), 8, 4)
    x = rng.random((8, 4))
    g = rng.normal(size=op.No) + 1j * rng.normal(size=op.No)
    assert abs(np.vdot(op.A(x), g) - np.vdot(x, op.At(g))) < 1e-12 * abs(np.vdot(op.A(x), g))
    # A against an explicit DFT matrix of the padded, scaled image (column-major vector)
    pad = np.zeros((op.Kty, op.Ktx))
    pad[:8, :4] = op.sn * x
    Fy = np.exp(-2j * np.pi * np.outer(np.arange(op.Kty), np.arange(op.Kty)) / op.Kty)
    Fx = np.exp(-2j * np.pi * np.outer(np.arange(op.Ktx), np.arange(op.Ktx)) / op.Ktx)
    assert np.allclose(op.A(x), (Fy @ pad @ Fx.T).ravel(order='F'), atol=1e-10)


def test_product_g_builder_matches_oracle():
    import fhs.nufft as pn
    rng = np.random.default_rng(5)
    u = rng.uniform(-np.pi, np.pi, 2000)
    v = rng.uniform(-np.pi, np.pi, 2000)
    P = pn.Nufft2D(32, 64)
    a, c = P.window(u, v)
    O = onu.NufftOracle(u, v, 32, 64)
    assert np.array_equal(a, O.anchor)
    assert np.max(np.abs(c - O.coef)) < 1e-10 * np.max(np.abs(O.coef))
    assert np.max(np.abs(P.ay.s - O.sy)) < 1e-12


def test_uniform_weights():
    u = np.array([0.0, 0.001, 1.0, -3.0])
    v = np.array([0.0, 0.001, 1.0, 2.0])
    w = onu.uniform_precond_weights(u, v, 64, 64)
    assert np.allclose(w, [0.5, 0.5, 1.0, 1.0])


# --------------------------------------------------------------- wavelets --
DB4 = [0.2303778133, 0.7148465706, 0.6308807679, -0.0279837694,
       -0.1870348117, 0.0308413818, 0.0328830117, -0.0105974018]


def test_daubechies_published_values():
    r3 = np.sqrt(3)
    db2 = np.array([1 + r3, 3 + r3, 3 - r3, 1 - r3]) / (4 * np.sqrt(2))
    assert np.allclose(owl.daubechies_lo_r(2), db2, atol=1e-14)
    assert np.allclose(owl.daubechies_lo_r(4), DB4, atol=1e-9)
    assert np.allclose(owl.daubechies_lo_r(1), [2 ** -0.5, 2 ** -0.5])


@pytest.mark.parametrize('N', range(1, 9))
def test_daubechies_orthonormal_and_moments(N):
    lo, hi = owl.filters(f'db{N}')
    lf = len(lo)
    assert lf == 2 * N and abs(lo.sum() - np.sqrt(2)) < 1e-12
    for m in range(N):
        s = np.dot(lo[:lf - 2 * m], lo[2 * m:])
        assert abs(s - (1.0 if m == 0 else 0.0)) < 1e-11
    for p in range(N):   # vanishing moments of the wavelet
        assert abs(np.sum(np.arange(lf) ** p * hi)) < 1e-8 * max(1, lf ** p)


def test_product_filters_match_oracle():
    from fhs.sara import daubechies_scaling_filter
    for N in range(1, 9):
        assert np.allclose(daubechies_scaling_filter(N), owl.daubechies_lo_r(N), atol=1e-11)


def test_haar_by_hand():
    x = np.arange(16, dtype=float).reshape(4, 4, order='F')
    c = owl.wavedec2(x, 'db1', 1)
    a = np.array([[x[0, 0] + x[1, 0] + x[0, 1] + x[1, 1], x[0, 2] + x[1, 2] + x[0, 3] + x[1, 3]],
                  [x[2, 0] + x[3, 0] + x[2, 1] + x[3, 1], x[2, 2] + x[3, 2] + x[2, 3] + x[3, 3]]]) / 2
    assert np.allclose(c[:4], a.ravel(order='F'))
    # horizontal detail h = hi_y lo_x: (row difference) summed across columns
    h00 = ((x[0, 0] + x[0, 1]) - (x[1, 0] + x[1, 1])) / 2
    assert np.isclose(c[4], h00)


@pytest.mark.parametrize('Ny,Nx', [(37, 29), (64, 64), (5, 3), (1, 9)])
def test_sara_dictionary_isometry_and_adjoint(Ny, Nx):
    rng = np.random.default_rng(Ny * Nx)
    D = owl.SaraDictionary(Ny, Nx)
    x = rng.random((Ny, Nx))
    c = D.psit(x)
    assert abs(np.linalg.norm(c) - np.linalg.norm(x)) < 1e-10 * np.linalg.norm(x)
    assert np.max(np.abs(D.psi(c) - x)) < 1e-10
    cc = rng.normal(size=D.s)
    assert abs(np.dot(c, cc) - np.sum(x * D.psi(cc))) < 1e-10 * np.linalg.norm(cc) * np.linalg.norm(x)


def test_wavelet_vanishing_moments_2d():
    """db3 detail coefficients of a quadratic image vanish away from the border."""
    y, x = np.mgrid[0:64, 0:64].astype(float)
    img = 0.3 * y ** 2 - 0.2 * x * y + x + 2
    c = owl.wavedec2(img, 'db3', 1)
    n = owl.n_out(64, 6)
    d = c[4 * 0 + n * n:].reshape(3, n, n)   # H, V, D of level 1 (nlevel = 1)
    inner = d[:, 3:-3, 3:-3]
    assert np.max(np.abs(inner)) < 1e-8 * np.max(np.abs(img))


def test_coefficient_count():
    D = owl.SaraDictionary(512, 512)
    assert abs(D.s / 512 ** 2 - 9.4433) < 1e-3


# ------------------------------------------------------------------ prox ---
def test_proj_l2ball_cases():
    rng = np.random.default_rng(1)
    x = rng.normal(size=50) + 1j * rng.normal(size=50)
    y = rng.normal(size=50) + 1j * rng.normal(size=50)
    d = np.linalg.norm(x - y)
    assert np.allclose(opx.proj_l2ball(x, 2 * d, y), x)             # inside: identity
    p = opx.proj_l2ball(x, 0.5 * d, y)
    assert abs(np.linalg.norm(p - y) - 0.5 * d) < 1e-12            # on the sphere
    assert np.allclose((p - y) / np.linalg.norm(p - y), (x - y) / d)  # same direction
    assert np.allclose(opx.proj_l2ball(p, 0.5 * d, y), p)          # idempotent


def test_ellipsoid_constant_metric_is_l2_projection():
    rng = np.random.default_rng(2)
    M = 200
    v2 = rng.normal(size=M) + 1j * rng.normal(size=M)
    r = rng.normal(size=M) + 1j * rng.normal(size=M)
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    pU = np.full(M, 0.7)
    eps = 3.0
    p, _ = opx.solver_proj_elipse_fb(v2 / pU, r, y, pU, eps, np.zeros(M, complex), 20, 1, 1e-8)
    assert np.allclose(p, opx.proj_l2ball(v2 / pU + r, eps, y), atol=1e-12)


def test_ellipsoid_optimality():
    """Converged FB point is feasible and satisfies the KKT condition
    pU (p - w) = -lam (p - y), lam >= 0, of the weighted projection."""
    rng = np.random.default_rng(4)
    M = 100
    w = 3 * (rng.normal(size=M) + 1j * rng.normal(size=M))
    y = np.zeros(M, complex)
    pU = rng.uniform(0.5, 1.0, M)
    eps = 2.0
    p, it = opx.solver_proj_elipse_fb(w, 0 * w, y, pU, eps, np.zeros(M, complex), 5000, 1, 1e-14)
    assert np.linalg.norm(p - y) <= eps * (1 + 1e-10)
    g = pU * (p - w)
    lam = -np.real(np.vdot(p - y, g)) / np.linalg.norm(p - y) ** 2
    assert lam >= 0
    assert np.linalg.norm(g + lam * (p - y)) < 1e-6 * np.linalg.norm(g)


def test_l21_is_row_projection():
    rng = np.random.default_rng(6)
    w = rng.normal(size=(300, 5))
    t = rng.uniform(0, 3, 300)
    v = opx.soft_l21(w, t)
    l2 = np.linalg.norm(w, axis=1)
    lv = np.linalg.norm(v, axis=1)
    assert np.allclose(lv, np.minimum(l2, t), atol=1e-12)
    assert np.allclose(v[l2 <= t], w[l2 <= t])
    cos = np.sum(v * w, axis=1) / (lv * l2 + 1e-300)
    assert np.all(cos[lv > 0] > 1 - 1e-12)
    # L = 1 is the l1 step of SARA (src/sara/sara.m:925) up to eps
    r = rng.normal(size=(50, 1))
    assert np.allclose(opx.soft_l21(r, np.full(50, 0.3))[:, 0], opx.soft_l1_dual(r[:, 0], 0.3), atol=1e-14)


def test_nuclear_prox_and_norm():
    rng = np.random.default_rng(7)
    Z = rng.normal(size=(40, 6))
    s = np.linalg.svd(Z, compute_uv=False)
    t = np.full(6, np.median(s))
    v = opx.nuclear_prox_dual(Z, t)
    assert np.allclose(np.sort(np.linalg.svd(v, compute_uv=False)), np.sort(np.minimum(s, t)), atol=1e-12)
    assert np.allclose(opx.nuclear_prox_dual(Z, np.zeros(6)), 0, atol=1e-12)
    assert np.allclose(opx.nuclear_prox_dual(Z, np.full(6, 1e9)), Z)
    a, b = rng.normal(size=(5, 4, 1)), rng.normal(size=3)
    x = a * b                                  # rank one (5, 4, 3)
    assert abs(opx.nuclear_norm(x) - np.linalg.norm(a) * np.linalg.norm(b)) < 1e-12
    X = rng.random((6, 5, 4))
    ev = np.linalg.eigvalsh(X.reshape(-1, 4, order='F').T @ X.reshape(-1, 4, order='F'))
    assert abs(opx.nuclear_norm(X) - np.sum(np.sqrt(np.maximum(ev, 0)))) < 1e-10


def test_update_primal():
    x = np.array([1.0, -1.0, 2.0])
    g = np.array([2.0, -3.0, 0.5])
    xn, xhat, rel, nrm = opx.update_primal(x, g)
    assert np.allclose(xn, [0.0, 2.0, 1.5]) and np.allclose(xhat, [-1.0, 5.0, 1.0])
    assert rel == 1 + 9 + 0.25 and nrm == 4 + 2.25


def test_update_epsilon_behaviour():
    eps, tb = opx.update_epsilon([[10.0, 10.0, 10.0, 10.0]], 50, [[0, 0, 0, 45]], [[5.0, 20.0, 10.0, 5.0]],
                                 0.99, 1.001, 40, 0.5)
    assert eps[0] == [7.5, 15.0, 10.0, 10.0]      # down, up, unchanged, too recent
    assert tb[0] == [50, 50, 0, 45]


# ------------------------------------------------------------ data term ---
def test_data_fidelity_step_dense_composition():
    """update_dual_data_fidelity with pU = 1 against a dense re-derivation."""
    rng = np.random.default_rng(8)
    Ny = Nx = 8
    M = 60
    u = rng.uniform(-np.pi, np.pi, M)
    v = rng.uniform(-np.pi, np.pi, M)
    op = onu.NufftOracle(u, v, Ny, Nx)
    Gd = op.G.toarray()
    Adense = np.stack([op.A(e.reshape(Ny, Nx, order='F')) for e in np.eye(Ny * Nx)], axis=1)
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    x = rng.random((Ny, Nx, 1))
    Fx_old = (Adense @ rng.random(Ny * Nx))[:, None]
    v2 = [[rng.normal(size=M) + 0j]]
    eps = 0.3 * np.linalg.norm(y)
    out = odf.update_dual_data_fidelity(v2, [[y]], x, Fx_old, [[np.zeros(M, complex)]], [op.A], [op.At],
                                        [[op.G]], [[np.arange(op.No)]], [[np.ones(M)]], [[eps]],
                                        20, 1, 1e-12, [0.4])
    Fx = Adense @ x[:, :, 0].ravel(order='F')
    r2 = Gd @ (2 * Fx - Fx_old[:, 0])
    w = v2[0][0] + r2
    proj = y + (w - y) * min(1, eps / np.linalg.norm(w - y))
    v2n = v2[0][0] + r2 - proj
    assert np.allclose(out[0][0][0], v2n)
    Ftx = 0.4 * np.real(Adense.conj().T @ (Gd.conj().T @ v2n)).reshape(Ny, Nx, order='F')
    assert np.allclose(out[1][:, :, 0], Ftx)
    assert np.isclose(out[4][0][0], np.linalg.norm(Gd @ Fx - y))
    # linearity used by the GPU state: G(2Fx - Fx_old) = 2 G Fx - G Fx_old
    assert np.allclose(r2, 2 * Gd @ Fx - Gd @ Fx_old[:, 0])


@pytest.mark.parametrize('kind', ['random', 'nufft'])
def test_holographic_operator_dense(kind):
    """The DR data (H = tril(G'G)(W, W) with a halved diagonal, W, Lambda,
    gridded data and noise bound) against a dense G^H G brute force."""
    import scipy.sparse as sp
    rng = np.random.default_rng(21)
    if kind == 'random':
        M, No = 40, 30
        G = sp.random(M, No, density=0.15, random_state=3, format='csr').astype(np.complex128)
        G.data = rng.normal(size=G.nnz) + 1j * rng.normal(size=G.nnz)
        G = G.tolil()
        G[:, 7] = 0                                   # an empty column: not in W
        G = G.tocsr()
        G.eliminate_zeros()
    else:
        op = onu.NufftOracle(rng.normal(0, 0.8, 50), rng.normal(0, 0.8, 50), 8, 8)
        G, No, M = op.G.tocsr(), op.No, 50
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    n = (rng.normal(size=M) + 1j * rng.normal(size=M)) / np.sqrt(2)
    H, W, lam, yg, l2, sig = odf.holographic_operator(G, y, n)
    Gd = G.toarray()
    GhG = Gd.conj().T @ Gd
    Wd = np.flatnonzero(np.linalg.norm(Gd, axis=0) > 0)
    assert np.array_equal(W, Wd)
    if kind == 'random':
        assert 7 not in W and len(W) < No
    Hd = H.toarray()
    assert np.allclose(np.triu(Hd, 1), 0)
    assert np.allclose(Hd + Hd.conj().T, GhG[np.ix_(W, W)])
    assert np.allclose(lam ** 2 * np.real(np.diag(GhG))[W], 1.0)
    assert np.allclose(yg, lam * (Gd.conj().T @ y)[W])
    assert np.isclose(l2, np.linalg.norm(lam * (Gd.conj().T @ n)[W]))
    gn = lam * (Gd.conj().T @ n)[W]
    assert np.isclose(sig, np.sqrt(np.sum(np.abs(gn - gn.mean()) ** 2) / (len(gn) - 1)))
    # the DR operator Lambda (H + H') Lambda is Lambda G'G Lambda on W
    f = rng.normal(size=len(W)) + 1j * rng.normal(size=len(W))
    assert np.allclose(lam * (H @ f + H.conj().T @ f), lam * (GhG[np.ix_(W, W)] @ f))


def test_holographic_operator_sets_sum_data_sets():
    """Several data sets of one channel (the reference's per-set loop):
    against a dense brute force of the summed Grams, and equal to one set of
    the stacked rows (a sum of Grams is the Gram of the stack)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    ops = [onu.NufftOracle(rng.normal(0, 0.6, m), rng.normal(0, 0.6, m), 8, 8) for m in (30, 45, 12)]
    Gs = [op.G.tocsr() for op in ops]
    ys = [rng.normal(size=G.shape[0]) + 1j * rng.normal(size=G.shape[0]) for G in Gs]
    ns = [(rng.normal(size=G.shape[0]) + 1j * rng.normal(size=G.shape[0])) / np.sqrt(2) for G in Gs]
    H, W, lam, yg, l2, sig = odf.holographic_operator_sets(Gs, ys, ns)
    GhG = sum(G.toarray().conj().T @ G.toarray() for G in Gs)
    Gty = sum(G.toarray().conj().T @ y for G, y in zip(Gs, ys))
    assert np.array_equal(W, np.flatnonzero(np.real(np.diag(GhG)) > 0))
    Hd = H.toarray()
    assert np.allclose(Hd + Hd.conj().T, GhG[np.ix_(W, W)])
    assert np.allclose(lam ** 2 * np.real(np.diag(GhG))[W], 1.0)
    assert np.allclose(yg, lam * Gty[W])
    # one set from each data set alone is not the sum
    H1, W1 = odf.holographic_operator_sets(Gs[:1], ys[:1])[:2]
    assert len(W1) < len(W) or not np.allclose(H1.toarray(), Hd[np.ix_(np.searchsorted(W, W1), np.searchsorted(W, W1))])
    # stacked rows as one set: the same DR data
    Hs, Ws, lams, ygs, l2s, sigs = odf.holographic_operator(sp.vstack(Gs).tocsr(), np.concatenate(ys),
                                                            np.concatenate(ns))
    assert np.array_equal(Ws, W)
    assert np.allclose(Hs.toarray(), Hd) and np.allclose(lams, lam) and np.allclose(ygs, yg)
    assert np.isclose(l2s, l2) and np.isclose(sigs, sig)


def test_dirty_image_of_point_source_is_normalised_psf():
    rng = np.random.default_rng(9)
    Ny = Nx = 16
    u = rng.normal(0, 1, 400)
    v = rng.normal(0, 1, 400)
    op = onu.NufftOracle(u, v, Ny, Nx)
    d = np.zeros((Ny, Nx, 1))
    d[Ny // 2, Nx // 2, 0] = 1.0
    y = op.G @ op.A(d[:, :, 0])
    img = odf.compute_residual_images(np.zeros((Ny, Nx, 1)), [[y]], [op.A], [op.At], [[op.G]],
                                      [[np.arange(op.No)]])
    assert abs(img[Ny // 2, Nx // 2, 0] - 1.0) < 1e-12
    assert np.argmax(img[:, :, 0]) == np.ravel_multi_index((Ny // 2, Nx // 2), (Ny, Nx))


# ---------------------------------------------------------- heuristics ----
def test_noise_level_chi_moments():
    for L in (1, 4, 16):
        sig, sig_bar, mu_chi, sig_chi, _ = oh.compute_noise_level_hs(64, 64, L, np.ones(L), np.full(L, 4.0))
        assert abs(mu_chi - chi.mean(L)) < 1e-12 and abs(sig_chi - chi.std(L)) < 1e-10
        assert abs(sig - 0.5 * (chi.mean(L) + chi.std(L))) < 1e-10


def test_data_fidelity_bounds():
    e = oh.data_fidelity_bounds([100, 300])
    Ml = 400
    tot = np.sqrt(Ml + 2 * np.sqrt(Ml))
    assert np.isclose(np.sqrt(e[0] ** 2 + e[1] ** 2), tot)


def test_op_norm_power_iteration():
    rng = np.random.default_rng(10)
    A = rng.normal(size=(30, 20))
    val, _ = oh.op_norm(lambda x: A @ x, lambda y: A.T @ y, (20,), 1e-12, 2000)
    assert abs(val - np.linalg.norm(A, 2) ** 2) < 1e-8 * val


# ---------------------------------------------------------------- facets --
def test_facets_partition_and_windows():
    # MATLAB round(linspace(0, 10, 5)) = [0 3 5 8 10] (halves away from zero)
    assert np.array_equal(ofc.split_range(4, 10), [[0, 3], [3, 5], [5, 8], [8, 10]])
    geom = ofc.facet_geometry(96, 120, 3, 4, (0.5, 0.25))
    cover = np.zeros((96, 120))
    for q in range(12):
        y0, x0, ny, nx = geom['oy'][q], geom['ox'][q], geom['py'][q], geom['px'][q]
        cover[y0:y0 + ny, x0:x0 + nx] += geom['windows'][q]
    assert np.allclose(cover, 1.0)
    g1 = ofc.facet_geometry(40, 30, 1, 1, (0.5, 0.5))
    assert g1['oy'] == [0] and g1['py'] == [40] and np.allclose(g1['windows'][0], 1)


def test_faceted_lowrank_single_facet_is_global():
    rng = np.random.default_rng(11)
    x = rng.random((12, 10, 3))
    geom = ofc.facet_geometry(12, 10, 1, 1, (0, 0))
    v0s, g0 = ofc.faceted_lowrank_step([np.zeros((120, 3))], x, geom, np.ones((1, 3)), 0.5, 1.0)
    ref = opx.nuclear_prox_dual(x.reshape(120, 3, order='F'), np.full(3, 0.5))
    assert np.allclose(v0s[0], ref) and np.allclose(g0, ref.reshape(12, 10, 3, order='F'))


@pytest.mark.parametrize('N,Q,f', [(64, 2, 0.25), (1024, 4, 0.25), (1000, 3, 0.5), (512, 1, 0.25), (77, 5, 0.1)])
def test_product_facet_geometry_matches_oracle(N, Q, f):
    """The product's facet builder (used by the faceted solver and the bench)
    equals the oracle's restatement, windows included."""
    from fhs.lowrank import facet_geometry as product
    from oracle.facets import facet_geometry as ref
    a, b = product(N, N // 2 + 1, Q, max(Q - 1, 1), (f, f)), ref(N, N // 2 + 1, Q, max(Q - 1, 1), (f, f))
    for k in ('Qy', 'Qx', 'oy', 'ox', 'py', 'px', 'overlap'):
        assert list(np.atleast_1d(a[k])) == list(np.atleast_1d(b[k])), k
    for wa, wb in zip(a['windows'], b['windows']):
        assert np.array_equal(wa, wb)


@pytest.mark.parametrize('Ny,Nx,ys,xs', [(64, 48, [0, 20, 41, 64], [0, 25, 48]),
                                         (100, 90, [0, 50, 100], [0, 30, 61, 90]),
                                         (37, 29, [0, 37], [0, 29])])
def test_faceted_sara_is_the_global_transform(Ny, Nx, ys, xs):
    """The faceted SARA transform (uneven facets): every global coefficient is
    owned by exactly one facet, each facet's analysis of its extended image
    gives those global coefficients, and the facets' syntheses overlap-added
    give the global synthesis (the exact pieces the reference's faceted
    wavelet transform is built to produce)."""
    from oracle.wavelets import SaraFacets
    rng = np.random.default_rng(Ny)
    F = SaraFacets(Ny, Nx, ys, xs)
    x = rng.random((Ny, Nx))
    g = F.glob.psit(x)
    owned = np.concatenate([f['global'] for f in F.facets])
    assert np.array_equal(np.sort(owned), np.arange(F.glob.s))
    for q, f in enumerate(F.facets):
        assert np.allclose(F.analysis(q, F.extended(q, x)), g[f['global']], rtol=0, atol=1e-13)
        assert f['ey0'] % 16 == 0 and f['ex0'] % 16 == 0
    c = rng.random(F.glob.s)
    out = np.zeros((Ny, Nx))
    for q, f in enumerate(F.facets):
        F.add_extended(q, F.synthesis(q, c[f['global']]), out)
    assert np.allclose(out, F.glob.psi(c), rtol=0, atol=1e-12)
    # a facet's coefficients do not depend on pixels outside its extended image
    q = len(F.facets) - 1
    f = F.facets[q]
    x2 = x.copy()
    x2[:max(f['ey0'], 0), :] = 5.0
    x2[:, :max(f['ex0'], 0)] = 5.0
    assert np.allclose(F.glob.psit(x2)[f['global']], g[f['global']], rtol=0, atol=1e-13)
