"""Oracle: data-fidelity operators and dual update, float64.

Test infrastructure only (see ``oracle/__init__.py``).

Cells ``{L}{nblocks}`` are nested Python lists; ``G[i][j]`` is a scipy
sparse matrix (M_b x nnz(W)), ``W[i][j]`` an index array into the
oversampled grid (MATLAB logical mask -> ``np.flatnonzero``), ``A``/``At``
callables for one channel (``oracle.nufft.NufftOracle``).
"""
import numpy as np

from .prox import sc, solver_proj_elipse_fb


def apply_direct_operator(Fx, G, Lambda=None):
    """``src/apply_direct_operator.m:20``: ``Lambda .* (G * Fx)``."""
    y = G @ Fx
    return y if Lambda is None else Lambda * y


def apply_adjoint_operator(y, G, Lambda=None):
    """``src/apply_adjoint_operator.m:21``: ``G' * (Lambda .* y)``."""
    yy = y if Lambda is None else Lambda * y
    return G.conj().T @ yy


def apply_scaled_fourier_transform(x, A, No):
    """``src/apply_scaled_fourier_transform.m:21-24``: column l = A(x(:,:,l))."""
    L = x.shape[2]
    Fx = np.zeros((No, L), dtype=np.complex128)
    for l in range(L):
        Fx[:, l] = A(x[:, :, l])
    return Fx


def _is_tril(G):
    coo = G.tocoo()
    return bool(np.all(coo.row >= coo.col))


def update_dual_data_fidelity(v2, y, x, Fx_old, proj, A, At, G, W, pU, epsilon,
                              elipse_proj_max_iter, elipse_proj_min_iter,
                              elipse_proj_eps, sigma22,
                              flag_dimensionality_reduction=False, Lambda=None, mapper=map):
    """``src/update_dual_data_fidelity.m:78-147`` (both branches).

    ``A``/``At``: per-channel lists of callables (the reference passes one
    handle; channels of a synthetic cube have their own operators).
    ``mapper``: how the independent channels are visited (``map``, or a
    thread pool's map; the arithmetic of each channel is the same).
    Returns (v2, Ftx, Fx_old, proj, norm_res, global_norm_res, norm_epsilon)
    where ``global_norm_res`` and ``norm_epsilon`` are sums of squares.
    """
    n_channels = x.shape[2]
    Ftx = np.zeros(x.shape)
    v2 = [list(b) for b in v2]
    proj = [list(b) for b in proj]
    Fx_old = Fx_old.copy()

    def channel(i):
        Fx = A[i](x[:, :, i])
        g2 = np.zeros(Fx.shape, dtype=np.complex128)
        nr = []
        for j in range(len(G[i])):
            Gij, Wij = G[i][j], W[i][j]
            if flag_dimensionality_reduction:
                lam = Lambda[i][j]
                fxp = 2 * Fx[Wij] - Fx_old[Wij, i]
                if _is_tril(Gij):
                    hf = Gij @ fxp + Gij.conj().T @ fxp
                else:
                    hf = Gij @ fxp
                v2[i][j] = sc(v2[i][j] + lam * hf - y[i][j], epsilon[i][j])
                if _is_tril(Gij):
                    u2 = lam * v2[i][j]
                    g2[Wij] += Gij.conj().T @ u2 + Gij @ u2
                else:
                    g2[Wij] += Gij.conj().T @ (lam * v2[i][j])
                fs = Fx[Wij]
                if _is_tril(Gij):
                    res = lam * (Gij @ fs + Gij.conj().T @ fs) - y[i][j]
                else:
                    res = lam * (Gij @ fs) - y[i][j]
                nr.append(float(np.sqrt(np.sum(np.abs(res) ** 2))))
            else:
                r2 = Gij @ (2 * Fx[Wij] - Fx_old[Wij, i])
                proj[i][j], _ = solver_proj_elipse_fb(
                    v2[i][j] / pU[i][j], r2, y[i][j], pU[i][j], epsilon[i][j],
                    proj[i][j], elipse_proj_max_iter, elipse_proj_min_iter,
                    elipse_proj_eps)
                v2[i][j] = v2[i][j] + pU[i][j] * (r2 - proj[i][j])
                g2[Wij] += Gij.conj().T @ v2[i][j]
                nr.append(float(np.linalg.norm(Gij @ Fx[Wij] - y[i][j])))
        return Fx, sigma22[i] * np.real(At[i](g2)), nr

    norm_res = []
    global_norm_res = 0.0
    norm_epsilon = 0.0
    for i, (Fx, ftx, nr) in enumerate(mapper(channel, range(n_channels))):
        Fx_old[:, i] = Fx
        Ftx[:, :, i] = ftx
        norm_res.append(nr)
        for j, r in enumerate(nr):        # sums in channel, block order
            global_norm_res += r ** 2
            norm_epsilon += epsilon[i][j] ** 2
    return v2, Ftx, Fx_old, proj, norm_res, global_norm_res, norm_epsilon


def compute_residual_images(x, y, A, At, G, W, flag_visibility_gridding=False, Lambda=None):
    """``src/compute_residual_images.m:39-103`` (non-DR branch and DR branch).

    ``real(At(sum_j G'(y - G Fx)))`` normalised by the PSF peak (the same
    chain applied to a Dirac at pixel (Ny/2, Nx/2), 0-based).
    """
    Ny, Nx, L = x.shape
    out = np.zeros(x.shape)
    dirac = np.zeros((Ny, Nx))
    dirac[Ny // 2, Nx // 2] = 1.0
    for i in range(L):
        Fx = A[i](x[:, :, i])
        Fd = A[i](dirac)
        r = np.zeros(Fx.shape, dtype=np.complex128)
        rd = np.zeros(Fx.shape, dtype=np.complex128)
        for j in range(len(G[i])):
            Gij, Wij = G[i][j], W[i][j]
            if flag_visibility_gridding:
                lam = Lambda[i][j]
                tril = _is_tril(Gij)
                def H(f):
                    return Gij @ f + Gij.conj().T @ f if tril else Gij @ f
                def Ht(f):
                    return Gij.conj().T @ f + Gij @ f if tril else Gij.conj().T @ f
                res_f = lam * (y[i][j] - lam * H(Fx[Wij]))
                r[Wij] += Ht(res_f)
                rd[Wij] += Ht(lam ** 2 * H(Fd[Wij]))
            else:
                r[Wij] += Gij.conj().T @ (y[i][j] - Gij @ Fx[Wij])
                rd[Wij] += Gij.conj().T @ (Gij @ Fd[Wij])
        out[:, :, i] = np.real(At[i](r)) / np.max(np.real(At[i](rd)))
    return out


def holographic_operator(G, y=None, noise=None):
    """Dimensionality-reduced data of one channel,
    ``lib/operators/util_load_data_gen_dr_measurement_operator.m:160-225``,
    from one data set (``holographic_operator_sets`` with one set).

    ``G``: scipy sparse (M visibilities x No grid points), the measurement
    rows with W = the whole grid (the reference's ``G_curr(W_curr, :) =
    G{1}'`` places them so, ``:167-168``).  Flagged rows are assumed removed
    (the reference masks them with ``abs(nW) > 0``, ``:171``).

    Returns ``(H, W, Lambda, y_grid, l2bound, sigma)``:
      H       scipy csr |W| x |W|: ``tril(G_curr * G_curr')`` restricted to
              ``W`` with its diagonal halved (``:196, :215-222``);
      W       ascending grid indices with ``diag > 0`` (``:200-203``);
      Lambda  ``1 ./ sqrt(diag(W))`` (``:206``);
      y_grid  ``(G_curr * y)(W) .* Lambda`` (``:171, :209``), None without y;
      l2bound ``norm((G_curr * noise)(W) .* Lambda)`` and sigma its
              ``std`` (``:186-192, :211-213``), None without noise.

    Pinned (``tests/test_oracle.py``) by a dense ``G^H G`` brute force on
    small random matrices: ``H + H^H`` equals it on ``W x W``, H is lower
    triangular, ``Lambda^2 diag = 1``, W are the non-empty columns of G.
    """
    return holographic_operator_sets([G], None if y is None else [y], None if noise is None else [noise])


def holographic_operator_sets(Gs, ys=None, noises=None):
    """The reference's loop over the data sets (and physical channels) of
    one effective channel, ``util_load_data_gen_dr_measurement_operator.m:
    112-196``: ``H += tril(G_curr * G_curr')`` (``:196``), ``gridded_y +=
    G_curr * y`` (``:171``), the noise ``+= G_curr * noise`` (``:176-192``)
    per set, then W, Lambda, the data and the noise bound of the sums
    (``:200-225``).  ``Gs``: list of scipy sparse (M_d x No); ``ys`` /
    ``noises``: lists of M_d vectors or None.  Same returns as
    ``holographic_operator``.  Pinned (``tests/test_oracle.py``) against
    ``holographic_operator`` of the row-stacked G (a sum of Grams is the Gram
    of the stack)."""
    import scipy.sparse as sp
    No = Gs[0].shape[1]
    Hfull = sp.csr_matrix((No, No), dtype=np.complex128)
    yg = np.zeros(No, dtype=np.complex128)
    gn = np.zeros(No, dtype=np.complex128)
    for d, G in enumerate(Gs):
        Gc = sp.csr_matrix(G).conj().T.tocsr()      # G_curr: No x M_d
        Hfull = Hfull + sp.tril(Gc @ Gc.conj().T)
        if ys is not None:
            yg += Gc @ np.asarray(ys[d])
        if noises is not None:
            gn += Gc @ np.asarray(noises[d])
    Hfull = Hfull.tocsr()
    diagH = np.real(Hfull.diagonal())
    W = np.flatnonzero(diagH > 0)
    Lambda = 1.0 / np.sqrt(diagH[W])
    H = Hfull[W][:, W].tocsr()
    H.setdiag(0.5 * H.diagonal())
    H.eliminate_zeros()
    y_grid = None if ys is None else yg[W] * Lambda
    l2bound = sigma = None
    if noises is not None:
        g = gn[W] * Lambda
        l2bound = float(np.linalg.norm(g))
        sigma = float(np.std(g, ddof=1))
    return H, W, Lambda, y_grid, l2bound, sigma
