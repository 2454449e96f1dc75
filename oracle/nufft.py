"""Oracle: min-max NUFFT (degridding matrix G, scaled FFT A / At), float64.

Test infrastructure only (see ``oracle/__init__.py``).

Reference semantics
-------------------
* ``op_p_nufft([{v} {u}], [Ny Nx], [Ky Kx], [oy*Ny ox*Nx], [Ny/2 Nx/2], nW,
  [], kernel)`` is called at ``lib/operators/util_gen_measurement_operator_synth.m:85``
  (submodule RI-measurement-operator, un-vendored).  It returns
  ``A = @(x) fft2(pad(sn .* x))(:)``, ``At`` its exact adjoint, the sparse
  interpolation matrix ``G`` (M x No, J = Ky*Kx nonzeros per row) and the
  grid mask ``W``.  The first image axis (rows, y) pairs with ``v``, the
  second (columns, x) with ``u``; the phase centre is pixel ``N/2``.
* The interpolator is Fessler & Sutton's min-max interpolator (IEEE TSP
  2003, ``docs/source/biblio.bib:65``): for each frequency the J neighbour
  coefficients minimise the worst-case error over unit-norm signals given
  scaling factors ``sn``.  With ``sn`` a cosine series the required sums are
  Dirichlet kernels (closed form); the J x J Gram matrix is frequency
  independent.
* Scaling factors: cosine series (``SCALING_TERMS`` terms) least-squares
  fitted to the reciprocal Kaiser-Bessel transform, the 'minmax:kb' recipe
  of the same toolbox.  The exact 'minmax:tuned' tables of the reference
  default (``imaging/default_parameters.json:45``) are not in the container:
  coefficient values are **parity unpinned** against the reference; they are
  pinned against the exact non-uniform DFT instead (tests/test_oracle_nufft.py).
* W is the full oversampled grid (every column kept), so ``Fx(W) == Fx``.

Layout: column-major (MATLAB) vectorisation everywhere; grid index of
(ky, kx) is ``ky + Kty*kx``.
"""
import numpy as np
import scipy.sparse as sp

SCALING_TERMS = 6          # cosine terms in the fitted scaling factors
KB_ALPHA_PER_J = 2.34      # Kaiser-Bessel shape alpha = 2.34 J (K/N = 2)


def kb_scaling_target(N, J, K):
    """Reciprocal of the continuous Kaiser-Bessel transform at (n - n0)/K."""
    n0 = (N - 1) / 2.0
    u = (np.arange(N) - n0) / K
    alpha = KB_ALPHA_PER_J * J
    z = np.sqrt(alpha ** 2 - (np.pi * J * u) ** 2 + 0j)
    ft = np.real(np.sinh(z) / z)
    return 1.0 / ft


def fit_scaling(N, J, K, terms=None):
    """alpha_l such that sum_l alpha_l cos(gamma l (n-n0)) ~ kb target."""
    terms = SCALING_TERMS if terms is None else terms
    n0 = (N - 1) / 2.0
    gam = 2 * np.pi / K
    n = np.arange(N) - n0
    basis = np.cos(gam * np.outer(n, np.arange(terms)))
    target = kb_scaling_target(N, J, K)
    alpha, *_ = np.linalg.lstsq(basis, target, rcond=None)
    return alpha / alpha[0]


def scaling_factors(N, K, alpha):
    n0 = (N - 1) / 2.0
    gam = 2 * np.pi / K
    n = np.arange(N) - n0
    return np.cos(gam * np.outer(n, np.arange(len(alpha)))) @ alpha


def dirichlet(t, N):
    """sum_{n=0}^{N-1} exp(i t (n - (N-1)/2)) = sin(N t/2)/sin(t/2)."""
    t = np.asarray(t, dtype=np.float64)
    s = np.sin(t / 2)
    small = np.abs(s) < 1e-12
    out = np.empty_like(t)
    out[~small] = np.sin(N * t[~small] / 2) / s[~small]
    out[small] = N
    return out


def minmax_gram_inverse(N, J, K, alpha):
    """T = (A^H A)^{-1}, (A^H A)_{jj'} = sum_n s_n^2 cos(gamma (j-j')(n-n0))."""
    s = scaling_factors(N, K, alpha)
    n0 = (N - 1) / 2.0
    gam = 2 * np.pi / K
    n = np.arange(N) - n0
    d = np.arange(J)[:, None] - np.arange(J)[None, :]
    gram = np.einsum('n,jkn->jk', s ** 2, np.cos(gam * d[:, :, None] * n[None, None, :]))
    return np.linalg.inv(gram)


def minmax_coef_1d(om, N, J, K, alpha, nc):
    """Min-max interpolation coefficients along one axis.

    Returns (idx, coef): idx (M, J) grid indices mod K, coef (M, J) complex,
    such that sum_n x_n exp(-i om (n - nc)) ~ sum_j coef_j X[idx_j] with
    X = fft(s * x, K).
    """
    om = np.asarray(om, dtype=np.float64)
    gam = 2 * np.pi / K
    n0 = (N - 1) / 2.0
    T = minmax_gram_inverse(N, J, K, alpha)
    k0 = np.floor(om / gam - J / 2.0)
    kj = k0[:, None] + np.arange(1, J + 1)[None, :]                 # (M, J)
    t = gam * kj - om[:, None]
    r = np.zeros_like(t)
    for l, a in enumerate(alpha):
        if l == 0:
            r += a * dirichlet(t, N)
        else:
            r += 0.5 * a * (dirichlet(t + gam * l, N) + dirichlet(t - gam * l, N))
    c = r @ T.T
    coef = c * np.exp(1j * gam * kj * n0) * np.exp(-1j * om[:, None] * (n0 - nc))
    idx = np.mod(kj, K).astype(np.int64)
    return idx, coef


class NufftOracle:
    """A, At and G for one channel (one set of uv points)."""

    def __init__(self, u, v, Ny, Nx, Ky=7, Kx=7, oy=2, ox=2):
        self.Ny, self.Nx = Ny, Nx
        self.Kty, self.Ktx = oy * Ny, ox * Nx
        self.No = self.Kty * self.Ktx
        self.ay = fit_scaling(Ny, Ky, self.Kty)
        self.ax = fit_scaling(Nx, Kx, self.Ktx)
        self.sy = scaling_factors(Ny, self.Kty, self.ay)
        self.sx = scaling_factors(Nx, self.Ktx, self.ax)
        self.sn = np.outer(self.sy, self.sx)
        iy, cy = minmax_coef_1d(v, Ny, Ky, self.Kty, self.ay, Ny // 2)
        ix, cx = minmax_coef_1d(u, Nx, Kx, self.Ktx, self.ax, Nx // 2)
        M = len(u)
        cols = (iy[:, None, :] + self.Kty * ix[:, :, None]).reshape(M, -1)
        vals = (cy[:, None, :] * cx[:, :, None]).reshape(M, -1)
        rows = np.repeat(np.arange(M), Ky * Kx)
        self.G = sp.csr_matrix((vals.ravel(), (rows, cols.ravel())), shape=(M, self.No))
        self.G.sort_indices()
        # dense-window form: anchor = first tap (ky0 + Kty*kx0), taps jx*Ky + jy
        self.anchor = (iy[:, 0] + self.Kty * ix[:, 0]).astype(np.int64)
        self.coef = vals
        self.cy_sep, self.cx_sep = cy, cx        # G[m, (jy, jx)] = cy[m, jy] cx[m, jx]
        self.Ky, self.Kx = Ky, Kx

    def A(self, x):
        """fft2 of the scaled, zero-padded image, column-major vector (No,)."""
        pad = np.zeros((self.Kty, self.Ktx), dtype=np.complex128)
        pad[:self.Ny, :self.Nx] = self.sn * x
        return np.fft.fft2(pad).ravel(order='F')

    def At(self, g):
        """Exact adjoint of A (complex image; callers take the real part)."""
        grid = np.reshape(g, (self.Kty, self.Ktx), order='F')
        full = np.fft.ifft2(grid) * self.No
        return self.sn * full[:self.Ny, :self.Nx]


def nudft(x, u, v):
    """Exact NUDFT with phase centre (Ny/2, Nx/2): the quantity G A x approximates."""
    Ny, Nx = x.shape
    ny = np.arange(Ny) - Ny // 2
    nx = np.arange(Nx) - Nx // 2
    Ey = np.exp(-1j * np.outer(v, ny))
    Ex = np.exp(-1j * np.outer(u, nx))
    return np.einsum('my,yx,mx->m', Ey, x, Ex)


def uniform_precond_weights(u, v, Kty, Ktx):
    """Uniform-weighting preconditioner aW = 1 / (uv points in the same grid cell).

    Restates util_gen_preconditioning_matrix (submodule) with
    gen_uniform_weight_matrix = true and uniform_weight_sub_pixels = 1
    (``imaging/default_parameters.json:47-51``).  Parity unpinned (the
    submodule source is absent); semantics restated.
    """
    cy = np.floor((np.asarray(v) + np.pi) / (2 * np.pi) * Kty).astype(np.int64) % Kty
    cx = np.floor((np.asarray(u) + np.pi) / (2 * np.pi) * Ktx).astype(np.int64) % Ktx
    cell = cy + Kty * cx
    counts = np.bincount(cell, minlength=Kty * Ktx)
    return 1.0 / counts[cell]
