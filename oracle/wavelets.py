"""Oracle: SARA dictionary (db1-db8 + Dirac, zero-padded multilevel DWT), float64.

Test infrastructure only (see ``oracle/__init__.py``).

Reference semantics
-------------------
* The SARA dictionary is ``op_sp_wlt_basis(wavelet, nlevel, Ny, Nx)``
  (submodule SARA-dictionary, un-vendored), built at
  ``src/hs/hyperSARA.m`` (``spmd [Psi_, Psit_] = op_sp_wlt_basis(...)``) and
  ``imaging/imaging.m:797-810``, with bases ``db1..db8`` + ``'self'`` and
  4 levels (``imaging/default_parameters.json:52-57``), extension mode 'zpd'
  (``lib/utils/util_set_parpool.m:51``).  ``Psit`` stacks the per-basis
  ``wavedec2`` coefficient vectors (Dirac last), each basis scaled by
  ``1/sqrt(P)`` so that ``Psi Psit = I`` (``nu1 = 1``, ``imaging/imaging.m:845``).
* Filters: Daubechies extremal-phase (db N has 2N taps, N vanishing moments),
  obtained by spectral factorisation of the Daubechies polynomial, keeping
  the roots inside the unit circle.  Pinned against the published db2
  closed form and db4 values, orthonormality and vanishing moments.
* One 'zpd' analysis step on a length-n signal (MATLAB ``dwt``): full
  convolution with the decomposition filter, keep odd samples ->
  ``floor((n + lf - 1)/2)`` coefficients.  With the reconstruction filter f
  (``Lo_R`` or ``Hi_R``) that is ``out[k] = sum_j f[j] x[2k + 2 - lf + j]``,
  x zero outside [0, n).  ``Hi_R[k] = (-1)^k Lo_R[lf-1-k]`` (MATLAB ``qmf``).
* 2-D step (MATLAB ``dwt2``): axis 1 is y (rows), axis 2 is x (columns);
  ``a = lo_y lo_x``, ``h = hi_y lo_x``, ``v = lo_y hi_x``, ``d = hi_y hi_x``.
* Coefficient vector (MATLAB ``wavedec2``): ``[A_J, H_J, V_J, D_J, H_{J-1},
  V_{J-1}, D_{J-1}, ..., H_1, V_1, D_1]``, each column-major.
* Synthesis ``Psi`` is the exact adjoint of ``Psit``; for zpd the analysis
  is an isometry, so it is also the left inverse (``waverec2``).

The exact coefficient ordering of ``sdwt2_sara_faceting`` (faceted variant)
cannot be checked (submodule absent): parity unpinned; the faceted ordering
used here is documented in ``oracle/facets.py``.
"""
import numpy as np
from math import comb

DEFAULT_BASES = ('db1', 'db2', 'db3', 'db4', 'db5', 'db6', 'db7', 'db8', 'self')
DEFAULT_NLEVEL = 4


def daubechies_lo_r(N):
    """Reconstruction low-pass filter of dbN (sum = sqrt(2)), length 2N."""
    if N == 1:
        return np.array([1.0, 1.0]) / np.sqrt(2.0)
    # P(y) = sum_{k<N} C(N-1+k, k) y^k ; y = (2 - z - 1/z)/4 = -(z-1)^2/(4z)
    # z^{N-1} P(y(z)) is a polynomial in z of degree 2N-2 with roots in pairs (r, 1/r)
    poly = np.zeros(2 * N - 1)
    for k in range(N):
        # y^k z^{N-1} = (-1)^k (z-1)^{2k} z^{N-1-k} / 4^k
        term = np.array([1.0])
        for _ in range(2 * k):
            term = np.convolve(term, [1.0, -1.0])
        term = term * ((-1.0) ** k) * comb(N - 1 + k, k) / 4.0 ** k
        # multiply by z^{N-1-k}: coefficient array highest power first
        term = np.concatenate([term, np.zeros(N - 1 - k)])
        poly[-len(term):] += term
    roots = np.roots(poly)
    inside = roots[np.abs(roots) < 1.0]
    q = np.real(np.poly(inside))
    h = q
    for _ in range(N):
        h = np.convolve(h, [1.0, 1.0])
    h = h / np.sum(h) * np.sqrt(2.0)
    # extremal phase, energy at the start (MATLAB Lo_R ordering)
    if np.abs(h[0]) < np.abs(h[-1]):
        h = h[::-1]
    return h


def filters(name):
    """(lo_r, hi_r) reconstruction filters of a basis 'dbN'."""
    N = int(name[2:])
    lo = daubechies_lo_r(N)
    lf = len(lo)
    hi = np.array([(-1.0) ** k * lo[lf - 1 - k] for k in range(lf)])
    return lo, hi


def n_out(n, lf):
    return (n + lf - 1) // 2


def analysis_1d(x, f, axis):
    """One zpd analysis step along ``axis`` with reconstruction filter f:
    ``out[k] = sum_j f[j] x[2k + 2 - lf + j]`` (x zero outside [0, n))."""
    x = np.moveaxis(np.asarray(x, dtype=np.float64), axis, 0)
    n = x.shape[0]
    lf = len(f)
    m = n_out(n, lf)
    xp = np.zeros((n + 2 * lf,) + x.shape[1:])
    xp[lf - 2:lf - 2 + n] = x            # xp[i + lf - 2] = x[i]
    out = np.zeros((m,) + x.shape[1:])
    for j in range(lf):
        out += f[j] * xp[j:j + 2 * m:2]
    return np.moveaxis(out, 0, axis)


def synthesis_1d(c, f, n, axis):
    """Adjoint of analysis_1d: length-n output along ``axis``."""
    c = np.moveaxis(np.asarray(c, dtype=np.float64), axis, 0)
    m = c.shape[0]
    lf = len(f)
    op = np.zeros((2 * m + lf,) + c.shape[1:])
    for j in range(lf):
        op[j:j + 2 * m:2] += f[j] * c
    out = np.zeros((n,) + c.shape[1:])
    hi = min(n, op.shape[0] - (lf - 2))
    out[:hi] = op[lf - 2:lf - 2 + hi]
    return np.moveaxis(out, 0, axis)


def level_shapes(Ny, Nx, lf, nlevel):
    shapes = [(Ny, Nx)]
    for _ in range(nlevel):
        ny, nx = shapes[-1]
        shapes.append((n_out(ny, lf), n_out(nx, lf)))
    return shapes


def n_coefs(Ny, Nx, lf, nlevel):
    sh = level_shapes(Ny, Nx, lf, nlevel)
    return sh[-1][0] * sh[-1][1] + 3 * sum(a * b for a, b in sh[1:])


def wavedec2(x, name, nlevel):
    lo, hi = filters(name)
    a = np.asarray(x, dtype=np.float64)
    details = []
    for _ in range(nlevel):
        lx = analysis_1d(a, lo, 1)
        hx = analysis_1d(a, hi, 1)
        h = analysis_1d(lx, hi, 0)
        v = analysis_1d(hx, lo, 0)
        d = analysis_1d(hx, hi, 0)
        a = analysis_1d(lx, lo, 0)
        details.append((h, v, d))
    parts = [a.ravel(order='F')]
    for h, v, d in reversed(details):
        parts += [h.ravel(order='F'), v.ravel(order='F'), d.ravel(order='F')]
    return np.concatenate(parts)


def waverec2_adjoint(c, name, nlevel, Ny, Nx):
    lo, hi = filters(name)
    lf = len(lo)
    shapes = level_shapes(Ny, Nx, lf, nlevel)
    pos = 0
    ny, nx = shapes[-1]
    a = c[pos:pos + ny * nx].reshape((ny, nx), order='F'); pos += ny * nx
    for lev in range(nlevel, 0, -1):
        ny, nx = shapes[lev]
        py, px = shapes[lev - 1]
        sz = ny * nx
        h = c[pos:pos + sz].reshape((ny, nx), order='F'); pos += sz
        v = c[pos:pos + sz].reshape((ny, nx), order='F'); pos += sz
        d = c[pos:pos + sz].reshape((ny, nx), order='F'); pos += sz
        lx = synthesis_1d(a, lo, py, 0) + synthesis_1d(h, hi, py, 0)
        hx = synthesis_1d(v, lo, py, 0) + synthesis_1d(d, hi, py, 0)
        a = synthesis_1d(lx, lo, px, 1) + synthesis_1d(hx, hi, px, 1)
    return a


class SaraDictionary:
    """Psit / Psi of the SARA dictionary for one Ny x Nx channel."""

    def __init__(self, Ny, Nx, bases=DEFAULT_BASES, nlevel=DEFAULT_NLEVEL):
        self.Ny, self.Nx = Ny, Nx
        self.bases = tuple(bases)
        self.nlevel = nlevel
        self.P = len(self.bases)
        self.sizes = []
        for b in self.bases:
            if b == 'self':
                self.sizes.append(Ny * Nx)
            else:
                self.sizes.append(n_coefs(Ny, Nx, len(filters(b)[0]), nlevel))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.s = int(self.offsets[-1])

    def psit(self, x):
        parts = []
        for b in self.bases:
            if b == 'self':
                parts.append(np.asarray(x, dtype=np.float64).ravel(order='F'))
            else:
                parts.append(wavedec2(x, b, self.nlevel))
        return np.concatenate(parts) / np.sqrt(self.P)

    def psi(self, c):
        out = np.zeros((self.Ny, self.Nx))
        for i, b in enumerate(self.bases):
            ci = c[self.offsets[i]:self.offsets[i + 1]]
            if b == 'self':
                out += ci.reshape((self.Ny, self.Nx), order='F')
            else:
                out += waverec2_adjoint(ci, b, self.nlevel, self.Ny, self.Nx)
        return out / np.sqrt(self.P)


# ---- faceted SARA transform -------------------------------------------------
# The reference splits the analysis and synthesis over non-overlapping facets
# (sdwt2_setup / sdwt2_sara_faceting / isdwt2_sara_faceting of its faceted
# wavelet submodule, absent from the reference tree; call sites
# src/fhs/facetHyperSARA.m:325-326, src/fhs/fhs_update_dual_sparsity.m:71-93,
# src/fhs/fhs_initialize_dual_and_weights.m:80-103).  The facet transforms
# are exact pieces of the global one: every global coefficient belongs to
# one facet, and a facet computes its coefficients from its own pixels plus a
# top/left extension that the neighbouring facets provide.  The ownership
# rule below is this project's (the submodule's own layout is not in the
# tree): at level j a facet owns the coefficients k whose last input sample
# 2k + 1 lies in the range it owns at level j - 1 (the level-0 ranges being
# its pixels), the right-border coefficients of the zpd extension going to
# the last facet.  So the level-j boundaries are t = ceil((s - 1) / 2) of
# the level-(j - 1) boundaries s.

def facet_level_bounds(N, starts, lf, nlevel):
    """Boundaries of every level's owned ranges along one axis: bounds[j][i]
    .. bounds[j][i + 1] is facet i's range at level j (level 0 = pixels);
    ``starts``: ascending facet origins, starts[0] = 0."""
    bounds = [list(starts) + [N]]
    n = N
    for _ in range(nlevel):
        n = n_out(n, lf)
        prev = bounds[-1]
        bounds.append([0] + [s // 2 for s in prev[1:-1]] + [n])   # ceil((s - 1) / 2) = floor(s / 2)
    return bounds


def facet_extension_start(t_J, lf, nlevel):
    """First input pixel the owned coefficients of a facet need, from its
    first owned level-J coefficient: m_(j-1) = 2 m_j + 2 - lf; rounded down
    to a multiple of 2^nlevel so that the extended image decimates in phase
    with the whole image."""
    m = t_J
    for _ in range(nlevel):
        m = 2 * m + 2 - lf
    step = 1 << nlevel
    return (m // step) * step


class SaraFacets:
    """Faceted Psit / Psi: facet (qy, qx) = q = qx * Qy + qy owns pixels
    [ys[qy], ys[qy+1]) x [xs[qx], xs[qx+1]) and reads the extended image
    [ey0, ys[qy+1]) x [ex0, xs[qx+1]) (zero where the origin is negative).

    Facet coefficient vector: per basis the owned tiles of A_J, then H_j,
    V_j, D_j for j = J .. 1 (each column-major), the Dirac part last (the
    owned pixels); scaled by 1/sqrt(P) like ``SaraDictionary``."""

    def __init__(self, Ny, Nx, ys, xs, bases=DEFAULT_BASES, nlevel=DEFAULT_NLEVEL):
        self.Ny, self.Nx, self.nlevel = Ny, Nx, nlevel
        self.bases = tuple(bases)
        self.ys, self.xs = list(ys), list(xs)
        self.Qy, self.Qx = len(ys) - 1, len(xs) - 1
        self.glob = SaraDictionary(Ny, Nx, bases, nlevel)
        lfs = [len(filters(b)[0]) for b in self.bases if b != 'self']
        by = {lf: facet_level_bounds(Ny, self.ys[:-1], lf, nlevel) for lf in set(lfs)}
        bx = {lf: facet_level_bounds(Nx, self.xs[:-1], lf, nlevel) for lf in set(lfs)}
        self.facets = []
        for qx in range(self.Qx):
            for qy in range(self.Qy):
                ey0 = min([facet_extension_start(by[lf][nlevel][qy], lf, nlevel) for lf in lfs] + [self.ys[qy]])
                ex0 = min([facet_extension_start(bx[lf][nlevel][qx], lf, nlevel) for lf in lfs] + [self.xs[qx]])
                eny, enx = self.ys[qy + 1] - ey0, self.xs[qx + 1] - ex0
                sub = SaraDictionary(eny, enx, bases, nlevel)
                loc, glb = [], []
                for i, b in enumerate(self.bases):
                    if b == 'self':
                        r = np.arange(self.ys[qy], self.ys[qy + 1])
                        c = np.arange(self.xs[qx], self.xs[qx + 1])
                        loc.append(sub.offsets[i] + ((r - ey0)[:, None] + eny * (c - ex0)[None, :]).ravel(order='F'))
                        glb.append(self.glob.offsets[i] + (r[:, None] + Ny * c[None, :]).ravel(order='F'))
                        continue
                    lf = len(filters(b)[0])
                    ssh = level_shapes(eny, enx, lf, nlevel)
                    gsh = level_shapes(Ny, Nx, lf, nlevel)
                    # subband order inside a basis: A_J, then (H, V, D)_j for j = J .. 1
                    order = [('a', nlevel)] + [(sb, j) for j in range(nlevel, 0, -1) for sb in 'hvd']
                    so, go = 0, 0
                    for sb, j in order:
                        sny, snx = ssh[j]
                        gny, gnx = gsh[j]
                        r = np.arange(by[lf][j][qy], by[lf][j][qy + 1])
                        c = np.arange(bx[lf][j][qx], bx[lf][j][qx + 1])
                        dy, dx = ey0 >> j, ex0 >> j
                        loc.append(sub.offsets[i] + so + ((r - dy)[:, None] + sny * (c - dx)[None, :]).ravel(order='F'))
                        glb.append(self.glob.offsets[i] + go + (r[:, None] + gny * c[None, :]).ravel(order='F'))
                        so += sny * snx
                        go += gny * gnx
                loc = np.concatenate(loc).astype(np.int64)
                glb = np.concatenate(glb).astype(np.int64)
                self.facets.append({'ey0': ey0, 'ex0': ex0, 'eny': eny, 'enx': enx, 'sub': sub,
                                    'local': loc, 'global': glb})

    def extended(self, q, x):
        """The extended image of facet q cut from a whole image x (zero padded)."""
        f = self.facets[q]
        out = np.zeros((f['eny'], f['enx']))
        y0, x0 = max(f['ey0'], 0), max(f['ex0'], 0)
        y1, x1 = f['ey0'] + f['eny'], f['ex0'] + f['enx']
        out[y0 - f['ey0']:, x0 - f['ex0']:] = x[y0:y1, x0:x1]
        return out

    def analysis(self, q, x_ext):
        """sdwt2_sara_faceting: the facet's coefficients from its extended image."""
        f = self.facets[q]
        return f['sub'].psit(x_ext)[f['local']]

    def synthesis(self, q, z):
        """isdwt2_sara_faceting: the facet's contribution (extended region)."""
        f = self.facets[q]
        c = np.zeros(f['sub'].s)
        c[f['local']] = z
        return f['sub'].psi(c)

    def add_extended(self, q, g, out):
        """out += the facet's extended-region image g, cropped to the image."""
        f = self.facets[q]
        y0, x0 = max(f['ey0'], 0), max(f['ex0'], 0)
        out[y0:f['ey0'] + f['eny'], x0:f['ex0'] + f['enx']] += g[y0 - f['ey0']:, x0 - f['ex0']:]
