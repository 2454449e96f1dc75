"""Oracle: w-projection rows of the measurement operator G, float64.

Test infrastructure only (see ``oracle/__init__.py``).

Reference semantics: ``op_p_nufft_wproj_dde(param_nufft, [{v} {u}], {w},
{nW}, param_wproj)`` (``lib/operators/util_load_data_gen_measurement_operator.m:119-121``;
submodule RI-measurement-operator, un-vendored: parity unpinned, semantics
restated) returns G whose row m holds the NUFFT interpolation row
convolved with the Fourier kernel of visibility m's w-term, both kernels
sparsified by energy: the w-kernel keeps its largest entries up to a
fraction ``CEnergyL2`` of its energy, the combined row up to ``GEnergyL2``
(``imaging/default_parameters.json:58-61``, 0.9999 each; 1 disables).  The
field of view enters through ``FoVx = sin(pixelSize Nx)``
(``imaging/imaging.m:427-431``); w is in wavelengths.

Measurement equation with the w-term (visibility m, pixel n = (ny, nx),
direction cosines l = (nx - Nx/2) FoV/Nx, m = (ny - Ny/2) FoV/Ny):
    y_m = sum_n x_n c_m(n) exp(-i (u_m (nx - Nx/2) + v_m (ny - Ny/2))),
    c_m(n) = exp(-2 pi i w_m (sqrt(1 - l^2 - m^2) - 1))          (exact)
           ~ exp(i pi w_m (l^2 + m^2))                             (Fresnel).
The Fresnel chirp is separable, c = c_y(m) c_x(l), so its Fourier kernel is
the outer product of two 1-D kernels.  With A x = FFT_K(pad(sn .* x)) and
the convolution theorem on the oversampled grid,
    G'_m[j] = (1/K) sum_k G_m[k] Chat_m[k - j]   per axis,
Chat = FFT_K(c_ext), where c_ext equals the chirp on the image support and
is free elsewhere (A's input is zero there).  Filling the padding with a
smooth periodic phase bridge (C^2 at both edges) makes Chat compact; the
zero-padded chirp would have sinc tails.  With both energy fractions 1 the
rows reproduce the NUDFT with the Fresnel phase to NUFFT accuracy; the
Fresnel form differs from the exact phase by O(w (l^2 + m^2)^2).

Pinned in ``tests/test_oracle_wproj.py`` against the direct NUDFT with the
w-phase (exact and Fresnel) on small images.
"""
import numpy as np

from .nufft import fit_scaling, minmax_coef_1d, minmax_gram_inverse  # noqa: F401


def bridge_phase(t, w, F):
    """Phase of the extended chirp on the oversampled axis, period 2F:
    pi w t^2 on the image support |t| <= F/2, and on the padding an even
    polynomial bridge q(u) = pi w (2.5 h^2 - 2 u^2 + u^4 / (2 h^2)),
    u = t - F (mod 2F) in [-h, h], h = F/2, which matches the chirp's value,
    slope and curvature at both ends of the padding."""
    t = np.asarray(t, dtype=np.float64)
    h = F / 2
    tt = np.mod(t + F, 2 * F) - F            # [-F, F)
    inside = np.abs(tt) <= h
    u = np.where(tt > 0, tt - F, tt + F)     # distance from the padding's centre (t = +-F)
    q = np.pi * w * (2.5 * h * h - 2 * u * u + u ** 4 / (2 * h * h))
    return np.where(inside, np.pi * w * tt * tt, q)


def axis_positions(N, K, F):
    """Direction cosine of each point of the K-point oversampled axis: the
    image occupies points [0, N) (pixel n at (n - N/2) F/N), the padding
    [N, K) continues periodically."""
    n = np.arange(K)
    t = (n - N // 2) * (F / N)
    return t


def chirp_kernel_1d(w, N, K, F):
    """Chat[d], d = 0..K-1 (circular offsets), of the extended Fresnel
    chirp exp(i bridge_phase) along one axis: FFT_K of c_ext."""
    c = np.exp(1j * bridge_phase(axis_positions(N, K, F), w, F))
    return np.fft.fft(c)


def energy_mask(v, frac):
    """Boolean mask of the largest |v| entries holding at least ``frac`` of
    the energy (all entries for frac >= 1)."""
    v = np.asarray(v)
    if frac >= 1.0:
        return np.ones(v.shape, dtype=bool)
    e = np.abs(v) ** 2
    order = np.argsort(-e, kind='stable')
    cum = np.cumsum(e[order])
    k = int(np.searchsorted(cum, frac * cum[-1])) + 1
    mask = np.zeros(v.shape, dtype=bool)
    mask[order[:k]] = True
    return mask


def wproj_axis_rows(idx, coef, w, N, K, F, c_energy=1.0, g_energy=1.0, window=None):
    """W-projected 1-D rows of one axis.

    idx (M, J): grid indices of the min-max row (mod K), coef (M, J); w (M,).
    Returns a list of (indices, values) per row: the correlation of the
    row with the w-kernel (kernel truncated to ``c_energy`` of its energy,
    then to offsets |d| <= window when given), truncated to ``g_energy`` of
    the row's energy."""
    out = []
    d = np.arange(K)
    d = np.where(d > K // 2, d - K, d)
    for m in range(len(w)):
        Ch = chirp_kernel_1d(w[m], N, K, F)
        Ch = np.where(energy_mask(Ch, c_energy), Ch, 0.0)
        if window is not None:
            Ch = np.where(np.abs(d) <= window, Ch, 0.0)
        row = np.zeros(K, dtype=np.complex128)
        row[idx[m]] += coef[m]
        # G'[j] = (1/K) sum_k G[k] Chat[(k - j) mod K]: circular correlation
        rp = np.fft.ifft(np.fft.fft(row) * np.conj(np.fft.fft(np.conj(Ch)))) / K
        # (fft identity: sum_k a[k] b[k - j] = ifft(fft(a) * conj(fft(conj(b))))[j])
        keep = energy_mask(rp, g_energy) & (np.abs(rp) > 0)
        j = np.nonzero(keep)[0]
        out.append((j, rp[j]))
    return out


def wproj_matrix(u, v, w, Ny, Nx, F, c_energy=1.0, g_energy=1.0, Ky=7, Kx=7, oy=2, ox=2, window=None):
    """Sparse G (scipy csr, M x Kty*Ktx, column ky + Kty kx) of the
    Fresnel w-projected min-max NUFFT, with the per-axis energy fractions
    sqrt(c_energy), sqrt(g_energy) (the energies of a separable row
    multiply).  Returns (G, rows_y, rows_x)."""
    import scipy.sparse as sp
    Kty, Ktx = oy * Ny, ox * Nx
    ay, ax = fit_scaling(Ny, Ky, Kty), fit_scaling(Nx, Kx, Ktx)
    iy, cy = minmax_coef_1d(v, Ny, Ky, Kty, ay, Ny // 2)
    ix, cx = minmax_coef_1d(u, Nx, Kx, Ktx, ax, Nx // 2)
    ce, ge = np.sqrt(c_energy), np.sqrt(g_energy)
    ry = wproj_axis_rows(iy, cy, w, Ny, Kty, F, ce, ge, window)
    rx = wproj_axis_rows(ix, cx, w, Nx, Ktx, F, ce, ge, window)
    rows, cols, vals = [], [], []
    for m in range(len(w)):
        jy, vy = ry[m]
        jx, vx = rx[m]
        c = (jy[:, None] + Kty * jx[None, :]).ravel()
        rows.append(np.full(c.size, m))
        cols.append(c)
        vals.append((vy[:, None] * vx[None, :]).ravel())
    G = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(len(w), Kty * Ktx))
    return G, ry, rx


def nudft_w(x, u, v, w, F, exact=True):
    """Direct sum with the w-phase (the quantity the w-projected G A x
    approximates); x (Ny, Nx), phase centre (Ny/2, Nx/2)."""
    Ny, Nx = x.shape
    ny = np.arange(Ny) - Ny // 2
    nx = np.arange(Nx) - Nx // 2
    lm = ny[:, None] * (F / Ny)
    ll = nx[None, :] * (F / Nx)
    out = np.empty(len(u), dtype=np.complex128)
    for k in range(len(u)):
        if exact:
            c = np.exp(-2j * np.pi * w[k] * (np.sqrt(1 - ll ** 2 - lm ** 2) - 1))
        else:
            c = np.exp(1j * np.pi * w[k] * (ll ** 2 + lm ** 2))
        ph = np.exp(-1j * (v[k] * ny[:, None] + u[k] * nx[None, :]))
        out[k] = np.sum(x * c * ph)
    return out
