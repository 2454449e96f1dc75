"""Seeded synthetic problems (sky, spectral cube, uv coverage, noisy data).

Semantics follow the reference's synthetic-data path:
* uv: Gaussian coverage, rejection to (-pi, pi)^2
  (``lib/generate_data/util_gen_sampling_pattern.m:54-87``), per channel
  scaled by nu_l / nu_max (``lib/operators/util_gen_measurement_operator_synth.m:67-69``);
* cube: curved power law x_l = x0 (nu_l/nu0)^(alpha + beta log(nu_l/nu0))
  (``imaging/main_generate_cyga_cubes.m:147-150``);
* noise: sigma_l = 10^(-SNR/20) ||y0_l|| / sqrt(M_l), complex Gaussian,
  whitening G <- G/sigma, y <- y/sigma
  (``lib/generate_data/util_gen_measurements_snr_new.m:40-81``);
* l2 bounds: eps_l = sqrt(M_l + 2 sqrt(M_l)) on whitened data, split over
  blocks as sqrt(m_b/M_l) eps_l
  (``lib/utils/util_gen_data_fidelity_bounds2.m:23-36``, sigma_ball = 2,
  ``imaging/main_generate_data.m:297``).
The sky model (elliptical Gaussian sources + k^-3 Gaussian random field) has
no reference generator; it is this project's own.
"""
from dataclasses import dataclass, field

import numpy as np


def gaussian_uv(rng, M, sigma=np.pi / 4):
    u = np.empty(0)
    v = np.empty(0)
    while len(u) < M:
        n = 2 * (M - len(u)) + 16
        uu = sigma * rng.standard_normal(n)
        vv = sigma * rng.standard_normal(n)
        keep = (np.abs(uu) < np.pi) & (np.abs(vv) < np.pi)
        u = np.concatenate([u, uu[keep]])
        v = np.concatenate([v, vv[keep]])
    return u[:M], v[:M]


def sky_image(rng, Ny, Nx, n_sources=200, grf_weight=0.3):
    """Elliptical Gaussians (log-normal fluxes mu=0, sigma=1, axes U[1,8] px,
    angle U[0, pi)) + a diffuse P(k) ~ k^-3 Gaussian random field clipped at 0;
    peak-normalised to 1."""
    yy, xx = np.mgrid[0:Ny, 0:Nx].astype(np.float64)
    src = np.zeros((Ny, Nx))
    cy = rng.uniform(0, Ny, n_sources)
    cx = rng.uniform(0, Nx, n_sources)
    a = rng.uniform(1, 8, n_sources)
    b = rng.uniform(1, 8, n_sources)
    th = rng.uniform(0, np.pi, n_sources)
    flux = rng.lognormal(0.0, 1.0, n_sources)
    for i in range(n_sources):
        r = max(a[i], b[i]) * 5
        y0, y1 = int(max(cy[i] - r, 0)), int(min(cy[i] + r + 1, Ny))
        x0, x1 = int(max(cx[i] - r, 0)), int(min(cx[i] + r + 1, Nx))
        dy = yy[y0:y1, x0:x1] - cy[i]
        dx = xx[y0:y1, x0:x1] - cx[i]
        c, s = np.cos(th[i]), np.sin(th[i])
        p = c * dx + s * dy
        q = -s * dx + c * dy
        src[y0:y1, x0:x1] += flux[i] * np.exp(-0.5 * ((p / a[i]) ** 2 + (q / b[i]) ** 2))
    src /= src.max()
    ky = np.fft.fftfreq(Ny)[:, None]
    kx = np.fft.fftfreq(Nx)[None, :]
    k = np.sqrt(ky ** 2 + kx ** 2)
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** -1.5
    noise = rng.standard_normal((Ny, Nx)) + 1j * rng.standard_normal((Ny, Nx))
    grf = np.real(np.fft.ifft2(noise * amp))
    grf = np.maximum(grf, 0.0)
    if grf.max() > 0:
        grf *= grf_weight / grf.max()
    img = src + grf
    return img / img.max()


def spectral_cube(rng, x0, L, dnu=0.05, nu=None):
    nu = 1.0 + dnu * np.arange(L) if nu is None else np.asarray(nu, dtype=np.float64)
    alpha = rng.normal(-0.7, 0.2, x0.shape)
    beta = rng.normal(0.0, 0.05, x0.shape)
    ln = np.log(nu)
    return np.stack([x0 * np.exp((alpha + beta * ln[l]) * ln[l]) for l in range(L)], axis=2), nu


@dataclass
class Problem:
    """Whitened synthetic problem (MATLAB layouts; x is (Ny, Nx, L))."""
    Ny: int
    Nx: int
    L: int
    X0: np.ndarray
    u: list
    v: list
    blocks: list             # per channel: list of block sizes
    nu: np.ndarray
    seed: int
    input_snr: float
    w: list = field(default_factory=list)          # per channel w (wavelengths), empty: no w-term
    fov: float = 0.0                               # field of view (direction cosines) of the w-term
    y: list = field(default_factory=list)          # whitened visibilities per channel
    sigma_noise: list = field(default_factory=list)
    epsilon: list = field(default_factory=list)    # per channel: list of block bounds
    datasets: list = field(default_factory=list)   # per channel: rows of each data set (in order)

    @property
    def M(self):
        return [len(a) for a in self.u]


def make_problem(Ny, Nx, L, M_base, nblocks, seed=0, flag_range=(0.1, 0.3), input_snr=40.0,
                 dnu=0.05, n_sources=200, w_max=None, fov=0.08, nu=None, nu_max=None, n_datasets=1):
    """Seeded problem.  ``w_max``: base w uniform in [0, w_max] wavelengths
    (scaled per channel with u and v) and field of view ``fov`` for the
    w-term; ``nu``: the channels' frequencies (default 1 + dnu l) and
    ``nu_max`` the frequency the uv coverage is scaled against (default
    max(nu)) -- one spectral sub-cube of a larger cube takes its share of
    the cube's frequencies and the cube's nu_max, and the same seed gives
    every sub-cube the same sky, spectral indices and base coverage.
    ``n_datasets``: data sets per channel (the reference's ``nDataSets``),
    each its own base coverage of M_base points, scaled and flagged per
    channel and split into ``nblocks`` blocks; a channel's rows are its sets
    one after the other (``Problem.datasets``: rows per set)."""
    rng = np.random.default_rng(seed)
    x0 = sky_image(rng, Ny, Nx, n_sources=n_sources)
    if L == 1:
        X0, nu_ = x0[:, :, None], np.ones(1)
    else:
        X0, nu_ = spectral_cube(rng, x0, L, dnu, nu)
    nu = nu_
    ub, vb = gaussian_uv(rng, M_base)
    wb = rng.uniform(0.0, w_max, M_base) if w_max else None
    sets = [(ub, vb, wb)]
    for _ in range(n_datasets - 1):
        u2, v2 = gaussian_uv(rng, M_base)
        sets.append((u2, v2, rng.uniform(0.0, w_max, M_base) if w_max else None))
    us, vs, ws, blocks, datasets = [], [], [], [], []
    for l in range(L):
        scale = nu[l] / (nu.max() if nu_max is None else nu_max)
        ul, vl, wl, bl, dl = [], [], [], [], []
        for ub, vb, wb in sets:
            if L > 1 and flag_range is not None:
                frac = rng.uniform(*flag_range)
                keep = np.sort(rng.permutation(M_base)[:int(round(M_base * (1 - frac)))])
            else:
                keep = np.arange(M_base)
            ul.append(ub[keep] * scale)
            vl.append(vb[keep] * scale)
            if wb is not None:
                wl.append(wb[keep] * scale)
            b = np.linspace(0, len(keep), nblocks + 1).astype(np.int64)
            bl += [int(b[i + 1] - b[i]) for i in range(nblocks)]
            dl.append(len(keep))
        us.append(np.concatenate(ul))
        vs.append(np.concatenate(vl))
        if wl:
            ws.append(np.concatenate(wl))
        blocks.append(bl)
        datasets.append(dl)
    return Problem(Ny, Nx, L, X0, us, vs, blocks, nu, seed, input_snr, w=ws, fov=fov if ws else 0.0,
                   datasets=datasets)


def host_visibilities(problem):
    """Noiseless visibilities G_l A x0_l of every channel, computed in FP64 on
    the host (the min-max coefficients of ``fhs.nufft.Nufft2D`` and numpy's
    FFT), so a problem's data do not depend on the device that builds it."""
    from .nufft import Nufft2D
    P = problem
    if getattr(P, 'w', None):
        raise NotImplementedError('host_visibilities: no w-term on the host path (use the GPU data path)')
    nf = Nufft2D(P.Ny, P.Nx)
    Kty, Ktx = nf.Kty, nf.Ktx
    sn = np.outer(nf.ay.s, nf.ax.s)
    out = []
    for l in range(P.L):
        pad = np.zeros((Kty, Ktx), dtype=np.complex128)
        pad[:P.Ny, :P.Nx] = sn * P.X0[:, :, l]
        F = np.fft.fft2(pad)
        a, cy, cx = nf.separable(P.u[l], P.v[l])
        iy = (a % Kty)[:, None] + np.arange(nf.Ky)[None, :]
        ix = (a // Kty)[:, None] + np.arange(nf.Kx)[None, :]
        iy %= Kty
        ix %= Ktx
        y = np.zeros(len(a), dtype=np.complex128)
        for jx in range(nf.Kx):
            y += cx[:, jx] * np.einsum('mj,mj->m', cy, F[iy, ix[:, jx:jx + 1]])
        out.append(y)
    return out


def add_host_data(problem):
    """Noisy whitened data from ``host_visibilities`` (``add_noise``)."""
    return add_noise(problem, host_visibilities(problem))


def add_noise(problem, y0):
    """Noisy whitened data from noiseless visibilities y0 (list per channel).
    Returns per-channel sigma (G must be divided by it)."""
    rng = np.random.default_rng(problem.seed + 7919)
    problem.y, problem.sigma_noise, problem.epsilon = [], [], []
    for l, yl in enumerate(y0):
        M = len(yl)
        sigma = 10 ** (-problem.input_snr / 20) * np.linalg.norm(yl) / np.sqrt(M)
        n = (rng.standard_normal(M) + 1j * rng.standard_normal(M)) / np.sqrt(2)
        problem.y.append((yl + sigma * n) / sigma)
        problem.sigma_noise.append(sigma)
        eps_l = np.sqrt(M + 2 * np.sqrt(M))
        problem.epsilon.append([np.sqrt(m / M) * eps_l for m in problem.blocks[l]])
    return problem.sigma_noise
