"""Oracle: scalar heuristics around the PDFB iteration, float64.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
from math import gamma as gamma_fn, sqrt


def compute_noise_level_hs(Ny, Nx, n_channels, std_noise, squared_operator_norm):
    """``src/heuristics/compute_noise_level.m:49-56`` ('hs' branch).

    Returns (sig, sig_bar, mu_chi, sig_chi, sig_sara)."""
    std_noise = np.asarray(std_noise, dtype=np.float64)
    son = np.asarray(squared_operator_norm, dtype=np.float64)
    sig_sara = sqrt(np.mean(std_noise ** 2 / son))
    mu_chi = sqrt(2) * gamma_fn((n_channels + 1) / 2) / gamma_fn(n_channels / 2)
    sig_chi = sqrt(n_channels - mu_chi ** 2)
    sig = sig_sara * (mu_chi + sig_chi)
    sig_bar = sqrt(Nx * Ny * np.sum(std_noise ** 2 / son) / min(Nx * Ny, n_channels))
    return sig, sig_bar, mu_chi, sig_chi, sig_sara


def compute_noise_level_sara(std_noise, squared_operator_norm):
    """``src/heuristics/compute_noise_level_sara.m:21``."""
    return sqrt(std_noise ** 2 / squared_operator_norm)


def data_fidelity_bounds(block_sizes, sigma_ball=2.0, sigma_noise=1.0):
    """``lib/utils/util_gen_data_fidelity_bounds2.m:23-36`` ('sigma' type).

    Channel bound ``sqrt(Ml + s1 sqrt(Ml)) sigma``, split per block as
    ``sqrt(m/Ml) * epsilon``."""
    Ml = float(np.sum(block_sizes))
    eps_ch = sqrt(Ml + sigma_ball * sqrt(Ml)) * sigma_noise
    return [sqrt(m / Ml) * eps_ch for m in block_sizes]


def op_norm(F, Ft, shape, tol, max_iter, seed=0):
    """Power iteration for ||F||^2 (op_norm of the RI-measurement-operator
    submodule, called at ``lib/utils/util_operator_norm.m:46,55``).
    Restated: x <- Ft(F(x))/||.||, stop when the relative change of the
    estimate is below tol.  Returns (value, rel_var)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    x /= np.linalg.norm(x)
    val = 1.0
    rel = 1.0
    for _ in range(max_iter):
        xn = Ft(F(x))
        newval = np.linalg.norm(xn)
        rel = abs(newval - val) / newval
        x = xn / newval
        val = newval
        if rel < tol:
            break
    return val, rel
