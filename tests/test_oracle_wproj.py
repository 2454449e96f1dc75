"""Pins of the w-projection oracle (oracle/wproj.py) and of the product's
host-side w-kernel table (fhs.nufft.wkernel_table), CPU only.

The reference's w-projection (op_p_nufft_wproj_dde, un-vendored submodule)
has no golden vectors; the oracle is pinned by what the mathematics fixes:
* w = 0 reduces the rows to the min-max rows exactly;
* the extended chirp equals the chirp on the image and is C^2 at its edges;
* with no energy truncation the w-projected G A x equals the direct sum with
  the Fresnel w-phase to NUFFT accuracy (the construction is exact);
* with the reference's 0.9999 energy fractions the error stays at the
  truncation level;
* the Fresnel phase differs from the exact phase by at most its analytic
  bound at the field of view of the 2048x2048 workload (0.08 rad).
"""
import numpy as np
import pytest

from oracle.nufft import NufftOracle
from oracle.wproj import (bridge_phase, chirp_kernel_1d, nudft_w, wproj_matrix)


def _case(N=32, M=40, wmax=200.0, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-np.pi, np.pi, M)
    v = rng.uniform(-np.pi, np.pi, M)
    w = rng.uniform(0, wmax, M)
    x = rng.random((N, N))
    return u, v, w, x, NufftOracle(u, v, N, N)


def test_w_zero_gives_minmax_rows():
    u, v, w, x, op = _case()
    G, _, _ = wproj_matrix(u, v, np.zeros_like(w), 32, 32, 0.1)
    assert abs(G - op.G).max() < 1e-12


def test_extended_chirp_is_chirp_on_image_and_c2_at_edges():
    F, w = 0.1, 137.0
    t = np.linspace(-F / 2, F / 2, 101)
    assert np.allclose(bridge_phase(t, w, F), np.pi * w * t * t, rtol=0, atol=1e-12)
    h = 1e-6
    for e in (F / 2, -F / 2):
        s = np.sign(e)
        inner = np.array([e - 2 * s * h, e - s * h, e])
        outer = np.array([e, e + s * h, e + 2 * s * h])
        pi_, po = bridge_phase(inner, w, F), bridge_phase(outer, w, F)
        assert abs(pi_[2] - po[0]) < 1e-12                                   # value
        d_in, d_out = (pi_[2] - pi_[1]) / (s * h), (po[1] - po[0]) / (s * h)
        assert abs(d_in - d_out) < 1e-3 * abs(np.pi * w * F)                 # slope
        c_in = (pi_[2] - 2 * pi_[1] + pi_[0]) / h ** 2
        c_out = (po[2] - 2 * po[1] + po[0]) / h ** 2
        assert abs(c_in - c_out) < 1e-2 * 2 * np.pi * w                      # curvature
    # periodic with period 2F
    assert np.allclose(bridge_phase(t + 2 * F, w, F), bridge_phase(t, w, F), atol=1e-9)


def test_untruncated_rows_are_exact_fresnel():
    u, v, w, x, op = _case()
    G, _, _ = wproj_matrix(u, v, w, 32, 32, 0.3)
    y, yr = G @ op.A(x), nudft_w(x, u, v, w, 0.3, exact=False)
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-4


def test_energy_truncated_rows_error_at_truncation_level():
    u, v, w, x, op = _case()
    G, _, _ = wproj_matrix(u, v, w, 32, 32, 0.3, 0.9999, 0.9999)
    y, yr = G @ op.A(x), nudft_w(x, u, v, w, 0.3, exact=False)
    e = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    assert 1e-4 < e < 2e-2, e


def test_fresnel_phase_bound_at_workload_fov():
    """|2 pi w (sqrt(1 - r^2) - 1 + r^2 / 2)| <= 2 pi w r^4 / 8 / (1 - r^2)
    with r^2 <= F^2 / 2: at F = 0.08 and w <= 200 the two sums differ by at
    most that phase (relative)."""
    F, wmax = 0.08, 200.0
    u, v, w, x, _ = _case(N=32, M=30, wmax=wmax, seed=3)
    ye, yf = nudft_w(x, u, v, w, F, exact=True), nudft_w(x, u, v, w, F, exact=False)
    r2 = F * F / 2
    bound = 2 * np.pi * wmax * r2 * r2 / 8 / (1 - r2)
    err = np.max(np.abs(ye - yf) / np.sum(np.abs(x)))
    assert err <= bound, (err, bound)
    assert bound < 2e-3


@pytest.mark.parametrize('N,K,F,S', [(32, 64, 0.3, 6), (256, 512, 0.08, 4)])
def test_product_table_matches_oracle_kernel(N, K, F, S):
    """The product's host table (its own extended chirp) against the
    oracle's kernel, at the table's w values, with and without CEnergyL2."""
    from fhs.nufft import wkernel_table
    from oracle.wproj import energy_mask
    for keep in (1.0, np.sqrt(0.9999)):
        tab, dw, dmax = wkernel_table(N, K, F, 10.0, S, keep, dw=1.0)
        d = np.arange(-S, S + 1) % K
        for b in (0, 3, 7, 10):
            Ch = chirp_kernel_1d(b * dw, N, K, F)
            Ch = np.where(energy_mask(Ch, keep), Ch, 0.0)
            assert np.allclose(tab[b], Ch[d], rtol=0, atol=1e-9 * K), (N, b, keep)
