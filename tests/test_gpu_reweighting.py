"""The reweighting step on the GPU against the oracle: the l21 weights
upsilon ./ (upsilon + ||Psit x||_row) (src/fhs/fhs_update_weights.m:82-92,
src/hs/hs_update_weights_sparsity_distributed.m) and the nuclear-norm
weights upsilon_bar ./ (upsilon_bar + sigma(w_q .* x_q)) per facet
(src/fhs/fhs_update_weights.m:73-80, src/hs/hs_update_weights_lowrankness_serial.m),
computed by the solver driver's reweighting (fhs.solvers._reweight) at the
end of a PDFB cycle."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle.facets import facet_geometry, facet_singular_values
from oracle.solvers import lowrank_weights, sparsity_weights
from oracle.wavelets import SaraDictionary as SaraRef

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def _cube(rng, Ny, Nx, L):
    base = rng.random((Ny * Nx, 2)) @ rng.random((2, L))
    x = (base + 0.1 * rng.random((Ny * Nx, L))).reshape((Ny, Nx, L), order='F')
    x[rng.random((Ny, Nx)) < 0.3] = 0.0          # a sparse-ish nonnegative cube
    return x.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize('kind,Ny,Nx,L,facets', [('hs', 48, 40, 6, None), ('fhs', 64, 72, 5, (2, 3, 0.25)),
                                                  ('fhs', 96, 96, 16, (3, 3, 0.5))])
def test_reweighting_weights_match_the_oracle(ctx, kind, Ny, Nx, L, facets):
    import fhs
    from fhs import solvers
    rng = np.random.default_rng(Ny + L)
    x = _cube(rng, Ny, Nx, L)
    geom = facet_geometry(Ny, Nx, *facets[:2], (facets[2], facets[2])) if facets else None
    S = fhs.SaraDictionary(ctx, Ny, Nx)
    R = fhs.LowRankPrior(ctx, Ny, Nx, L, facets=geom)
    Q = R.Q
    sig, alpha = 0.37, 1.0
    sig_bar = np.linspace(0.5, 1.5, Q)
    xd = torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0)).astype(np.float32)).cuda()
    eng = SimpleNamespace(ctx=ctx, sara=S, x=xd, meas=SimpleNamespace(L=L), lowrank=R,
                          weights1=torch.zeros(S.s, dtype=torch.float32, device=xd.device))
    solvers._reweight(eng, alpha, sig, sig_bar)
    w1_ref = sparsity_weights(x, SaraRef(Ny, Nx), sig * alpha)
    assert _rel(eng.weights1.cpu().numpy(), w1_ref) < TOL
    if kind == 'hs':
        w0_ref = lowrank_weights(x, sig_bar[0] * alpha)[None, :]
    else:
        sv = facet_singular_values(x, geom)
        ub = sig_bar[:, None] * alpha
        w0_ref = ub / (ub + np.abs(sv))
        # the apodisation window matters: without it the weights differ
        flat = dict(geom, windows=None)
        w0_flat = ub / (ub + np.abs(facet_singular_values(x, flat)))
        assert _rel(w0_flat, w0_ref) > 100 * TOL
    assert _rel(eng._w0, w0_ref) < TOL
    # weights in (0, 1], the largest coefficients get the smallest weights
    assert np.all(eng._w0 > 0) and np.all(eng._w0 <= 1)
    assert np.all(np.diff(eng._w0, axis=1) >= -1e-7)   # descending singular values -> ascending weights
