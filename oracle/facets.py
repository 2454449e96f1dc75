"""Oracle: facet geometry and the faceted low-rank dual step, float64.

Test infrastructure only (see ``oracle/__init__.py``).

Reference semantics
-------------------
* Facets: ``rg = split_range(Q, N)`` and ``split_range(Q, N, overlap)``
  (SARA-dictionary submodule, absent) at ``src/fhs/facetHyperSARA.m:297-321``;
  facet q = (qx-1)*Qy + qy; ``I`` = 0-based origin, ``dims`` sizes.  The
  overlapping (nuclear-norm) facets extend the plain ones towards the
  previous facet (top / left): their extra rows are cropped with
  ``overlap(1)+1:end`` (``src/fhs/facetHyperSARA.m:758``) and exchanged with
  the N/W/NW neighbours (``src/fhs/fhs_setup_priors.m:164-186``).
  Restated: boundaries ``round(i N / Q)`` (MATLAB round), overlapping facet
  q > 0 starts ``overlap`` pixels earlier.  Rounding is parity unpinned.
* Overlap size: ``get_overlap_size(N, Q, fraction)`` (submodule, called at
  ``imaging/imaging.m:421``) restated as ``floor(fraction * N / Q)``
  (parity unpinned).
* Apodisation ``generate_weights(..., 'triangular', ...)`` (submodule,
  ``src/fhs/fhs_setup_priors.m:189``) restated as a separable linear ramp
  over each overlap strip, rising on the leading strip and falling on the
  trailing one so that overlapping windows sum to one (parity unpinned).
"""
import numpy as np

from .prox import nuclear_prox_dual


def matlab_round(x):
    return np.floor(np.asarray(x, dtype=np.float64) + 0.5).astype(np.int64)


def split_range(Q, N, overlap=0):
    """(Q, 2) 0-based [start, stop) ranges."""
    b = matlab_round(np.arange(Q + 1) * N / Q)
    rg = np.stack([b[:-1], b[1:]], axis=1)
    if overlap > 0:
        rg[1:, 0] = np.maximum(rg[1:, 0] - overlap, 0)
    return rg


def overlap_size(N, Q, fraction):
    return int(np.floor(fraction * N / Q))


def ramp_1d(n_o, lead, trail):
    w = np.ones(n_o)
    if lead > 0:
        w[:lead] = np.arange(1, lead + 1) / (lead + 1)
    if trail > 0:
        w[n_o - trail:] = 1.0 - np.arange(1, trail + 1) / (trail + 1)
    return w


def facet_geometry(Ny, Nx, Qy, Qx, overlap_fraction):
    """Overlapping nuclear-norm facets: dict with Qy, Qx, oy, ox, py, px
    (facet q = qx*Qy + qy) and triangular windows (py, px)."""
    dy = overlap_size(Ny, Qy, overlap_fraction[0]) if Qy > 1 else 0
    dx = overlap_size(Nx, Qx, overlap_fraction[1]) if Qx > 1 else 0
    ry, rx = split_range(Qy, Ny, dy), split_range(Qx, Nx, dx)
    oy, ox, py, px, wins = [], [], [], [], []
    for qx in range(Qx):
        for qy in range(Qy):
            y0, y1 = ry[qy]
            x0, x1 = rx[qx]
            oy.append(int(y0)); ox.append(int(x0))
            py.append(int(y1 - y0)); px.append(int(x1 - x0))
            lead_y = dy if qy > 0 else 0
            lead_x = dx if qx > 0 else 0
            trail_y = dy if qy < Qy - 1 else 0
            trail_x = dx if qx < Qx - 1 else 0
            wins.append(np.outer(ramp_1d(y1 - y0, lead_y, trail_y), ramp_1d(x1 - x0, lead_x, trail_x)))
    return {'Qy': Qy, 'Qx': Qx, 'oy': oy, 'ox': ox, 'py': py, 'px': px, 'windows': wins,
            'overlap': (dy, dx)}


def faceted_lowrank_step(v0s, xhat, geom, weights0, beta0, sigma0):
    """``src/fhs/fhs_update_dual_lowrankness.m:40-45`` for every facet and the
    overlap-add of ``sigma00 * g0`` (``src/fhs/facetHyperSARA.m:746-754``,
    ``comm2d_reduce``).  v0s: list of (P_q, L); xhat (Ny, Nx, L);
    weights0 (Q, L).  Returns (v0s, g0 image)."""
    Ny, Nx, L = xhat.shape
    g0 = np.zeros_like(xhat)
    out = []
    for q in range(len(geom['oy'])):
        y0, x0, ny, nx = geom['oy'][q], geom['ox'][q], geom['py'][q], geom['px'][q]
        w = geom['windows'][q] if geom.get('windows') is not None else np.ones((ny, nx))
        xq = xhat[y0:y0 + ny, x0:x0 + nx, :]
        Z = v0s[q] + (w[:, :, None] * xq).reshape((ny * nx, L), order='F')
        v = nuclear_prox_dual(Z, beta0 * np.asarray(weights0[q]))
        out.append(v)
        g0[y0:y0 + ny, x0:x0 + nx, :] += sigma0 * w[:, :, None] * v.reshape((ny, nx, L), order='F')
    return out, g0


def facet_singular_values(x, geom):
    Ny, Nx, L = x.shape
    res = []
    for q in range(len(geom['oy'])):
        y0, x0, ny, nx = geom['oy'][q], geom['ox'][q], geom['py'][q], geom['px'][q]
        w = geom['windows'][q] if geom.get('windows') is not None else np.ones((ny, nx))
        X = (w[:, :, None] * x[y0:y0 + ny, x0:x0 + nx, :]).reshape((ny * nx, L), order='F')
        res.append(np.linalg.svd(X, compute_uv=False))
    return np.array(res)
