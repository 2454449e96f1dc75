"""A spectral cube imaged as C independent sub-cubes on one GPU.

The reference images each spectral sub-cube of a wide-band cube in a
separate run (``imaging/imaging.m:306``: one ``facetHyperSARA`` solve per
sub-cube of ``nChannels / C`` channels); here the C solves share one GPU.
All sub-cubes see the same sky, spectral indices and base uv coverage (the
cube's seed), each its own share of the cube's frequencies, with the uv
coverage scaled by nu_l / nu_max of the whole cube
(``lib/operators/util_gen_measurement_operator_synth.m:67-69``).
"""
import numpy as np

from . import solvers, synth
from .lowrank import facet_geometry

# Faceted HyperSARA 2048 x 2048 x 64: C = 4 sub-cubes of 16 channels at
# nu_l = 1 + 0.05 l, 6 x 6 nuclear-norm facets with 50 % overlap, 1M
# visibilities per channel (no flagging), 4 blocks per channel, w-term with
# w uniform in [0, 200] wavelengths over a 0.08 rad field of view,
# uniform-weight preconditioning, 3 reweighting steps.
CUBE_2048x64 = {
    'N': 2048, 'channels': 64, 'subcubes': 4, 'vis': 1000000, 'blocks': 4, 'seed': 0,
    'facets': (6, 6, 0.5), 'w_max': 200.0, 'fov': 0.08, 'dnu': 0.05, 'reweighting_max_iter': 3,
}


def cube_frequencies(cfg):
    return 1.0 + cfg['dnu'] * np.arange(cfg['channels'])


def build_subcube(ctx, cfg, c, precond=True):
    """(problem, measurement, facets, solver parameters) of sub-cube c."""
    nu = cube_frequencies(cfg)
    Lc = cfg['channels'] // cfg['subcubes']
    N = cfg['N']
    prob = synth.make_problem(N, N, Lc, cfg['vis'], cfg['blocks'], seed=cfg['seed'], flag_range=None,
                              w_max=cfg['w_max'] or None, fov=cfg['fov'], nu=nu[c * Lc:(c + 1) * Lc],
                              nu_max=float(nu.max()))
    meas = solvers.Measurement(ctx, prob, precond=precond)
    son, nu2 = meas.op_norm(False), meas.op_norm(precond)
    Qy, Qx, ov = cfg['facets']
    facets = facet_geometry(N, N, Qy, Qx, (ov, ov))
    param = {'squared_operator_norm': son, 'nu2': nu2, 'reweighting_max_iter': cfg['reweighting_max_iter']}
    return prob, meas, facets, param
