"""Writes tests/golden/solver_<case>.npz: full FP64 oracle solves (PDFB
cycles, reweighting to convergence) of the solver cases of
tests/solver_cases.py, for the GPU solver tests to compare against.

    python tests/golden/make_solver_fixtures.py [case ...]

Each fixture holds the operator norms the oracle computed (FP64 power
iteration; the GPU run reuses them so both solve the same problem), the
iteration and reweighting counts, the per-iteration relative variation and
SNR, and the final cube.  The data are rebuilt from the case's seed on
every machine (host FP64 visibilities), so only results are stored.
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, 'faceted-hypersara_amd'), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import solver_cases as sc  # noqa: E402
from oracle.solvers import pdfb  # noqa: E402


def make(name):
    kind, Ny, L, M, nb, seed, ns, pre, fac, over = sc.SOLVER_CASES[name]
    prob, op = sc.make_case(Ny, L, M, nb, seed, ns, pre)
    son, nu2 = sc.oracle_op_norms(op)
    geom = sc.case_geometry(name)
    p = sc.solver_param(kind, Ny, L, son, nu2, geom, **over)
    t0 = time.time()
    x, info = pdfb(kind, op, p, geom=geom, X0=prob.X0)
    dt = time.time() - t0
    out = dict(son=son, nu2=nu2, x=x, iterations=info['iterations'], reweights=np.array(info['reweights']),
               rel_val=np.array(info['rel_val']), snr=np.array(info['snr']),
               norm_res=np.array(info['norm_res']), norm_eps=np.array(info['norm_eps']),
               converged=info['converged'], seconds=dt)
    if geom is not None:
        for k in ('oy', 'ox', 'py', 'px'):
            out['geom_' + k] = np.array(geom[k])
    path = os.path.join(HERE, f'solver_{name}.npz')
    np.savez_compressed(path, **out)
    print(f'{name}: {info["iterations"]} iterations, reweights {info["reweights"]}, '
          f'SNR {info["snr"][-1]:.4f} dB, {dt:.1f} s -> {path}', flush=True)


if __name__ == '__main__':
    for n in (sys.argv[1:] or list(sc.SOLVER_CASES)):
        make(n)
