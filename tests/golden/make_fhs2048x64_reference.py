"""Writes gpurun_out/fhs2048x64_reference.npz (committed as
tests/golden/fhs2048x64_reference.npz): the converged plain GPU solves of
the four 16-channel spectral sub-cubes of the Faceted HyperSARA 2048x2048x64
workload (6x6 facets, 50% overlap, 1M visibilities per channel, w-term with
w in [0, 200] wavelengths over a 0.08 rad field, 3 reweighting steps, the
reference's stopping rules).  No FP64 solve of this size is feasible; the
FP64 oracle pins the same solver loop at small sizes
(tests/test_gpu_solver_convergence.py).  Per sub-cube it stores the SNR of
every iteration, the reweighting iterations and, per PDFB cycle, why it
stopped: its last relative variation and data residual over the l2 bound.

    python tests/golden/make_fhs2048x64_reference.py        (on a GPU, ~15 min)
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, 'faceted-hypersara_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import fhs  # noqa: E402
from fhs import solvers  # noqa: E402
from fhs.cube import CUBE_2048x64, build_subcube  # noqa: E402


def main():
    out = os.path.join(ROOT, 'gpurun_out', 'fhs2048x64_reference.npz')
    os.makedirs(os.path.dirname(out), exist_ok=True)
    ctx = fhs.default_context(0)
    cfg = CUBE_2048x64
    res = {}
    t_all = time.time()
    for c in range(cfg['subcubes']):
        t0 = time.time()
        prob, meas, facets, param = build_subcube(ctx, cfg, c)
        print(f'sub-cube {c}: set up in {time.time() - t0:.0f} s', flush=True)
        t0 = time.time()

        def progress(t, eng, info):
            if t % 250 == 0:
                print(f'  sub-cube {c} iteration {t}: SNR {info["snr"][-1]:.4f} dB, rel_val {info["rel_val"][-1]:.2e}, '
                      f'res/eps {info["norm_res"][-1] / info["norm_eps"][-1]:.5f}, '
                      f'reweights {info["reweights"]}, {time.time() - t0:.0f} s', flush=True)
        _, info = solvers._solve('fhs', meas, dict(param, track_snr=True), X0=prob.X0, facets=facets,
                                 callback=progress)
        rel = np.array(info['rel_val'])
        ratio = np.array(info['norm_res']) / np.array(info['norm_eps'])
        ends = info['reweights'] + [info['iterations']]
        starts = [0] + info['reweights']
        cycles = []
        for a, b in zip(starts, ends):
            cyc_rel, cyc_ratio = rel[a:b], ratio[a:b]
            small = np.nonzero(cyc_rel <= 5e-5)[0]
            cycles.append([b - a, cyc_rel[-1], cyc_ratio[-1], float(cyc_ratio.min()),
                           float(cyc_ratio[small[0]]) if len(small) else np.nan,
                           len(small) / max(1, b - a)])
        print(f'sub-cube {c}: {info["iterations"]} iterations, reweights {info["reweights"]}, '
              f'SNR {info["snr"][-1]:.4f} dB, {time.time() - t0:.0f} s; cycles (length, last rel_val, '
              f'last res/eps, min res/eps, res/eps at the first rel_val <= 5e-5, share of iterations with '
              f'rel_val <= 5e-5): {np.round(cycles, 6).tolist()}', flush=True)
        res[f'snr_{c}'] = np.array(info['snr'], dtype=np.float64)
        res[f'reweights_{c}'] = np.array(info['reweights'])
        res[f'iterations_{c}'] = info['iterations']
        res[f'converged_{c}'] = info['converged']
        res[f'cycles_{c}'] = np.array(cycles)
        res[f'rel_val_{c}'] = rel.astype(np.float32)
        res[f'res_over_eps_{c}'] = ratio.astype(np.float32)
        np.savez_compressed(out, **res)
        del meas, prob
    print(f'all sub-cubes in {time.time() - t_all:.0f} s -> {out}', flush=True)


if __name__ == '__main__':
    main()
