"""Benchmark: Faceted HyperSARA PDFB iterations on one MI355X.

Default workload (--config fhs2048x64): Faceted HyperSARA on a 2048 x 2048 x
64 cube as C = 4 spectral sub-cubes of 16 channels, all resident on one GPU;
6 x 6 nuclear-norm facets with 50 % overlap, 1M visibilities per channel
(Gaussian coverage scaled by nu_l / nu_max), 4 blocks per channel, the
w-term (w uniform in [0, 200] wavelengths, w-projected separable rows of
15 x 15 taps), uniform-weight preconditioning, db1-db8 + Dirac at 4 levels,
input SNR 40 dB.  Synthetic seeded data (fhs/cube.py).  A step is one PDFB
iteration of every sub-cube (src/fhs/facetHyperSARA.m:701-889).

The JSON line carries iterations/s of the chosen form (the PDFB engine: one
hipGraph replay per iteration, priors on side streams, the mixed-precision
Gram of --gram on the matrix cores), the dominant kernel against the HBM
roofline (HIP-event timing on the engine stream, algorithmic bytes), and
``solve``: per sub-cube the wall-clock and iterations to reach the
reference SNR minus 0.1 ... 3 dB for the plain form (the reference-named
drop-in entry points called one after the other every iteration, FP64
Gram, zero start) and the chosen form, with the errors of the mixed
Gram.  The CPU baseline is a bounded sample of one cube iteration by the
FP64 oracle on the host.

--config hs512 (HyperSARA 512 x 512 x 16, 200k base visibilities, reference
SNR from the FP64 oracle's converged solve, tests/golden/hs512x16_reference.npz)
also times the batched FP64-Gram engine and the learned warm start;
hs512dr, fhs1024, fhs2048 and sara256 measure the other workloads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, 'faceted-hypersara_amd')):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = 'time-to-converged cube (s) at ref SNR−0.1 dB, chosen vs plain form; iters/s'
PEAK_HBM_GBS = 8000.0


def log(msg):
    """One progress line on stderr per finished stage (stdout keeps the JSON line)."""
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--size', type=int, default=512)
    ap.add_argument('--channels', type=int, default=16)
    ap.add_argument('--vis', type=int, default=200000)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-solve', action='store_true', help='skip the time-to-SNR solve')
    ap.add_argument('--solve-seconds', type=float, default=60.0)
    ap.add_argument('--solve-repeats', type=int, default=3,
                    help='timed solves per start after the full reference solve; the median time is reported')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run a plain 256 MiB device-to-device copy and a 256 MiB elementwise add, 3 times '
                         'each (the reference kernels for scaling the FETCH_SIZE / WRITE_SIZE counters), and exit')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--gram', choices=('fp64', 'bf16x3', 'bf16x2'), default='bf16x3',
                    help='nuclear-norm Gram precision of the chosen form (throughput engine and chosen solves)')
    ap.add_argument('--rows', type=int, choices=(16, 32), default=16,
                    help='storage precision of the G rows the data kernels read in the chosen form (throughput '
                         'engine and chosen solves; tile-gridded separable G only, 32 for the DR form)')
    ap.add_argument('--no-autotune', action='store_true',
                    help='keep the default launch shapes (the tuned ones give bit-identical results)')
    ap.add_argument('--w-max', type=float, default=None,
                    help='w-term: base w uniform in [0, W] wavelengths (fhs2048: 200), field of view --fov')
    ap.add_argument('--fov', type=float, default=0.08, help='field of view of the w-term (direction cosines)')
    ap.add_argument('--config', choices=('fhs2048x64', 'hs512', 'hs512dr', 'fhs1024', 'fhs2048', 'sara256'),
                    default='fhs2048x64',
                    help='fhs2048x64: Faceted HyperSARA 2048x2048x64 (the headline workload) as C = 4 spectral '
                         'sub-cubes of 16 channels solved on one GPU, 6x6 facets with 50%% overlap, 1M visibilities '
                         'per channel, w-term (w in [0, 200] wavelengths, 0.08 rad field); hs512: HyperSARA '
                         '512x512x16 against the FP64 oracle\'s converged SNR; fhs1024: one 16-channel spectral '
                         'sub-cube of Faceted HyperSARA 1024x1024x32 (C = 2), 4x4 nuclear-norm facets with 25%% '
                         'overlap, 300k-500k visibilities per channel; fhs2048: one 16-channel spectral sub-cube '
                         'of Faceted HyperSARA 2048x2048x64 (C = 4), 6x6 facets with 50%% overlap, 1M visibilities '
                         'per channel, w-term with w in [0, 200] wavelengths over a 0.08 rad field (--w-max 0 '
                         'turns it off); sara256: SARA 256x256, 65,536 '
                         'visibilities, 2 blocks, 300 iterations, against the FP64 oracle run to the same end '
                         '(the CPU-runnable case: its SNR is the reference); hs512dr: the hs512 problem in its '
                         'dimensionality-reduced form (holographic H = tril(G\'G) per channel, no preconditioning, '
                         'imaging/imaging.m:619-631), its own converged SNR the reference')
    a = ap.parse_args()
    if a.config == 'fhs2048x64':
        return a
    if a.config == 'fhs1024':
        a.size, a.channels, a.vis = 1024, 16, 500000
        a.no_cpu = True     # one FP64 oracle iteration of this size takes minutes
    if a.config == 'fhs2048':
        a.size, a.channels, a.vis = 2048, 16, 1000000
        a.no_cpu = True
        if a.w_max is None:
            a.w_max = 200.0
    if a.config == 'sara256':
        a.size, a.channels, a.vis = 256, 1, 65536
    if a.config == 'hs512dr':
        a.no_cpu = True     # the FP64 oracle's sparse G'G of 16 channels takes minutes
    return a


def lowrank_bytes(lowrank, L, N):
    """Per-launch algorithmic bytes of the nuclear-norm step (Q facets of
    P_q pixels, sum P = sum_q P_q): the Gram reads v0 and the windowed facet
    of x-hat (8 L sum P) and the windows (4 sum P); the reduce + eigen-solve
    read the per-chunk FP64 partial Grams (Q * chunks * 256 * 8) and write
    the L x L mixing matrices; the apply reads v0, x-hat and the windows and
    writes v0 (12 L sum P + 4 sum P); the overlap-add reads every facet's v0
    and window once and writes g0 (4 L sum P + 4 sum P + 4 L N)."""
    if lowrank is None:
        return {}
    P = [int(a) * int(b) for a, b in zip(lowrank.geom['py'], lowrank.geom['px'])]
    sP, Q = sum(P), len(P)
    chunks = -(-max(P) // 1024)
    out = {'nuclear_gram': 8 * L * sP + 4 * sP,
           'nuclear_eig': Q * chunks * 256 * 8 + Q * 256 * 8 * 2,
           'nuclear_apply': 12 * L * sP + 4 * sP}
    if not (Q == 1 and sP == N):
        out['overlap_add'] = 4 * L * sP + 4 * sP + 4 * L * N
    return out


def algorithmic_bytes(meas, sara_s, lowrank=None):
    """Per-launch algorithmic HBM bytes of the kernels for the layouts in use
    (complex = 8 B, real FP32 = 4 B; every operand read or written once).
    G is the separable min-max layout: Ky + Kx coefficients (8 B each, 4 B
    when the data state reads FP16 rows) and an anchor per row.  Its adjoint
    grids 32 x 32 tiles from per-tile row lists (row index + local cell, 8 B
    per listed row): each row's coefficients, v2 and pU counted once, each
    grid point written once."""
    L, No = meas.L, meas.nufft.No
    nf = meas.nufft
    M = [g.M for g in meas.G]
    N = meas.Ny * meas.Nx
    Ky, Kx = getattr(meas.G[0], 'Ky', nf.Ky), getattr(meas.G[0], 'Kx', nf.Kx)
    coef = 4 if getattr(getattr(meas, 'D', None), 'row_bits', 32) == 16 else 8
    row = coef * (Ky + Kx) + 4
    fwd = sum(row * m + 8 * No + 24 * m for m in M)                      # coefs, Fx, gx_old r/w, r2
    adj = 0
    for a in meas.anchor:
        a = a.astype(np.int64)
        ay, ax = a % nf.Kty, a // nf.Kty
        ny = 1 + (ay // 32 != ((ay + Ky - 1) % nf.Kty) // 32)
        nx = 1 + (ax // 32 != ((ax + Kx - 1) % nf.Ktx) // 32)
        entries = int(np.sum(ny * nx))
        adj += (row - 4 + 8 + 4) * len(a) + 8 * entries + 4 * (No // 1024 + 1) + 8 * No
    return {
        'data_adjoint': adj,
        'data_forward': fwd,
        'fft_fwd_y': 4 * N * L + 8 * (No // 2) * L,          # read image, write half-grid
        'fft_fwd_x': 8 * (No // 2) * L + 8 * No * L,
        'fft_adj_x': 8 * No * L + 8 * (No // 2) * L,
        'fft_adj_y': 8 * (No // 2) * L + 4 * N * L,
        'wavelet_analysis': 4 * L * N + 8 * L * sara_s,
        'wavelet_synthesis': 4 * L * sara_s + 4 * L * N,
        'l21_threshold': 8 * L * sara_s + 4 * L * N + 4 * sara_s,
        'primal': 4 * L * N * 6,
        'fb_dual': 44 * sum(M),
        **lowrank_bytes(lowrank, L, N),
    }


def algorithmic_bytes_dr(meas, sara_s, lowrank=None):
    """The DR form's per-launch algorithmic bytes: every active mode's full
    Hermitian stencil ((2Ky-1)(2Kx-1) complex values) read once per apply, the
    grid vectors once; FFT and prior kernels as in the standard form."""
    out = algorithmic_bytes(meas.meas, sara_s, lowrank)
    nf = meas.meas.nufft
    No = nf.No
    ns = (2 * nf.Ky - 1) * (2 * nf.Kx - 1)
    W = [f.n_modes for f in meas.forms]
    st = sum(8 * ns * w for w in W)
    out['data_forward'] = st + sum(8 * No + 28 * w for w in W)     # Fx in; Lambda, gx_old r/w, r2 on W
    out['data_adjoint'] = st + len(W) * 28 * No                    # v2, Lambda in; g2 out
    out['l2_dual'] = 64 * len(W) * No                              # two passes over v2, r2, y (+ gx), v2 out
    del out['fb_dual']
    return out


def sara_against_oracle(args, solvers, meas, prob, son, nu2, n_iter=300):
    """SARA 256x256: the plain FP32 GPU run and the FP64 numpy oracle, both
    n_iter iterations without reweighting from the same data.  The oracle's
    final SNR is the reference; returns (solve, cpu_baseline)."""
    from threadpoolctl import threadpool_limits
    from fhs import warmstart
    from oracle.solvers import OracleProblem, sara_solver
    prm = {'squared_operator_norm': son, 'nu2': nu2, 'pdfb_min_iter': n_iter, 'pdfb_max_iter': n_iter,
           'reweighting_min_iter': 0, 'reweighting_max_iter': 0, 'track_snr': True}
    _, full = solvers._solve('sara', meas, prm, X0=prob.X0)           # untimed: reference curve, warm-up
    b = np.concatenate([[0], np.cumsum(prob.blocks[0])])
    blocks = [[(int(b[i]), int(b[i + 1])) for i in range(len(prob.blocks[0]))]]
    with threadpool_limits(limits=1):
        op = OracleProblem(prob.Ny, prob.Nx, meas.u, meas.v, meas.y_host,
                           [np.asarray(p, dtype=np.float64) for p in meas.pU_host], blocks, prob.epsilon,
                           g_scale=[1.0 / s for s in prob.sigma_noise])
        pr = dict(solvers.DEFAULT_PARAM)
        # the same loop as the GPU run: n_iter PDFB iterations, no reweighting
        pr.update({k: prm[k] for k in ('pdfb_min_iter', 'pdfb_max_iter', 'reweighting_min_iter',
                                       'reweighting_max_iter')})
        pr.update({'nu2': nu2, 'gamma': float(np.sqrt(1.0 / son[0]))})
        tc = time.perf_counter()
        _, oinfo = sara_solver(op, pr, X0=prob.X0[:, :, 0], max_iter_total=n_iter)
        tcpu = time.perf_counter() - tc
    if len(oinfo['snr']) != n_iter or full['iterations'] != n_iter or oinfo['reweights']:
        raise RuntimeError(f"SARA reference runs differ in length: oracle {len(oinfo['snr'])} iterations, "
                           f"{len(oinfo['reweights'])} reweights; GPU {full['iterations']}")
    ref = oinfo['snr'][-1]
    tols = (0.1, 0.5, 1.0, 2.0, 3.0)
    runs = [solvers._solve('sara', meas, prm, X0=prob.X0, stop_snr=ref - min(tols))[1]
            for _ in range(max(1, args.solve_repeats))]
    gpu, cpu_reach = {}, {}
    for d in tols:
        key = f'ref-{d}dB'
        k = warmstart.iterations_to(full['snr'], ref - d)
        ok = k is not None and all(len(r['time']) >= k for r in runs)
        gpu[key] = {'iterations': k, 's': round(float(np.median([r['time'][k - 1] for r in runs])), 4)} if ok else None
        kc = warmstart.iterations_to(oinfo['snr'], ref - d)
        cpu_reach[key] = None if kc is None else {'iterations': kc, 's': round(tcpu * kc / n_iter, 2)}
    solve = {'ref_snr_db': round(ref, 3),
             'ref_definition': f'SNR of the FP64 numpy oracle after {n_iter} SARA iterations (no reweighting) on '
                               'the same data; SNR = 20 log10(||X0||/||X0 - X||)',
             'gpu_fp32_snr_db_at_end': round(full['snr'][-1], 3),
             'snr_gap_db': round(full['snr'][-1] - ref, 4),
             'timing': f'GPU: median of {max(1, args.solve_repeats)} solves after a full one; CPU: one run, '
                       'time at iteration k taken as k/n of the run',
             'plain': {'iterations': full['iterations'], 'reach': gpu},
             'cpu_oracle': {'iterations': n_iter, 'wall_s': round(tcpu, 2), 'reach': cpu_reach}}
    cpu = {'value': round(n_iter / tcpu, 4), 'unit': 'iters/s', 'cores': 1, 'kind': 'port',
           'sample': f'the whole {n_iter}-iteration SARA 256x256 run of the FP64 numpy oracle, single thread '
                     f'({tcpu:.1f} s)'}
    return solve, cpu


def gram_precision_errors(ctx, meas, facets, x, mode):
    """Error of the mixed-precision Gram against the FP64 Gram (itself within
    1e-6 of the FP64 oracle in the GPU tests) on a solve's cube x: relative
    2-norm errors of the facets' singular values and of one low-rank dual
    step from v0 = 0 (threshold = median singular value, unit weights)."""
    import numpy as np
    from fhs import LowRankPrior
    out = {}
    res = {}
    for m in ('fp64', mode):
        R = LowRankPrior(ctx, meas.Ny, meas.Nx, meas.L, facets=facets)
        R.set_gram_precision(m)
        sv = R.singular_values(x)
        beta = float(np.median(res['fp64'][0])) if m != 'fp64' else float(np.median(sv))
        v0, g0 = R.update_dual(R.zeros_v0(), x, np.ones((R.Q, R.L)), beta, 1.0)
        res[m] = (sv, v0.double(), g0.double())
        del R
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    out['singular_values_rel_err'] = rel(res[mode][0], res['fp64'][0])
    out['lowrank_dual_step_rel_err'] = float((res[mode][1] - res['fp64'][1]).norm() / res['fp64'][1].norm())
    out['lowrank_g0_rel_err'] = float((res[mode][2] - res['fp64'][2]).norm() / res['fp64'][2].norm())
    out['reference'] = 'the FP64-Gram GPU path on the same cube (tested within 1e-6 of the FP64 oracle)'
    return out


def row_precision_errors(meas, bits):
    """Error of the data forward / adjoint with the G rows stored in `bits`
    against the rows as created (FP32; those are tested within 1e-5 of the
    FP64 oracle), on a random cube and random data: relative 2-norm errors."""
    import torch
    from fhs import solvers
    lib, D, ctx = solvers.lib(), meas.D, meas.ctx
    g = torch.Generator(device='cuda').manual_seed(1)
    x = torch.rand((meas.L, meas.Nx, meas.Ny), dtype=torch.float32, device='cuda', generator=g)
    y = torch.randn(D.Mtot, dtype=torch.complex64, device='cuda', generator=g)
    prev = D.row_bits
    res = {}
    for b in (32, bits):
        D.set_row_precision(b)
        f = torch.empty(D.Mtot, dtype=torch.complex64, device='cuda')
        a = torch.empty_like(x)
        ctx.check(lib.fhs_data_forward(ctx.h, D.h, x.data_ptr(), None, f.data_ptr()), 'data_forward')
        ctx.check(lib.fhs_data_adjoint(ctx.h, D.h, y.data_ptr(), None, None, a.data_ptr()), 'data_adjoint')
        res[b] = (f.to(torch.complex128), a.double())
    D.set_row_precision(prev)
    rel = lambda u, v: float((u - v).norm() / v.norm())  # noqa: E731
    return {'forward_rel_err': rel(res[bits][0], res[32][0]), 'adjoint_rel_err': rel(res[bits][1], res[32][1]),
            'reference': 'the same data state with FP32 rows (tested within 1e-5 of the FP64 oracle)'}


_CPU_BARRIER = None


def _cpu_pool_init(barrier):
    global _CPU_BARRIER
    _CPU_BARRIER = barrier


def _cube_cpu_worker(task):
    """One host worker of the cube's CPU sample: builds the sampled rows'
    G, waits until every worker is ready, then times one channel's
    transforms, rows and wavelets and one facet's prox (FP64 oracle, one
    thread).  Returns (t_transforms, t_rows, t_wavelets, t_prox)."""
    import scipy.sparse as sp
    from threadpoolctl import threadpool_limits
    from oracle.nufft import NufftOracle
    from oracle.prox import nuclear_prox_dual, solver_proj_elipse_fb
    from oracle.wavelets import SaraDictionary
    d, seed = task
    N, Kty, Ktx, a, cy, cx = d['N'], d['Kty'], d['Ktx'], d['a'], d['cy'], d['cx']
    m, Ky, Kx = len(a), cy.shape[1], cx.shape[1]
    ky = (a % Kty)[:, None] + np.arange(Ky)[None, :]
    kx = (a // Kty)[:, None] + np.arange(Kx)[None, :]
    cols = ((ky % Kty)[:, :, None] + Kty * (kx % Ktx)[:, None, :]).reshape(m, -1)
    vals = (cy.astype(np.complex128)[:, :, None] * cx.astype(np.complex128)[:, None, :]).reshape(m, -1)
    G = sp.csr_matrix((vals.ravel(), (np.repeat(np.arange(m), Ky * Kx), cols.ravel())), shape=(m, Kty * Ktx))
    GH = G.conj().T.tocsr()
    op = NufftOracle(np.zeros(1), np.zeros(1), N, N)
    sara = SaraDictionary(N, N)
    rng = np.random.default_rng(seed)
    x = rng.random((N, N))
    y = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    Z = rng.standard_normal((d['facet_pixels'], d['facet_channels']))
    with threadpool_limits(1):
        _CPU_BARRIER.wait(timeout=600)   # all workers time their phases together
        t0 = time.perf_counter()
        Fx = op.A(x)                                   # transforms: A and At
        _ = np.real(op.At(Fx))
        t_tr = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = G @ Fx                                     # rows: G, the ellipsoid projection, G'
        z, _ = solver_proj_elipse_fb(np.zeros(m), r, y, np.ones(m), 0.5 * np.linalg.norm(r - y),
                                     np.zeros(m, complex), 20, 1, 1e-8)
        _ = GH @ (r - z)
        t_rows = time.perf_counter() - t0
        t0 = time.perf_counter()
        c = sara.psit(x)
        _ = sara.psi(c)
        t_wl = time.perf_counter() - t0
        t0 = time.perf_counter()
        nuclear_prox_dual(Z, np.ones(d['facet_channels']))
        t_fac = time.perf_counter() - t0
    return t_tr, t_rows, t_wl, t_fac


def cpu_sample_cube(cfg, prob, meas, facets, pool, workers):
    """The FP64 oracle's cost of one PDFB iteration of the whole cube on the
    host's cores, from a bounded sample run by `workers` processes at once
    (one core each): every worker times one channel's data step at full
    size (A, At and the ellipsoid projection) on the first 1/16 of the
    channel's w-projected rows, the channel's wavelet analysis and
    synthesis, and one facet's low-rank prox (SVD of the facet's 16
    channels).  A cube iteration takes ceil(channels / workers) rounds of
    the slowest worker's channel (transforms + 16 x row-proportional part
    + wavelets) plus ceil(facet proxes / workers) rounds of the prox."""
    N = cfg['N']
    C, L = cfg['subcubes'], cfg['channels']
    frac = 16
    m = len(prob.u[0]) // frac
    nf = meas.nufft
    a, cy, cx = nf.separable_wproj_gpu(meas.ctx, prob.u[0][:m], prob.v[0][:m], prob.w[0][:m], cfg['fov'])
    d = {'N': N, 'Kty': nf.Kty, 'Ktx': nf.Ktx, 'a': np.asarray(a), 'cy': np.asarray(cy), 'cx': np.asarray(cx),
         'facet_pixels': max(facets['py']) * max(facets['px']), 'facet_channels': L // C}
    t0 = time.perf_counter()
    res = np.array(pool.map(_cube_cpu_worker, [(d, s) for s in range(workers)], chunksize=1))
    wall = time.perf_counter() - t0
    t_tr, t_rows, t_wl, t_fac = (res[:, j] for j in range(4))
    per_channel = t_tr + frac * t_rows + t_wl
    nq = len(facets['oy'])
    rounds = -(-L // workers)
    t_iter = rounds * float(per_channel.max()) + (-(-C * nq // workers)) * float(t_fac.max())
    Ky, Kx = cy.shape[1], cx.shape[1]
    return {'value': round(1.0 / t_iter, 6), 'unit': 'iters/s', 'cores': workers, 'nproc': os.cpu_count(),
            'kind': 'port',
            'sample': (f'FP64 numpy oracle, {workers} processes at once (one core each, {wall:.1f} s wall): each '
                       f'times one channel at {N}x{N} -- A and At (median {np.median(t_tr):.2f} s), G, the 20-step '
                       f'ellipsoid projection and G\' on 1/{frac} of its {len(prob.u[0])} w-projected rows '
                       f'({Ky}x{Kx} taps: {np.median(t_rows):.2f} s), the SARA analysis and synthesis '
                       f'({np.median(t_wl):.2f} s) -- and one facet\'s low-rank prox ({np.median(t_fac):.2f} s); '
                       f'one cube iteration = {rounds} rounds of the slowest worker\'s channel (transforms + {frac} '
                       f'x rows + wavelets, {per_channel.max():.1f} s) + {-(-C * nq // workers)} rounds of the prox '
                       f'({C} x {nq} facets) = {t_iter:.0f} s on {workers} cores')}


def cube_main(args):
    """The 2048x2048x64 cube as C independent sub-cube solves on one GPU."""
    import torch
    import fhs
    from fhs import solvers, warmstart
    from fhs.cube import CUBE_2048x64, build_subcube
    cfg = dict(CUBE_2048x64)
    if args.w_max is not None:
        cfg['w_max'] = args.w_max
    cfg['fov'] = args.fov
    # the CPU baseline's worker processes are forked before the first HIP call
    pool, workers = None, min(16, os.cpu_count() or 1)   # the box's CPU share for one GPU
    if not args.no_cpu:
        import multiprocessing as mp
        fork = mp.get_context('fork')
        pool = fork.Pool(workers, initializer=_cpu_pool_init, initargs=(fork.Barrier(workers),))
    ctx = fhs.default_context(0)
    ctx.set_autotune(not args.no_autotune)
    C = cfg['subcubes']
    t_setup = time.perf_counter()
    subs = []
    for c in range(C):
        subs.append(build_subcube(ctx, cfg, c))
        log(f'sub-cube {c} set up ({time.perf_counter() - t_setup:.0f} s)')
    setup_s = time.perf_counter() - t_setup
    # ---- throughput: one PDFB iteration of every sub-cube per step ----------
    engines = []
    for prob, meas, facets, param in subs:
        holder = {}

        def grab(t, eng, info, holder=holder):
            holder['eng'] = eng
            return t < args.warmup
        solvers._solve('fhs', meas, param, facets=facets, callback=grab, max_total_iter=args.warmup + 1,
                       gram=args.gram, rows=args.rows)
        engines.append(holder['eng'])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for eng in engines:
            eng.iterate(1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cube_per_s = args.steps / elapsed
    log(f'{args.steps} timed cube iterations ({C} sub-cubes each): {cube_per_s:.3f} iterations/s')
    # ---- per-kernel timing on sub-cube 0 ----------------------------------
    eng = engines[0]
    eng.set_concurrent(False)
    ctx.timing_reset()
    ctx.set_timing(True)
    n_timed = 5
    eng.iterate(n_timed)
    ctx.set_timing(False)
    rep = ctx.timing_report()
    ctx.timing_reset()
    meas0 = subs[0][1]
    alg = algorithmic_bytes(meas0, eng.sara.s, eng.lowrank)
    dominant = max((k for k in rep if k in alg), key=lambda k: rep[k][0])
    avg_ms = rep[dominant][0] / rep[dominant][1]
    achieved = alg[dominant] / (avg_ms * 1e-3) / 1e9
    kernels = {k: {'ms_per_launch': rep[k][0] / rep[k][1], 'launches_per_iter': rep[k][1] / n_timed,
                   'GBps': (alg[k] / (rep[k][0] / rep[k][1] * 1e-3) / 1e9) if k in alg else None}
               for k in rep}
    del engines, eng
    torch.cuda.synchronize()
    # HBM bytes per launch of the dominant kernel on sub-cube 0, from the
    # copy-calibrated FETCH_SIZE / WRITE_SIZE passes of this workload
    traffic = None
    tpath = os.path.join(ROOT, 'profiles', 'fhs2048x64_pmc_traffic.json')
    if os.path.exists(tpath) and dict(CUBE_2048x64, fov=args.fov) == cfg and args.seed == 0:
        with open(tpath) as fh:
            traffic = json.load(fh).get('traffic_bytes_per_launch', {}).get(dominant)
    # ---- time to reference SNR: per sub-cube, plain (drop-in) and chosen --
    solve = None
    fx_path = os.path.join(ROOT, 'tests', 'golden', 'fhs2048x64_reference.npz')
    default_cfg = dict(CUBE_2048x64, fov=args.fov) == cfg and args.seed == 0
    if not args.no_solve and os.path.exists(fx_path) and default_cfg:
        fx = np.load(fx_path)
        tols = (0.1, 0.5, 1.0, 2.0, 3.0)
        per = []
        mixed_err = None
        for c, (prob, meas, facets, param) in enumerate(subs):
            ref = float(fx[f'snr_{c}'][-1])
            prm = dict(param, track_snr=True)
            row = {'ref_snr_db': round(ref, 4), 'reference_iterations': int(fx[f'iterations_{c}']),
                   'reference_reweights': [int(t) for t in fx[f'reweights_{c}']]}
            for name, kw in (('plain', dict(form='dropin')),
                             ('chosen', dict(form='engine', gram=args.gram, rows=args.rows))):
                eng_s, info = solvers._solve('fhs', meas, prm, X0=prob.X0, facets=facets, stop_snr=ref - min(tols),
                                             time_limit=args.solve_seconds, **kw)
                reach = {}
                for d in tols:
                    k = warmstart.iterations_to(info['snr'], ref - d)
                    reach[f'ref-{d}dB'] = None if k is None else {'iterations': k, 's': round(info['time'][k - 1], 4)}
                row[name] = reach
                if name == 'chosen' and c == 0:
                    mixed_err = gram_precision_errors(ctx, meas, facets, eng_s.x, args.gram) \
                        if args.gram != 'fp64' else {}
                    if args.rows != 32:
                        mixed_err['g_rows'] = dict(row_precision_errors(meas, args.rows), bits=args.rows)
                del eng_s
            per.append(row)
            log(f'sub-cube {c}: timed solves done')
        total = {}
        for f in ('plain', 'chosen'):
            total[f] = {}
            for d in tols:
                key = f'ref-{d}dB'
                vals = [r[f][key] for r in per]
                total[f][key] = None if any(v is None for v in vals) else round(sum(v['s'] for v in vals), 4)
        speed = {key: (None if total['plain'][key] is None or total['chosen'][key] is None else
                       round(total['plain'][key] / total['chosen'][key], 3)) for key in total['plain']}
        solve = {'ref_definition': ('per sub-cube, the converged SNR of the plain GPU solve with the reference '
                                    'stopping rules and 3 reweighting steps (tests/golden/fhs2048x64_reference.npz; '
                                    'no FP64 solve of this size is feasible -- the solver loop is checked against '
                                    'the FP64 oracle at small sizes); SNR = 20 log10(||X0||/||X0 - X||)'),
                 'timing': 'the sub-cubes are independent solves run one after the other: the cube reaches a '
                           'tolerance when every sub-cube has, at the sum of their times',
                 'plain_form': 'drop-in entry points per iteration, eager, FP64 Gram, zero start',
                 'chosen_form': (f'PDFB engine (one hipGraph replay per iteration, priors on side streams, stopping '
                                 f'rule in the library), {args.gram} Gram on the matrix cores, FP{args.rows} G '
                                 f'rows, zero start'),
                 'warm_start': ('not part of the chosen form: the learned start does not shorten the time to '
                                'ref - 0.1 dB at 512x512x16, and the shipped network is trained on 512x512x16 cubes'),
                 'mixed_precision_errors': mixed_err,
                 'subcubes': per, 'total_s': total, 'speedup_chosen_vs_plain': speed}
    # ---- CPU baseline: bounded oracle sample --------------------------------
    cpu = None
    if not args.no_cpu:
        prob, meas, facets, _ = subs[0]
        cpu = cpu_sample_cube(cfg, prob, meas, facets, pool, workers)
        pool.close()
        pool.join()
        log('CPU baseline sample done')
    nnz = int(sum(g.nnz for (_, m, _, _) in subs for g in m.G))
    Ky = subs[0][1].G[0].Ky
    out = {
        'metric': METRIC,
        'value': round(cube_per_s, 4),
        'unit': 'iters/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * elapsed / args.steps, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': f'f32 (f64 norms, {args.gram} Gram, f64 eig, fp{args.rows} G rows)',
        'data': 'synthetic',
        'config': {'workload': (f'Faceted HyperSARA {cfg["N"]}x{cfg["N"]}x{cfg["channels"]} as C = {C} spectral '
                                f'sub-cubes of {cfg["channels"] // C} channels solved on one GPU (one step = one PDFB '
                                f'iteration of every sub-cube), 6x6 nuclear-norm facets, 50% overlap, '
                                f'{cfg["vis"]} vis/channel, 4 blocks/channel, w-term (w in [0, {cfg["w_max"]:g}] '
                                f'wavelengths, {cfg["fov"]} rad field, w-projected separable rows of {Ky}x{Ky} taps)'),
                   'Mtot': int(sum(m.D.Mtot for (_, m, _, _) in subs)), 'nnz_G_total': nnz,
                   'setup_s': round(setup_s, 2)},
        'roofline': {'bound': 'hbm', 'kernel': dominant, 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4),
                     'traffic': None if traffic is None else int(traffic),
                     'traffic_source': ('profiles/fhs2048x64_pmc_traffic.json (FETCH_SIZE x 2 + WRITE_SIZE, '
                                        'copy-calibrated; an upper bound for this gather kernel: its raw '
                                        'FETCH_SIZE is 64 bytes per memory request, and the factor 2 fits the '
                                        'streaming copy)'),
                     'avg_launch_ms': round(avg_ms, 5), 'algorithmic_bytes_per_launch': int(alg[dominant]),
                     'measured_on': 'sub-cube 0 (16 channels)'},
        'kernels': kernels,
        'tuned_launch_shapes': ctx.autotune_report(),
        'solve': solve,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus must be 1',
                          'n_gpus': args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import fhs
    from fhs import solvers, synth

    if args.copy_calibration:
        n = 256 << 20
        a = torch.ones(n // 4, dtype=torch.float32, device='cuda')
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)             # the runtime's copy kernel
        for _ in range(3):
            torch.add(a, 1.0, out=b)   # an elementwise kernel (vector loads and stores)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_read': n, 'copy_bytes_written': n, 'repeats': 3,
                          'kernels': ['copy', 'elementwise add']}))
        return
    if args.config == 'fhs2048x64':
        return cube_main(args)
    ctx = fhs.default_context(0)
    ctx.set_autotune(not args.no_autotune)
    t_setup = time.perf_counter()
    facets = None
    if args.config == 'fhs1024':
        from fhs.lowrank import facet_geometry
        prob = synth.make_problem(args.size, args.size, args.channels, args.vis, 4, seed=args.seed,
                                  flag_range=(0.0, 0.4))
        facets = facet_geometry(args.size, args.size, 4, 4, (0.25, 0.25))
    elif args.config == 'fhs2048':
        from fhs.lowrank import facet_geometry
        prob = synth.make_problem(args.size, args.size, args.channels, args.vis, 4, seed=args.seed,
                                  flag_range=None, w_max=args.w_max or None, fov=args.fov)
        facets = facet_geometry(args.size, args.size, 6, 6, (0.5, 0.5))
    elif args.config == 'sara256':
        prob = synth.make_problem(args.size, args.size, 1, args.vis, 2, seed=args.seed)
    else:
        prob = synth.make_problem(args.size, args.size, args.channels, args.vis, 4, seed=args.seed)
    # The headline workload's reference: the FP64 oracle's solve of the same
    # problem run to convergence (tests/golden/make_hs512_reference.py), on
    # data computed in FP64 on the host so both solve identical problems.
    ref_fx = None
    if args.config == 'hs512':
        synth.add_host_data(prob)
        fx_path = os.path.join(ROOT, 'tests', 'golden', 'hs512x16_reference.npz')
        if (args.size, args.channels, args.vis, args.seed) == (512, 16, 200000, 0) and os.path.exists(fx_path):
            with np.load(fx_path) as z:
                ref_fx = {k: z[k] for k in z.files}
    kind = 'fhs' if facets is not None else ('sara' if args.config == 'sara256' else 'hs')
    dr = args.config == 'hs512dr'
    if dr:
        from fhs.holographic import HolographicMeasurement
        meas = HolographicMeasurement(solvers.Measurement(ctx, prob, precond=False))
        son = meas.op_norm()
        nu2 = son
    else:
        meas = solvers.Measurement(ctx, prob, precond=True)
        son = meas.op_norm(False)
        nu2 = meas.op_norm(True)
    opnorm_gap = None
    if ref_fx is not None:
        # the oracle's FP64 operator norms, so the step sizes and thresholds
        # are the oracle run's; the GPU power iteration's distance to them
        opnorm_gap = float(max(np.max(np.abs(son / ref_fx['son'] - 1)), np.max(np.abs(nu2 / ref_fx['nu2'] - 1))))
        son, nu2 = ref_fx['son'], ref_fx['nu2']
    setup_s = time.perf_counter() - t_setup
    log(f'setup {setup_s:.1f} s')
    param = {'squared_operator_norm': son, 'nu2': nu2, 'reweighting_max_iter': 3}

    # ---- throughput: steps through the hipGraph engine --------------------
    holder = {}

    def grab(t, eng, info):
        holder['eng'] = eng
        return t < args.warmup

    # the chosen form's G rows (FP32 for SARA, which has no chosen form, and the DR form's untiled G)
    rows_c = 32 if (kind == 'sara' or dr) else args.rows
    solvers._solve(kind, meas, param, facets=facets, callback=grab, max_total_iter=args.warmup + 1, gram=args.gram,
                   rows=rows_c)
    eng = holder['eng']
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.iterate(1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    iters_per_s = args.steps / elapsed
    log(f'{args.steps} timed iterations: {iters_per_s:.2f} iterations/s')

    # ---- live per-kernel timing (HIP events on the engine stream) ---------
    # kernels in issue order on one stream, so each duration is its own
    eng.set_concurrent(False)
    ctx.timing_reset()
    ctx.set_timing(True)
    n_timed = 10
    eng.iterate(n_timed)
    ctx.set_timing(False)
    eng.set_concurrent(True)
    rep = ctx.timing_report()
    ctx.timing_reset()
    alg = (algorithmic_bytes_dr(meas, eng.sara.s, eng.lowrank) if dr else
           algorithmic_bytes(meas, eng.sara.s, eng.lowrank))
    per_iter_ms = {k: v[0] / n_timed for k, v in rep.items()}
    dominant = max((k for k in rep if k in alg), key=lambda k: rep[k][0])
    avg_ms = rep[dominant][0] / rep[dominant][1]
    achieved = alg[dominant] / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch measured by the FETCH_SIZE / WRITE_SIZE passes
    # (copy-calibrated) for this workload, when that summary exists
    traffic = None
    tpath = os.path.join(ROOT, 'profiles', 'hs512x16_pmc_traffic.json')
    if args.config == 'hs512' and (args.size, args.channels, args.vis, args.seed) == (512, 16, 200000, 0) \
            and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get('traffic_bytes_per_launch', {}).get(dominant)
    kernels = {k: {'ms_per_launch': rep[k][0] / rep[k][1], 'launches_per_iter': rep[k][1] / n_timed,
                   'GBps': (alg[k] / (rep[k][0] / rep[k][1] * 1e-3) / 1e9) if k in alg else None}
               for k in rep}

    # ---- time to reference SNR: plain (drop-in), batched engine, warm start --
    solve = None
    cpu = None
    if args.config == 'sara256' and not args.no_solve:
        solve, cpu = sara_against_oracle(args, solvers, meas, prob, son, nu2)
    elif not args.no_solve:
        from fhs import warmstart
        prm = dict(param, track_snr=True)
        # a full batched solve (the untimed warm-up; without the oracle fixture
        # its converged SNR is the reference); the times come from
        # --solve-repeats further solves of each form, each stopped once it
        # reaches ref - 0.1 dB
        _, full = solvers._solve(kind, meas, prm, X0=prob.X0, facets=facets, time_limit=args.solve_seconds)
        ref = full['snr'][-1] if ref_fx is None else float(ref_fx['snr'][-1])
        log(f'batched solve: {full["iterations"]} iterations, {full["wall"]:.1f} s, SNR {full["snr"][-1]:.3f} dB '
            f'(reference {ref:.3f} dB)')
        tols = (0.1, 0.5, 1.0, 2.0, 3.0)

        def timed(**kw):
            return [solvers._solve(kind, meas, prm, X0=prob.X0, facets=facets, time_limit=args.solve_seconds,
                                   stop_snr=ref - min(tols), **kw)[1] for _ in range(max(1, args.solve_repeats))]

        def reach(runs, extra_s=0.0):
            out, every = {}, {}
            for d in tols:
                k = warmstart.iterations_to(runs[0]['snr'], ref - d)
                if k is None or any(len(r['time']) < k for r in runs):
                    out[f'ref-{d}dB'] = None
                    continue
                ts = [r['time'][k - 1] + extra_s for r in runs]
                out[f'ref-{d}dB'] = {'iterations': k, 's': round(float(np.median(ts)), 4)}
                every[f'ref-{d}dB'] = [round(t, 4) for t in ts]
            return out, every

        if ref_fx is not None:
            ref_def = ('converged SNR of the FP64 oracle on the same data (tests/golden/hs512x16_reference.npz: '
                       f'{int(ref_fx["iterations"])} iterations, reweights at {[int(t) for t in ref_fx["reweights"]]}); '
                       'SNR = 20 log10(||X0||/||X0 - X||)')
        else:
            ref_def = ('SNR of the batched GPU run at convergence (no FP64 oracle solve of this size; 3 reweighting '
                       'steps, reference stopping rules); SNR = 20 log10(||X0||/||X0 - X||)')
        # plain: the drop-in form (the reference-named entry points called one
        # after the other each iteration, eager, zero start); batched: the
        # engine (one hipGraph replay per iteration on three streams, stopping
        # rule in the library); warm_start: the engine from the learned start
        _, dropin = solvers._solve(kind, meas, prm, X0=prob.X0, facets=facets, time_limit=args.solve_seconds,
                                   form='dropin')
        log(f'drop-in solve: {dropin["iterations"]} iterations, {dropin["wall"]:.2f} s')
        d_reach, d_all = reach(timed(form='dropin'))
        b_reach, b_all = reach(timed())
        log('timed plain and batched solves done')
        solve = {'ref_snr_db': round(ref, 4), 'ref_definition': ref_def,
                 'gpu_final_snr_db': round(full['snr'][-1], 4),
                 'gpu_minus_ref_db': round(full['snr'][-1] - ref, 4),
                 'oracle_reweights': None if ref_fx is None else [int(t) for t in ref_fx['reweights']],
                 'oracle_iterations': None if ref_fx is None else int(ref_fx['iterations']),
                 'timing': f'median of {max(1, args.solve_repeats)} solves after a full solve of each form',
                 'plain': {'form': 'drop-in entry points per iteration (update_primal, update_dual_lowrankness, '
                                   'update_dual_sparsity, update_dual_data_fidelity), eager, zero start',
                           'iterations': dropin['iterations'], 'wall_s': round(dropin['wall'], 3),
                           'converged': dropin['converged'], 'reweights': dropin['reweights'],
                           'final_snr_db': round(dropin['snr'][-1], 4), 'reach': d_reach, 'reach_s_each': d_all},
                 'batched': {'form': 'PDFB engine: one hipGraph replay per iteration, priors on two side streams, '
                                     'stopping rule evaluated in the library, zero start',
                             'iterations': full['iterations'], 'wall_s': round(full['wall'], 3),
                             'converged': full['converged'], 'reweights': full['reweights'],
                             'final_snr_db': round(full['snr'][-1], 4), 'reach': b_reach, 'reach_s_each': b_all}}
        forms = ['batched']
        if kind != 'sara':
            # chosen: the engine with the mixed-precision Gram (matrix cores)
            eng_c, chosen = solvers._solve(kind, meas, prm, X0=prob.X0, facets=facets,
                                           time_limit=args.solve_seconds, gram=args.gram, rows=rows_c)
            mixed_err = gram_precision_errors(ctx, meas, facets, eng_c.x, args.gram) if args.gram != 'fp64' else {}
            if rows_c != 32:
                mixed_err['g_rows'] = dict(row_precision_errors(meas, rows_c), bits=rows_c)
            del eng_c
            log(f'chosen-form solve: {chosen["iterations"]} iterations, SNR {chosen["snr"][-1]:.4f} dB')
            c_reach, c_all = reach(timed(gram=args.gram, rows=rows_c))
            solve['chosen'] = {'form': f'PDFB engine as batched, with the {args.gram} Gram on the matrix cores and '
                                       f'FP{rows_c} G rows',
                               'iterations': chosen['iterations'], 'wall_s': round(chosen['wall'], 3),
                               'converged': chosen['converged'], 'reweights': chosen['reweights'],
                               'final_snr_db': round(chosen['snr'][-1], 4),
                               'final_minus_ref_db': round(chosen['snr'][-1] - ref, 4),
                               'mixed_precision_errors': mixed_err, 'reach': c_reach, 'reach_s_each': c_all}
            forms.append('chosen')
        wpath = warmstart.default_weights_path()
        if os.path.exists(wpath) and not dr:
            ws = warmstart.WarmStart.load(ctx, wpath)
            ws.predict(warmstart.dirty_cube(meas))      # one untimed call: workspaces and code objects
            torch.cuda.synchronize()
            tw = time.perf_counter()
            x0 = ws.predict(warmstart.dirty_cube(meas))
            torch.cuda.synchronize()
            t_pred = time.perf_counter() - tw
            X0d = solvers._cube(prob.X0).double()
            x0_snr = 20 * np.log10(torch.linalg.vector_norm(X0d).item() /
                                   torch.linalg.vector_norm(X0d - x0.double()).item())
            _, warm = solvers._solve(kind, meas, prm, X0=prob.X0, facets=facets, x_init=x0, weights_from_init=False,
                                     time_limit=args.solve_seconds)
            log(f'warm-start solve: {warm["iterations"]} iterations')
            w_reach, w_all = reach(timed(x_init=x0, weights_from_init=False), t_pred)
            log('timed warm-start solves done')
            solve['warm_start'] = {'form': 'PDFB engine from the learned warm start (inference time included), '
                                           'unit first weights',
                                   'network': os.path.relpath(wpath, ROOT), 'predict_s': round(t_pred, 4),
                                   'x_init_snr_db': round(x0_snr, 3),
                                   'iterations': warm['iterations'], 'final_snr_db': round(warm['snr'][-1], 4),
                                   'reach': w_reach, 'reach_s_each': w_all}
            forms.append('warm_start')
        # speed-up of each GPU form over the plain form, per tolerance
        solve['speedup_vs_plain'] = {}
        for f in forms:
            sp = {}
            for d in tols:
                key = f'ref-{d}dB'
                p_, q_ = solve['plain']['reach'][key], solve[f]['reach'][key]
                sp[key] = None if (p_ is None or q_ is None) else round(p_['s'] / q_['s'], 3)
            solve['speedup_vs_plain'][f] = sp

    # ---- CPU baseline: FP64 oracle on the host cores -----------------------
    if not args.no_cpu and args.config != 'sara256':
        from oracle.nufft import uniform_precond_weights
        from oracle.solvers import OracleProblem, pdfb
        nproc = os.cpu_count() or 1
        threads = min(16, nproc)     # the box's CPU share for one GPU
        blocks = []
        for l in range(prob.L):
            b = np.concatenate([[0], np.cumsum(prob.blocks[l])])
            blocks.append([(int(b[i]), int(b[i + 1])) for i in range(len(prob.blocks[l]))])
        pU = [uniform_precond_weights(prob.u[l], prob.v[l], 2 * prob.Ny, 2 * prob.Nx) for l in range(prob.L)]
        op = OracleProblem(prob.Ny, prob.Nx, prob.u, prob.v, prob.y, pU, blocks, prob.epsilon,
                           g_scale=[1.0 / s for s in prob.sigma_noise], threads=threads)
        sig, sig_bar = solvers.noise_levels_hs(prob.Ny, prob.Nx, prob.L, np.ones(prob.L), son)
        pr = dict(solvers.DEFAULT_PARAM)
        pr.update({'nu2': nu2, 'gamma': sig, 'gamma0': sig_bar, 'sig': sig, 'sig_bar': sig_bar, 'alpha': 1.0})
        n_cpu = 10
        tc = time.perf_counter()
        pdfb('hs', op, pr, max_iter_total=n_cpu, threads=threads)
        tcpu = time.perf_counter() - tc
        log(f'CPU baseline: {n_cpu} oracle iterations on {threads} threads in {tcpu:.1f} s')
        cpu = {'value': round(n_cpu / tcpu, 5), 'unit': 'iters/s', 'cores': threads, 'nproc': nproc,
               'kind': 'port',
               'sample': f'{n_cpu} HyperSARA iterations of the same 512x512x16 problem by the FP64 numpy oracle '
                         f'(channels on a pool of {threads} threads, BLAS threaded; 2.6x the one-core rate on '
                         f'the 16-core host), {tcpu:.1f} s'}
        if ref_fx is not None:
            from fhs import warmstart
            k = warmstart.iterations_to(list(ref_fx['snr']), float(ref_fx['snr'][-1]) - 0.1)
            if k is not None:
                cpu['time_to_ref-0.1dB_s'] = round(k * tcpu / n_cpu, 1)
                cpu['time_to_ref-0.1dB_note'] = (f'{k} oracle iterations to the reference - 0.1 dB (from the '
                                                 f'reference solve) times the measured seconds per iteration')

    out = {
        'metric': METRIC,
        'value': round(iters_per_s, 3),
        'unit': 'iters/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * elapsed / args.steps, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': f'f32 (f64 norms, {args.gram if kind != "sara" else "no"} Gram, f64 eig, fp{rows_c} G rows)',
        'data': 'synthetic',
        'config': {'workload': (f'SARA {args.size}x{args.size}, {args.vis} visibilities, 2 blocks'
                                if args.config == 'sara256' else
                                f'HyperSARA {args.size}x{args.size}x{args.channels}, {args.vis} base vis/channel, '
                                f'10-30% flags, dimensionality-reduced (holographic) form, one block/channel'
                                if dr else
                                f'HyperSARA {args.size}x{args.size}x{args.channels}, '
                                f'{args.vis} base vis/channel, 10-30% flags, 4 blocks/channel')
                   if facets is None else
                   (f'Faceted HyperSARA {args.size}x{args.size}x{args.channels} (one of C = 4 spectral sub-cubes '
                    f'of {args.size}x{args.size}x{4 * args.channels}), 6x6 nuclear-norm facets, 50% overlap, '
                    f'{args.vis} vis/channel, '
                    + (f'w-term (w in [0, {args.w_max:g}] wavelengths, field of view {args.fov} rad, '
                       f'w-projected separable rows of {meas.G[0].Ky} taps per axis)' if args.w_max else 'no w-term') + ', 4 blocks/channel'
                    if args.config == 'fhs2048' else
                    f'Faceted HyperSARA {args.size}x{args.size}x{args.channels} (one of C = 2 spectral sub-cubes '
                    f'of {args.size}x{args.size}x{2 * args.channels}), 4x4 nuclear-norm facets, 25% overlap, '
                    f'{int(args.vis * 0.6)}-{args.vis} vis/channel, 4 blocks/channel'),
                   'Mtot': int(meas.D.Mtot),
                   'nnz_G_total': int(sum(g.nnz for g in (meas.meas.G if dr else meas.G))),
                   **({'active_modes': int(sum(f.n_modes for f in meas.forms)),
                       'nnz_H_lower_total': int(sum(f.H.nnz for f in meas.forms))} if dr else {}),
                   'setup_s': round(setup_s, 2),
                   'opnorm_rel_gap_gpu_vs_fp64': opnorm_gap},
        'roofline': {'bound': 'hbm', 'kernel': dominant, 'achieved': round(achieved, 1),
                     'peak': PEAK_HBM_GBS, 'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4),
                     'traffic': traffic, 'avg_launch_ms': round(avg_ms, 5),
                     'algorithmic_bytes_per_launch': int(alg[dominant]),
                     'survey_formula_bytes_per_launch': None if dr else
                     int(sum(8 * g.nnz + 12 * g.M + 8 * meas.nufft.No for g in meas.G))},
        'kernels': kernels,
        'tuned_launch_shapes': ctx.autotune_report(),
        'solve': solve,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
