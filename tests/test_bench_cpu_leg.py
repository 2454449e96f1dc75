"""The benchmark's CPU baseline worker (bench._cube_cpu_worker): a pool of
worker processes forked before any GPU use, each building the sampled rows'
G and timing one channel's transforms, rows, wavelets and one facet's prox
with the FP64 oracle, all workers released together by a barrier.  Run
here on a small payload (64 x 64 image, 300 rows of 15 x 15 taps)."""
import multiprocessing as mp

import numpy as np


def test_cube_cpu_worker_pool():
    import bench
    N, Kty, Ktx, m, K = 64, 128, 128, 300, 15
    rng = np.random.default_rng(0)
    d = {'N': N, 'Kty': Kty, 'Ktx': Ktx, 'a': rng.integers(0, Kty * Ktx, m).astype(np.int32),
         'cy': (rng.standard_normal((m, K)) + 1j * rng.standard_normal((m, K))).astype(np.complex64),
         'cx': (rng.standard_normal((m, K)) + 1j * rng.standard_normal((m, K))).astype(np.complex64),
         'facet_pixels': 1024, 'facet_channels': 4}
    workers = 2
    fork = mp.get_context('fork')
    with fork.Pool(workers, initializer=bench._cpu_pool_init, initargs=(fork.Barrier(workers),)) as pool:
        res = np.array(pool.map(bench._cube_cpu_worker, [(d, s) for s in range(workers)], chunksize=1))
    assert res.shape == (workers, 4)
    assert np.all(np.isfinite(res)) and np.all(res > 0)
