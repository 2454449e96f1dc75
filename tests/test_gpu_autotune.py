"""Launch-shape autotuning (fhs_ctx_set_autotune): the tuned FFT x-passes
(rows per workgroup) give bit-identical transforms to the default shapes,
on the first (tuning) call and on later calls, and the choice is reported
per kernel and problem shape."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('N', [256, 512, 1024, 2048])
def test_tuned_fft_is_bit_identical(ctx, N):
    import fhs
    from oracle.nufft import NufftOracle
    op = NufftOracle(np.zeros(1), np.zeros(1), N, N)
    L = 3
    rng = np.random.default_rng(N)
    x = torch.from_numpy(rng.random((L, N, N)).astype(np.float32)).cuda()
    ref_ctx = fhs.Context(0)                      # autotuning off: the default shapes
    plan0 = fhs.NufftPlan(ref_ctx, N, N, op.Kty, op.Ktx, op.sy, op.sx)
    F0 = plan0.A(x)
    b0 = plan0.At_real(F0)
    torch.cuda.synchronize()
    tctx = fhs.Context(0)
    tctx.set_autotune(True)
    plan1 = fhs.NufftPlan(tctx, N, N, op.Kty, op.Ktx, op.sy, op.sx)
    for _ in range(2):                            # the tuning call, then the tuned shape
        F1 = plan1.A(x)
        b1 = plan1.At_real(F1)
        torch.cuda.synchronize()
        assert torch.equal(F1, F0) and torch.equal(b1, b0)
    rep = tctx.autotune_report()
    keys = [k for k in rep if k.startswith('fft_fwd_x/N%d' % (2 * N))]
    if 2 * N >= 1024:   # grids with more than one candidate are tuned
        assert keys and rep[keys[0]].endswith('rows'), rep
    # the transform itself is right (against the oracle's A)
    ref = op.A(x[0].cpu().numpy().astype(np.float64).T)
    got = F1[0].cpu().numpy().astype(np.complex128)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
