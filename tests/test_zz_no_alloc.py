"""fhs_proj_l2ball_noalloc (no device allocation; its scratch is a fixed
per-device array) against fhs_proj_l2ball, through the C ABI: bit-identical
outputs on two calls in a row, the second with a smaller input, and the
cases of the l2 ball (outside, inside, x = y)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(ctx, name, x, eps, y):
    from fhs import _lib
    p = torch.empty_like(x)
    st = getattr(_lib.lib(), name)(ctx.h, x.data_ptr(), float(eps), y.data_ptr(), x.numel(), p.data_ptr())
    ctx.check(st, name)
    torch.cuda.synchronize()
    return p


def test_noalloc_proj_l2ball_matches_allocating_entry(ctx):
    g = torch.Generator().manual_seed(0)
    cases = []
    for n, scale in ((3_000_001, 1.0), (70_001, 1.0), (1_000, 1e-3), (257, 0.0)):
        x = torch.complex(torch.randn(n, generator=g), torch.randn(n, generator=g)).to(torch.complex64).cuda()
        y = (x + scale * torch.complex(torch.randn(n, generator=g), torch.randn(n, generator=g))
             .to(torch.complex64).cuda()) if scale > 0 else x.clone()
        d = torch.linalg.vector_norm((x - y).to(torch.complex128)).item()
        for eps in (0.5 * d, 2.0 * d + 1e-3):
            cases.append((x, eps, y))
    for x, eps, y in cases:        # consecutive calls, sizes shrinking and growing
        a = _call(ctx, 'fhs_proj_l2ball', x, eps, y)
        b = _call(ctx, 'fhs_proj_l2ball_noalloc', x, eps, y)
        assert torch.equal(torch.view_as_real(a), torch.view_as_real(b))
        # and the ball itself: ||p - y|| = min(eps, ||x - y||)
        r = torch.linalg.vector_norm((b - y).to(torch.complex128)).item()
        d = torch.linalg.vector_norm((x - y).to(torch.complex128)).item()
        assert r == pytest.approx(min(eps, d), rel=1e-5, abs=1e-6)


def test_noalloc_rejects_bad_arguments(ctx):
    from fhs import _lib
    x = torch.zeros(8, dtype=torch.complex64, device='cuda')
    assert _lib.lib().fhs_proj_l2ball_noalloc(ctx.h, x.data_ptr(), -1.0, x.data_ptr(), 8, x.data_ptr()) != 0
    assert _lib.lib().fhs_proj_l2ball_noalloc(ctx.h, x.data_ptr(), 1.0, x.data_ptr(), -1, x.data_ptr()) != 0
    assert _lib.lib().fhs_proj_l2ball_noalloc(ctx.h, x.data_ptr(), 1.0, x.data_ptr(), 0, x.data_ptr()) == 0
