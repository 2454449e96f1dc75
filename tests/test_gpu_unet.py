"""GPU kernels of the learned warm start against the FP64 U-Net oracle.

Operands are rounded to BF16 in both (the oracle rounds exactly where the
GEMM stages them), so single layers agree at FP32-accumulation precision
(1e-4); the whole network, where a 1-ulp BF16 flip of an activation
propagates, within 2e-2."""
import numpy as np
import pytest
import torch

from oracle.unet import OracleUNet, bf16, col2im3, im2col3, maxpool2, upsample2

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize('M,N,K', [(16, 16, 16), (37, 129, 70), (128, 4096, 1152), (200, 64, 33),
                                   (16, 72, 20000), (40, 300, 8200)])   # the last two split K
@pytest.mark.parametrize('ta,tb', [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_bf16(ctx, M, N, K, ta, tb):
    import ctypes as C
    from fhs._lib import lib
    from fhs.core import _ptr
    rng = np.random.default_rng(M + N + K + 2 * ta + tb)
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    bias = rng.normal(size=M).astype(np.float32)
    Cd = torch.zeros((M, N), dtype=torch.float32, device='cuda')
    Ad, Bd, bd = dev(A), dev(B), dev(bias)     # held until the kernel has run
    ctx.check(lib().fhs_nn_gemm(ctx.h, M, N, K, _ptr(Ad), M if ta else K, ta, _ptr(Bd), K if tb else N, tb,
                                _ptr(Cd), N, 1.0, 0.0, _ptr(bd), 0), 'gemm')
    torch.cuda.synchronize()
    opA = bf16(A).T if ta else bf16(A)
    opB = bf16(B).T if tb else bf16(B)
    ref = opA @ opB + bias[:, None]
    assert rel(Cd.cpu().numpy(), ref) < 1e-5
    # beta = 1 accumulates into C (the weight-gradient use)
    ctx.check(lib().fhs_nn_gemm(ctx.h, M, N, K, _ptr(Ad), M if ta else K, ta, _ptr(Bd), K if tb else N, tb,
                                _ptr(Cd), N, 1.0, 1.0, None, 0), 'gemm beta')
    assert rel(Cd.cpu().numpy(), 2 * (opA @ opB) + bias[:, None]) < 1e-5


def test_layer_primitives(ctx):
    import ctypes as C
    from fhs._lib import lib
    from fhs.core import _ptr
    rng = np.random.default_rng(1)
    Cc, NB, H, W = 5, 2, 12, 10
    x = rng.normal(size=(Cc, NB * H * W)).astype(np.float32)
    xd = dev(x)
    col = torch.empty((Cc * 9, NB * H * W), dtype=torch.float32, device='cuda')
    ctx.check(lib().fhs_nn_im2col3(ctx.h, _ptr(xd), Cc, NB, H, W, _ptr(col)), 'im2col')
    assert rel(col.cpu().numpy(), im2col3(x.astype(np.float64), NB, H, W)) < 1e-7
    dcol = rng.normal(size=(Cc * 9, NB * H * W)).astype(np.float32)
    dcold = dev(dcol)
    din = torch.empty((Cc, NB * H * W), dtype=torch.float32, device='cuda')
    ctx.check(lib().fhs_nn_col2im3(ctx.h, _ptr(dcold), Cc, NB, H, W, _ptr(din)), 'col2im')
    assert rel(din.cpu().numpy(), col2im3(dcol.astype(np.float64), Cc, NB, H, W)) < 1e-6
    p = torch.empty((Cc, NB * H * W // 4), dtype=torch.float32, device='cuda')
    ctx.check(lib().fhs_nn_maxpool2(ctx.h, _ptr(xd), Cc * NB, H, W, _ptr(p)), 'pool')
    assert rel(p.cpu().numpy(), maxpool2(x.astype(np.float64), NB, H, W)[0]) == 0
    u = torch.empty((Cc, NB * H * W * 4), dtype=torch.float32, device='cuda')
    ctx.check(lib().fhs_nn_upsample2(ctx.h, _ptr(xd), Cc * NB, H, W, _ptr(u)), 'up')
    assert rel(u.cpu().numpy(), upsample2(x.astype(np.float64), Cc, NB, H, W)) == 0


def _oracle_of(net, bf):
    layers = [(l.W.cpu().numpy(), l.b.cpu().numpy(), l.relu) for l in net.layers()]
    return OracleUNet(layers, net.widths, bf=bf)


@pytest.mark.parametrize('widths,NB,H,W', [((4, 6, 8, 10), 2, 16, 16), ((16, 32, 64, 128), 1, 32, 48)])
def test_unet_forward_backward(ctx, widths, NB, H, W):
    from fhs.unet import UNet
    rng = np.random.default_rng(len(widths) + H)
    net = UNet(ctx, cin=8, cout=8, widths=widths, seed=3)
    x = rng.normal(size=(8, NB * H * W)).astype(np.float32)
    t = rng.normal(size=(8, NB * H * W)).astype(np.float32)
    y = net.forward(dev(x), NB, H, W)
    loss, g = net.mse(y, dev(t))
    net.zero_grad()
    gx = net.backward(g, NB, H, W)
    orc = _oracle_of(net, bf=True)
    yo = orc.forward(x.astype(np.float64), NB, H, W)
    assert rel(y.cpu().numpy(), yo) < 2e-2
    go = orc.backward(2.0 / y.numel() * (yo - t), NB, H, W)
    assert rel(gx.cpu().numpy(), go) < 2e-2
    for lg, lo in zip(net.layers(), orc.layers):
        assert rel(lg.dW.cpu().numpy(), lo['dW']) < 2e-2
        assert rel(lg.db.cpu().numpy(), lo['db']) < 2e-2


def test_unet_training_reduces_loss(ctx):
    from fhs.unet import UNet
    rng = np.random.default_rng(5)
    net = UNet(ctx, cin=8, cout=8, widths=(8, 16, 32, 64), seed=1)
    NB, H, W = 2, 32, 32
    x = rng.random((8, NB * H * W)).astype(np.float32)
    t = (0.5 * x + 0.1).astype(np.float32)
    xd, td = dev(x), dev(t)
    first = None
    for step in range(60):
        net.zero_grad()
        y = net.forward(xd, NB, H, W)
        loss, g = net.mse(y, td)
        first = loss if first is None else first
        net.backward(g, NB, H, W)
        net.adam(lr=2e-3)
    assert loss < 0.1 * first
