"""Pins of the U-Net oracle (the learned warm start has no reference
counterpart): each layer against an independent formulation, adjoint
identities of the data-movement layers, and the whole backward pass
against central finite differences."""
import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle.unet import (OracleUNet, bf16, col2im3, im2col3, maxpool2, maxpool2_bwd, upsample2,
                         upsample2_bwd)


def test_bf16_rounding():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: round to even keeps 1
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38], np.float32)
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == -2.5
    assert np.isfinite(r[4])
    y = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(bf16(y) - y) / np.abs(y)) <= 2.0 ** -8


def test_conv_is_correlation():
    rng = np.random.default_rng(1)
    Ci, Co, NB, H, W = 3, 2, 2, 7, 5
    x = rng.normal(size=(Ci, NB * H * W))
    Wt = rng.normal(size=(Co, Ci * 9))
    y = Wt @ im2col3(x, NB, H, W)
    xr = x.reshape(Ci, NB, H, W)
    k = Wt.reshape(Co, Ci, 3, 3)
    ref = np.zeros((Co, NB, H, W))
    for co in range(Co):
        for n in range(NB):
            for ci in range(Ci):
                ref[co, n] += correlate2d(xr[ci, n], k[co, ci], mode='same')
    assert np.allclose(y.reshape(Co, NB, H, W), ref, atol=1e-12)


@pytest.mark.parametrize('layer', ['im2col', 'pool', 'up'])
def test_adjoint_identities(layer):
    rng = np.random.default_rng(2)
    C, NB, H, W = 3, 2, 6, 8
    x = rng.normal(size=(C, NB * H * W))
    if layer == 'im2col':
        y = rng.normal(size=(C * 9, NB * H * W))
        lhs, rhs = np.vdot(im2col3(x, NB, H, W), y), np.vdot(x, col2im3(y, C, NB, H, W))
    elif layer == 'up':
        y = rng.normal(size=(C, NB * H * W * 4))
        lhs, rhs = np.vdot(upsample2(x, C, NB, H, W), y), np.vdot(x, upsample2_bwd(y, C, NB, H, W))
    else:
        p, arg = maxpool2(x, NB, H, W)
        g = rng.normal(size=p.shape)
        # the pooled value is the input at the selected position: <p, g> = <x, P^T g>
        lhs, rhs = np.vdot(p, g), np.vdot(x, maxpool2_bwd(arg, g, C, NB, H, W))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_maxpool_first_max_wins():
    x = np.array([[2.0, 2.0, 1.0, 5.0, 0.0, 0.0, 5.0, 1.0]])   # one 2x4 plane
    p, arg = maxpool2(x, 1, 2, 4)
    assert np.array_equal(p, [[2.0, 5.0]])
    g = maxpool2_bwd(arg, np.array([[1.0, 1.0]]), 1, 1, 2, 4)
    assert np.array_equal(g, [[1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0]])


def _net(widths, cin, cout, seed):
    rng = np.random.default_rng(seed)
    layers = []
    c = cin
    for w in widths:
        layers += [(rng.normal(size=(w, c * 9)) * 0.4, rng.normal(size=w) * 0.1, True),
                   (rng.normal(size=(w, w * 9)) * 0.3, rng.normal(size=w) * 0.1, True)]
        c = w
    for i in range(len(widths) - 2, -1, -1):
        layers += [(rng.normal(size=(widths[i], widths[i + 1] * 9)) * 0.3, rng.normal(size=widths[i]) * 0.1, True),
                   (rng.normal(size=(widths[i], 2 * widths[i] * 9)) * 0.3, rng.normal(size=widths[i]) * 0.1, True),
                   (rng.normal(size=(widths[i], widths[i] * 9)) * 0.3, rng.normal(size=widths[i]) * 0.1, True)]
    layers.append((rng.normal(size=(cout, widths[0] * 9)) * 0.3, np.zeros(cout), False))
    return OracleUNet(layers, widths)


def test_backward_finite_differences():
    widths, cin, cout, NB, H, W = (2, 3, 4), 2, 2, 1, 8, 8
    net = _net(widths, cin, cout, 3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(cin, NB * H * W))
    t = rng.normal(size=(cout, NB * H * W))

    def loss(xx):
        return 0.5 * np.sum((net.forward(xx, NB, H, W) - t) ** 2)

    y = net.forward(x, NB, H, W)
    gx = net.backward(y - t, NB, H, W)
    h = 1e-6
    for idx in [(0, 0), (1, 17), (0, 40), (1, 63)]:
        e = np.zeros_like(x); e[idx] = h
        fd = (loss(x + e) - loss(x - e)) / (2 * h)
        assert abs(fd - gx[idx]) < 1e-6 * max(1.0, abs(fd))
    # parameter gradients (recompute the cache at x first)
    net.forward(x, NB, H, W); net.backward(y - t, NB, H, W)
    for li, (r, c) in [(0, (1, 5)), (3, (0, 2)), (len(net.layers) - 1, (1, 7)), (7, (2, 30))]:
        l = net.layers[li]
        dW = l['dW'][r, c]
        W0 = l['W'][r, c]
        l['W'][r, c] = W0 + h; lp = loss(x)
        l['W'][r, c] = W0 - h; lm = loss(x)
        l['W'][r, c] = W0
        fd = (lp - lm) / (2 * h)
        assert abs(fd - dW) < 1e-6 * max(1.0, abs(fd))
