"""Learned warm start: a 4-level U-Net (3x3 convolutions, ReLU, BF16 weights,
FP32 accumulation) mapping back-projected dirty facet tiles to an initial
cube, trained and run with this project's own kernels (``fhs_nn_*``).

Activations are FP32 tensors in [C][N][H][W] order; every 3x3 convolution
is im2col + one BF16 matrix-core GEMM (``fhs_nn_gemm``), its gradients one
GEMM each plus a col2im gather.  torch only owns the buffers (and
concatenates skip connections).

The warm start plugs into the reference's hook for an initial cube
(``xinit``, ``imaging/imaging.m:881-891``; ``varargin{1}`` of
``src/fhs/facetHyperSARA.m:384-390``); the zero start stays the fallback and
the PDFB fixed point is unchanged: only the iterations to reach it change.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import lib
from .core import _ptr as _f


class Conv3x3:
    """One 3x3 convolution (zero padding 1) with bias and optional ReLU."""

    def __init__(self, ctx, cin, cout, relu, rng):
        self.ctx, self.cin, self.cout, self.relu = ctx, cin, cout, relu
        k = cin * 9
        w = rng.normal(0.0, np.sqrt(2.0 / k), (cout, k)).astype(np.float32)      # He init
        dev = ctx.device
        self.W = torch.from_numpy(w).to(dev)
        self.b = torch.zeros(cout, dtype=torch.float32, device=dev)
        self.dW = torch.zeros_like(self.W)
        self.db = torch.zeros_like(self.b)
        self.mW, self.vW = torch.zeros_like(self.W), torch.zeros_like(self.W)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.cache = None

    def params(self):
        return [(self.W, self.dW, self.mW, self.vW), (self.b, self.db, self.mb, self.vb)]

    def forward(self, x, NB, H, W, keep=True):
        ctx = self.ctx
        Q = NB * H * W
        col = torch.empty((self.cin * 9, Q), dtype=torch.float32, device=ctx.device)
        ctx.check(lib().fhs_nn_im2col3(ctx.h, _f(x), self.cin, NB, H, W, _f(col)), 'nn_im2col3')
        y = torch.empty((self.cout, Q), dtype=torch.float32, device=ctx.device)
        ctx.check(lib().fhs_nn_gemm(ctx.h, self.cout, Q, self.cin * 9, _f(self.W), self.cin * 9, 0, _f(col), Q, 0,
                                    _f(y), Q, 1.0, 0.0, _f(self.b), int(self.relu)), 'nn_gemm')
        self.cache = (col, y, NB, H, W) if keep else None
        return y

    def backward(self, gy):
        """gy: gradient w.r.t. the output (masked in place by the ReLU);
        accumulates dW, db; returns the gradient w.r.t. the input."""
        ctx = self.ctx
        col, y, NB, H, W = self.cache
        Q = NB * H * W
        if self.relu:
            ctx.check(lib().fhs_nn_relu_bwd(ctx.h, _f(y), _f(gy), gy.numel()), 'nn_relu_bwd')
        K = self.cin * 9
        ctx.check(lib().fhs_nn_gemm(ctx.h, self.cout, K, Q, _f(gy), Q, 0, _f(col), Q, 1,
                                    _f(self.dW), K, 1.0, 1.0, None, 0), 'nn_gemm dW')
        db = torch.empty_like(self.db)
        ctx.check(lib().fhs_nn_bias_grad(ctx.h, _f(gy), self.cout, Q, _f(db)), 'nn_bias_grad')
        self.db += db
        dcol = torch.empty((K, Q), dtype=torch.float32, device=ctx.device)
        ctx.check(lib().fhs_nn_gemm(ctx.h, K, Q, self.cout, _f(self.W), K, 1, _f(gy), Q, 0,
                                    _f(dcol), Q, 1.0, 0.0, None, 0), 'nn_gemm dcol')
        dx = torch.empty((self.cin, Q), dtype=torch.float32, device=ctx.device)
        ctx.check(lib().fhs_nn_col2im3(ctx.h, _f(dcol), self.cin, NB, H, W, _f(dx)), 'nn_col2im3')
        self.cache = None
        return dx


class UNet:
    """widths (16, 32, 64, 128): two 3x3 conv + ReLU per level, 2x2 max
    pooling down, nearest upsampling + 3x3 conv + ReLU up, concatenation with
    the skip, two 3x3 conv + ReLU, and a linear 3x3 output head."""

    def __init__(self, ctx, cin=8, cout=8, widths=(16, 32, 64, 128), seed=0):
        self.ctx, self.cin, self.cout, self.widths = ctx, cin, cout, tuple(widths)
        rng = np.random.default_rng(seed)
        w = self.widths
        self.enc = []
        c = cin
        for wi in w:
            self.enc.append((Conv3x3(ctx, c, wi, True, rng), Conv3x3(ctx, wi, wi, True, rng)))
            c = wi
        self.dec = []            # dec[j] works at level len(w) - 2 - j
        for i in range(len(w) - 2, -1, -1):
            self.dec.append((Conv3x3(ctx, w[i + 1], w[i], True, rng), Conv3x3(ctx, 2 * w[i], w[i], True, rng),
                             Conv3x3(ctx, w[i], w[i], True, rng)))
        self.head = Conv3x3(ctx, w[0], cout, False, rng)
        self.step_count = 0
        self._skips = None

    def layers(self):
        out = []
        for a, b in self.enc:
            out += [a, b]
        for u, a, b in self.dec:
            out += [u, a, b]
        return out + [self.head]

    def _new(self, C, Q):
        return torch.empty((C, Q), dtype=torch.float32, device=self.ctx.device)

    def _pool(self, x, C, NB, H, W):
        y = self._new(C, NB * (H // 2) * (W // 2))
        self.ctx.check(lib().fhs_nn_maxpool2(self.ctx.h, _f(x), C * NB, H, W, _f(y)), 'nn_maxpool2')
        return y

    def _pool_bwd(self, x, gy, C, NB, H, W):
        gx = self._new(C, NB * H * W)
        self.ctx.check(lib().fhs_nn_maxpool2_bwd(self.ctx.h, _f(x), _f(gy), C * NB, H, W, _f(gx)), 'nn_maxpool2_bwd')
        return gx

    def _up(self, x, C, NB, H, W):
        y = self._new(C, NB * 4 * H * W)
        self.ctx.check(lib().fhs_nn_upsample2(self.ctx.h, _f(x), C * NB, H, W, _f(y)), 'nn_upsample2')
        return y

    def _up_bwd(self, gy, C, NB, H, W):
        gx = self._new(C, NB * H * W)
        self.ctx.check(lib().fhs_nn_upsample2_bwd(self.ctx.h, _f(gy), C * NB, H, W, _f(gx)), 'nn_upsample2_bwd')
        return gx

    def forward(self, x, NB, H, W, keep=True):
        """x: (cin, NB*H*W) float32 ([C][N][H][W]) -> (cout, NB*H*W).
        H and W must be divisible by 2^(levels-1)."""
        h, Hc, Wc = x.contiguous(), H, W
        skips = []
        for i, (a, b) in enumerate(self.enc):
            h = b.forward(a.forward(h, NB, Hc, Wc, keep), NB, Hc, Wc, keep)
            if i < len(self.enc) - 1:
                skips.append(h)
                h = self._pool(h, self.widths[i], NB, Hc, Wc)
                Hc, Wc = Hc // 2, Wc // 2
        for j, (u, a, b) in enumerate(self.dec):
            i = len(self.widths) - 2 - j
            hu = self._up(h, self.widths[i + 1], NB, Hc, Wc)
            Hc, Wc = Hc * 2, Wc * 2
            hu = u.forward(hu, NB, Hc, Wc, keep)
            cat = torch.cat([hu, skips[i]], dim=0)
            h = b.forward(a.forward(cat, NB, Hc, Wc, keep), NB, Hc, Wc, keep)
        self._skips = skips if keep else None
        return self.head.forward(h, NB, Hc, Wc, keep)

    def backward(self, g, NB, H, W):
        """g: gradient of the loss w.r.t. the output; accumulates the
        parameter gradients and returns the gradient w.r.t. the input."""
        w = self.widths
        g = self.head.backward(g.contiguous())
        skip_grads = [None] * (len(w) - 1)
        for j in range(len(self.dec) - 1, -1, -1):          # finest decoder stage first
            u, a, b = self.dec[j]
            i = len(w) - 2 - j
            g = a.backward(b.backward(g))                     # gradient of [up ; skip_i]
            skip_grads[i] = g[w[i]:].contiguous()
            gu = u.backward(g[:w[i]].contiguous())
            g = self._up_bwd(gu, w[i + 1], NB, H >> (i + 1), W >> (i + 1))
        for i in range(len(self.enc) - 1, -1, -1):
            a, b = self.enc[i]
            if i < len(self.enc) - 1:                         # pooled path + skip path
                g = self._pool_bwd(self._skips[i], g.contiguous(), w[i], NB, H >> i, W >> i) + skip_grads[i]
            g = a.backward(b.backward(g.contiguous()))
        self._skips = None
        return g

    def zero_grad(self):
        for layer in self.layers():
            layer.dW.zero_()
            layer.db.zero_()

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        for layer in self.layers():
            for p, g, m, v in layer.params():
                self.ctx.check(lib().fhs_nn_adam(self.ctx.h, _f(p), _f(g), _f(m), _f(v), p.numel(), lr, beta1,
                                                 beta2, eps, self.step_count), 'nn_adam')

    def mse(self, pred, target, grad=True):
        """mean squared error and its gradient (2/n)(pred - target)."""
        n = pred.numel()
        g = torch.empty_like(pred) if grad else None
        loss = C.c_double()
        self.ctx.check(lib().fhs_nn_mse(self.ctx.h, _f(pred), _f(target), n, 2.0 / n, _f(g), C.byref(loss)),
                       'nn_mse')
        return loss.value / n, g
