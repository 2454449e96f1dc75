"""Oracle: the warm-start U-Net in numpy float64 (forward and backward).

Test infrastructure only (see ``oracle/__init__.py``).  No reference
counterpart (the learned warm start has no MATLAB function); pinned by
finite differences of its own backward pass and by textbook identities of
each layer.  ``bf16=True`` rounds every GEMM operand to bfloat16 (round to
nearest even) exactly where the GPU kernels do, so the GPU can be compared
at FP32-accumulation precision.
"""
import numpy as np


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def im2col3(x, NB, H, W):
    """x (C, NB*H*W) -> (C*9, NB*H*W), zero padding 1."""
    C = x.shape[0]
    xi = x.reshape(C, NB, H, W)
    p = np.zeros((C, NB, H + 2, W + 2))
    p[:, :, 1:-1, 1:-1] = xi
    cols = np.zeros((C, 9, NB, H, W))
    for ky in range(3):
        for kx in range(3):
            cols[:, ky * 3 + kx] = p[:, :, ky:ky + H, kx:kx + W]
    return cols.reshape(C * 9, NB * H * W)


def col2im3(dcol, C, NB, H, W):
    """adjoint of im2col3."""
    d = dcol.reshape(C, 9, NB, H, W)
    p = np.zeros((C, NB, H + 2, W + 2))
    for ky in range(3):
        for kx in range(3):
            p[:, :, ky:ky + H, kx:kx + W] += d[:, ky * 3 + kx]
    return p[:, :, 1:-1, 1:-1].reshape(C, NB * H * W)


def maxpool2(x, NB, H, W):
    C = x.shape[0]
    v = x.reshape(C, NB, H // 2, 2, W // 2, 2).transpose(0, 1, 2, 4, 3, 5).reshape(C, NB, H // 2, W // 2, 4)
    # window order (0,0), (0,1), (1,0), (1,1): first maximum wins
    return v.max(axis=-1).reshape(C, -1), v.argmax(axis=-1)


def maxpool2_bwd(arg, gy, C, NB, H, W):
    g = np.zeros((C, NB, H // 2, W // 2, 4))
    np.put_along_axis(g, arg[..., None], gy.reshape(C, NB, H // 2, W // 2, 1), axis=-1)
    return g.reshape(C, NB, H // 2, W // 2, 2, 2).transpose(0, 1, 2, 4, 3, 5).reshape(C, -1)


def upsample2(x, C, NB, H, W):
    return np.repeat(np.repeat(x.reshape(C, NB, H, W), 2, axis=2), 2, axis=3).reshape(C, -1)


def upsample2_bwd(g, C, NB, H, W):
    return g.reshape(C, NB, H, 2, W, 2).sum(axis=(3, 5)).reshape(C, -1)


class OracleUNet:
    """Same parameters and structure as ``fhs.unet.UNet`` (weights copied in)."""

    def __init__(self, layers_np, widths, bf=False):
        # layers_np: list of (W (cout, cin*9), b (cout,), relu) in fhs.unet.UNet.layers() order
        self.layers = [dict(W=np.asarray(W, np.float64), b=np.asarray(b, np.float64), relu=r) for W, b, r in layers_np]
        self.widths = tuple(widths)
        self.bf = bf

    def _r(self, a):
        return bf16(a) if self.bf else a

    def conv(self, l, x, NB, H, W):
        col = im2col3(x, NB, H, W)
        y = self._r(l['W']) @ self._r(col) + l['b'][:, None]
        if l['relu']:
            y = np.maximum(y, 0)
        l['cache'] = (col, y, NB, H, W)
        return y

    def conv_bwd(self, l, gy):
        col, y, NB, H, W = l['cache']
        if l['relu']:
            gy = gy * (y > 0)
        l['dW'] = self._r(gy) @ self._r(col).T
        l['db'] = gy.sum(axis=1)
        dcol = self._r(l['W']).T @ self._r(gy)
        return col2im3(dcol, l['W'].shape[1] // 9, NB, H, W)

    def forward(self, x, NB, H, W):
        w = self.widths
        L = self.layers
        k = 0
        h, Hc, Wc = x, H, W
        self.skips, self.args = [], []
        for i in range(len(w)):
            h = self.conv(L[k], h, NB, Hc, Wc); k += 1
            h = self.conv(L[k], h, NB, Hc, Wc); k += 1
            if i < len(w) - 1:
                self.skips.append(h)
                h, a = maxpool2(h, NB, Hc, Wc)
                self.args.append(a)
                Hc, Wc = Hc // 2, Wc // 2
        for j in range(len(w) - 1):
            i = len(w) - 2 - j
            hu = upsample2(h, w[i + 1], NB, Hc, Wc)
            Hc, Wc = Hc * 2, Wc * 2
            hu = self.conv(L[k], hu, NB, Hc, Wc); k += 1
            h = np.concatenate([hu, self.skips[i]], axis=0)
            h = self.conv(L[k], h, NB, Hc, Wc); k += 1
            h = self.conv(L[k], h, NB, Hc, Wc); k += 1
        return self.conv(L[k], h, NB, Hc, Wc)

    def backward(self, g, NB, H, W):
        w = self.widths
        L = self.layers
        ne = 2 * len(w)
        g = self.conv_bwd(L[-1], g)
        sg = [None] * (len(w) - 1)
        for j in range(len(w) - 2, -1, -1):
            i = len(w) - 2 - j
            base = ne + 3 * j
            g = self.conv_bwd(L[base + 1], self.conv_bwd(L[base + 2], g))
            sg[i] = g[w[i]:]
            gu = self.conv_bwd(L[base], g[:w[i]])
            g = upsample2_bwd(gu, w[i + 1], NB, H >> (i + 1), W >> (i + 1))
        for i in range(len(w) - 1, -1, -1):
            if i < len(w) - 1:
                g = maxpool2_bwd(self.args[i], g, w[i], NB, H >> i, W >> i) + sg[i]
            g = self.conv_bwd(L[2 * i], self.conv_bwd(L[2 * i + 1], g))
        return g
