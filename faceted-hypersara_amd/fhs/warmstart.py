"""Learned warm start: back-projected dirty cube -> initial cube for PDFB.

No reference function computes this.  The reference takes an initial cube
through ``xinit`` (``imaging/imaging.m:881-891``), passed as ``varargin{1}``
to ``src/fhs/facetHyperSARA.m:384-390`` (``src/hs/hyperSARA.m`` the same);
the reweighting weights are then initialised from that cube and every dual
variable from zero (``src/fhs/fhs_initialize_dual_and_weights.m``,
``src/hs/hs_initialize_dual_*.m``).  This module produces that cube:

  dirty cube d = compute_residual_images(x = 0)   (PSF-peak normalised)
  -> groups of ``channels`` spectral channels, each scaled by 1/max(d)
  -> U-Net (``fhs.unet``: BF16 matrix-core GEMMs, FP32 accumulation)
  -> ReLU, rescale  =  x_init.

The network is trained on (dirty, converged) pairs that the plain solver
generates itself (``make_pairs``), with this project's own forward,
backward and Adam kernels.  The optimisation problem is not changed: the
warm-started run keeps the zero start's unit weights (the reference would
derive them from x_init, which lets a poor prediction move the reweighted
fixed point), so it is the plain iteration from another starting point, and
the zero start stays the fallback (``WarmStart.predict`` is never required;
``load`` raises on a missing or mismatched file).
"""
import time

import numpy as np
import torch

from .solvers import Measurement, _solve
from .synth import make_problem
from .unet import UNet

LEVELS_DIV = 8          # three 2x2 poolings: image sides padded to a multiple of 8


class WarmStart:
    """A U-Net over ``channels``-channel groups of an (L, Nx, Ny) dirty cube."""

    def __init__(self, ctx, channels=8, widths=(16, 32, 64, 128), seed=0, gain=1.0):
        self.ctx, self.channels, self.widths = ctx, channels, tuple(widths)
        self.net = UNet(ctx, cin=channels, cout=channels, widths=widths, seed=seed)
        # the network sees d / s (s = max of the group's dirty cube) and emits
        # x * gain / s: a dirty cube integrates the PSF over extended emission
        # and peaks well above the sky, so gain brings the targets to O(1)
        self.gain = float(gain)
        self.history = []

    # ---------------------------------------------------------- layout ----
    def _groups(self, L):
        c = self.channels
        ng = (L + c - 1) // c
        idx = np.minimum(np.arange(ng * c), L - 1)        # last group repeats the final channel
        return ng, torch.from_numpy(idx).to(self.ctx.device)

    @staticmethod
    def _pad(n):
        return (n + LEVELS_DIV - 1) // LEVELS_DIV * LEVELS_DIV

    def _to_net(self, cube, scale=None):
        """(L, Nx, Ny) -> network input (channels, NB*H*W) with per-group scale."""
        L, Nx, Ny = cube.shape
        ng, idx = self._groups(L)
        H, W = self._pad(Nx), self._pad(Ny)
        g = cube.index_select(0, idx).reshape(ng, self.channels, Nx, Ny)
        if scale is None:
            scale = g.amax(dim=(1, 2, 3)).clamp_min(1e-12)
        g = g / scale[:, None, None, None]
        if (H, W) != (Nx, Ny):
            g = torch.nn.functional.pad(g, (0, W - Ny, 0, H - Nx))
        return g.permute(1, 0, 2, 3).contiguous().reshape(self.channels, ng * H * W), scale, ng, H, W

    def _from_net(self, out, scale, ng, H, W, L, Nx, Ny):
        o = out.reshape(self.channels, ng, H, W).permute(1, 0, 2, 3)[:, :, :Nx, :Ny]
        o = o * (scale / self.gain)[:, None, None, None]
        return o.reshape(ng * self.channels, Nx, Ny)[:L].clamp_min(0.0).contiguous()

    # --------------------------------------------------------- inference --
    def predict(self, dirty):
        """Initial cube (L, Nx, Ny) float32 on the GPU from a dirty cube of the
        same shape (``Measurement.residual_images`` of the zero cube)."""
        L, Nx, Ny = dirty.shape
        x, scale, ng, H, W = self._to_net(dirty.float())
        out = self.net.forward(x, ng, H, W, keep=False)
        return self._from_net(out, scale, ng, H, W, L, Nx, Ny)

    # ---------------------------------------------------------- training --
    def train(self, pairs, epochs=40, batch=4, lr=1e-3, lr_final=1e-4, augment=True, seed=0, log=None):
        """Adam on the mean squared error between U(d/s) and gain x/s over
        the (dirty, target) pairs ((L, Nx, Ny) GPU tensors, L = channels).
        On the first call gain is set to the median of max(d)/max(x) over the
        pairs.  Returns the per-epoch mean losses."""
        rng = np.random.default_rng(seed)
        if not self.history:
            ratios = [d.amax().item() / max(x.amax().item(), 1e-30) for d, x in pairs]
            self.gain = float(np.median(ratios))
        xs, ts = [], []
        for d, x in pairs:
            xin, sc, ng, H, W = self._to_net(d.float())
            tin, _, _, _, _ = self._to_net(x.float(), sc / self.gain)
            xs.append(xin.reshape(self.channels, ng, H, W))
            ts.append(tin.reshape(self.channels, ng, H, W))
        X = torch.cat(xs, dim=1)          # (C, n*ng, H, W)
        T = torch.cat(ts, dim=1)
        N = X.shape[1]
        H, W = X.shape[2], X.shape[3]
        steps = epochs * ((N + batch - 1) // batch)
        k = 0
        losses = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(N)).to(X.device)
            tot, cnt = 0.0, 0
            for b0 in range(0, N, batch):
                sel = perm[b0:b0 + batch]
                xb, tb = X.index_select(1, sel), T.index_select(1, sel)
                if augment:
                    if rng.random() < 0.5:
                        xb, tb = xb.flip(2), tb.flip(2)
                    if rng.random() < 0.5:
                        xb, tb = xb.flip(3), tb.flip(3)
                    if H == W and rng.random() < 0.5:
                        xb, tb = xb.transpose(2, 3), tb.transpose(2, 3)
                nb = xb.shape[1]
                xb = xb.contiguous().reshape(self.channels, nb * H * W)
                tb = tb.contiguous().reshape(self.channels, nb * H * W)
                # cosine learning-rate decay from lr to lr_final
                cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * k / max(steps - 1, 1)))
                self.net.zero_grad()
                y = self.net.forward(xb, nb, H, W)
                loss, g = self.net.mse(y, tb)
                self.net.backward(g, nb, H, W)
                self.net.adam(lr=cur)
                tot += loss * nb
                cnt += nb
                k += 1
            losses.append(tot / cnt)
            if log:
                log(f'epoch {ep + 1}/{epochs} loss {losses[-1]:.4e}')
        self.history += losses
        return losses

    # ------------------------------------------------------- persistence --
    def state(self):
        out = {'channels': np.array(self.channels), 'widths': np.array(self.widths), 'gain': np.array(self.gain)}
        for i, layer in enumerate(self.net.layers()):
            out[f'W{i}'] = layer.W.cpu().numpy()
            out[f'b{i}'] = layer.b.cpu().numpy()
        return out

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, ctx, path):
        with np.load(path, allow_pickle=False) as z:
            ws = cls(ctx, channels=int(z['channels']), widths=tuple(int(w) for w in z['widths']),
                     gain=float(z['gain']))
            layers = ws.net.layers()
            for i, layer in enumerate(layers):
                W, b = z[f'W{i}'], z[f'b{i}']
                if W.shape != tuple(layer.W.shape) or b.shape != tuple(layer.b.shape):
                    raise ValueError(f'warm-start file {path}: layer {i} shape mismatch')
                layer.W.copy_(torch.from_numpy(W.astype(np.float32)))
                layer.b.copy_(torch.from_numpy(b.astype(np.float32)))
        return ws


# ------------------------------------------------------- pairs / evaluation --
def dirty_cube(meas):
    """Back-projected dirty cube (``src/compute_residual_images.m`` at x = 0)."""
    zero = torch.zeros((meas.L, meas.Nx, meas.Ny), dtype=torch.float32, device=meas.ctx.device)
    return meas.residual_images(zero)


def make_pairs(ctx, seeds, Ny=256, Nx=256, L=8, M_base=65536, nblocks=4, param=None, kind='hs', log=None):
    """(dirty, converged) pairs from the plain solver (zero start), one per
    seeded synthetic cube.  Returns a list of ((L, Nx, Ny), (L, Nx, Ny)) GPU
    tensors and per-cube solve records."""
    pairs, recs = [], []
    for s in seeds:
        t0 = time.perf_counter()
        P = make_problem(Ny, Nx, L, M_base, nblocks, seed=int(s))
        meas = Measurement(ctx, P)
        d = dirty_cube(meas)
        eng, info = _solve(kind, meas, param)
        pairs.append((d, eng.x.detach().clone()))
        recs.append({'seed': int(s), 'iterations': info['iterations'], 'solve_s': info['wall'],
                     'total_s': time.perf_counter() - t0})
        if log:
            log(f'pair seed {s}: {info["iterations"]} iterations, {info["wall"]:.2f} s solve, '
                f'{recs[-1]["total_s"]:.2f} s total')
        del eng, meas
    return pairs, recs


def iterations_to(snr, target):
    """1-based index of the first iterate whose SNR reaches ``target`` (None if never)."""
    hit = np.nonzero(np.asarray(snr) >= target)[0]
    return int(hit[0]) + 1 if len(hit) else None


def compare_starts(meas, ws, param=None, kind='hs', tol_db=0.1, plain=None, max_total_iter=None):
    """Iterations and wall-clock to reach (plain converged SNR - tol_db) from
    the zero start and from the warm start, on one measurement.

    ``plain``: an earlier plain run's info dict (needs 'snr') to reuse."""
    X0 = meas.problem.X0
    prm = dict(param or {})
    prm['track_snr'] = True
    if plain is None:
        _, plain = _solve(kind, meas, prm, X0=X0, max_total_iter=max_total_iter)
    ref = plain['snr'][-1]
    target = ref - tol_db
    t0 = time.perf_counter()
    x0 = ws.predict(dirty_cube(meas))
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    x0_snr = 20 * np.log10(np.linalg.norm(X0) / np.linalg.norm(
        X0 - x0.double().cpu().numpy().transpose(2, 1, 0)))

    def stop(t, eng, info):
        return not (info['snr'][-1] >= target)

    _, warm = _solve(kind, meas, prm, x_init=x0, X0=X0, callback=stop, max_total_iter=max_total_iter,
                     weights_from_init=False)
    ip, iw = iterations_to(plain['snr'], target), iterations_to(warm['snr'], target)
    return {
        'ref_snr': ref, 'target_snr': target, 'x_init_snr': x0_snr,
        'plain_iterations': ip, 'plain_s': plain['time'][ip - 1] if ip else None,
        'warm_iterations': iw, 'warm_s': (warm['time'][iw - 1] + t_pred) if iw else None,
        'predict_s': t_pred, 'plain_total_iterations': plain['iterations'],
    }


def default_weights_path():
    """The network shipped with the package (trained by ``main`` below)."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'warmstart_hs512x16.npz')


def main(argv=None):
    """Train the shipped warm-start network on the GPU:
    ``python -m fhs.warmstart --out fhs/data/warmstart_hs512x16.npz``.
    Pairs: seeded HyperSARA cubes of the 512x512x16 workload (200,000 base
    visibilities per channel, 10-30 % flags, 4 blocks, 40 dB; seeds from 1000,
    disjoint from the benchmark's), converged by the plain solver with 3
    reweighting steps.  The defaults reproduce the shipped file: 120 cubes,
    100 epochs (about 7 minutes on one MI355X)."""
    import argparse
    from .core import default_context
    ap = argparse.ArgumentParser(prog='python -m fhs.warmstart')
    ap.add_argument('--cubes', type=int, default=120)
    ap.add_argument('--first-seed', type=int, default=1000)
    ap.add_argument('--size', type=int, default=512)
    ap.add_argument('--vis', type=int, default=200000)
    ap.add_argument('--channels', type=int, default=16, help='channels per training cube (the network sees groups of 8)')
    ap.add_argument('--epochs', type=int, default=100)
    ap.add_argument('--batch', type=int, default=4)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--out', default=default_weights_path())
    a = ap.parse_args(argv)
    ctx = default_context()
    t0 = time.perf_counter()
    pairs, recs = make_pairs(ctx, range(a.first_seed, a.first_seed + a.cubes), Ny=a.size, Nx=a.size, L=a.channels,
                             M_base=a.vis, param={'reweighting_max_iter': 3},
                             log=lambda m: print(m, flush=True))
    print(f'{len(pairs)} pairs in {time.perf_counter() - t0:.1f} s '
          f'(mean {np.mean([r["iterations"] for r in recs]):.0f} PDFB iterations per cube)', flush=True)
    ws = WarmStart(ctx)
    t0 = time.perf_counter()
    losses = ws.train(pairs, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=3e-5,
                      log=lambda m: print(m, flush=True) if m.split('/')[0].endswith('0') else None)
    torch.cuda.synchronize()
    print(f'trained in {time.perf_counter() - t0:.1f} s: loss {losses[0]:.3e} -> {losses[-1]:.3e}', flush=True)
    ws.save(a.out)
    print(f'saved {a.out}', flush=True)


if __name__ == '__main__':
    main()
