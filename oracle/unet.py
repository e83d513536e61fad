"""FP64 numpy restatement of the LSRTM U-Net preconditioner (forward, loss,
backward) used as the checker of the library's BF16-MFMA network.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by __graft_entry__.smoke();
the product library never calls it.

The reference has no learned component, so there is no reference file to
follow; the network is defined by include/erm_unet.h:
  level l = 0..3 with channels c[l]; encoder: two 3x3 conv+ReLU per level,
  2x2 average pooling between levels; decoder: nearest 2x upsampling of the
  level below, concatenation [upsampled, skip] along channels, two 3x3
  conv+ReLU; output 3x3 conv (no ReLU) plus the input as a residual.
Parameter vector: per layer, weights [cout][cin][3][3] then biases [cout],
layers in forward order (15 layers).

Pinning: the 3x3 convolution is checked against direct loops, the whole
backward against central finite differences in FP64 and against
torch.autograd in float64 (tests/test_oracle_unet.py).  The optimizer
(`adam`, Kingma & Ba 2015 Algorithm 1 with the constants of
include/erm_unet.h) is pinned by two closed forms: a constant gradient moves
every parameter by lr*sign(g) per step, and g followed by -g moves it by
-lr*sign(g) and then +lr*sign(g)/19.

`bf16=True` rounds every convolution operand (weights, inputs, output
gradients) to bfloat16 with round-to-nearest-even, the rounding the MFMA
path applies, so the GPU can be compared at FP32-accumulation accuracy.
"""
import numpy as np


def bf16_round(a):
    """Round float values to the nearest bfloat16 (ties to even); float64 out."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def layer_specs(c):
    c1, c2, c3, c4 = c
    # (cin, cout, relu) in forward order
    return [(2, c1, 1), (c1, c1, 1), (c1, c2, 1), (c2, c2, 1), (c2, c3, 1), (c3, c3, 1), (c3, c4, 1), (c4, c4, 1),
            (c4 + c3, c3, 1), (c3, c3, 1), (c3 + c2, c2, 1), (c2, c2, 1), (c2 + c1, c1, 1), (c1, c1, 1), (c1, 2, 0)]


def num_params(c):
    return sum(ci * co * 9 + co for ci, co, _ in layer_specs(c))


def unpack(p, c):
    """Split the parameter vector into [(W [co][ci][3][3], b [co]), ...]."""
    out, off = [], 0
    for ci, co, _ in layer_specs(c):
        w = np.asarray(p[off:off + ci * co * 9], np.float64).reshape(co, ci, 3, 3)
        off += ci * co * 9
        b = np.asarray(p[off:off + co], np.float64)
        off += co
        out.append((w, b))
    assert off == len(p)
    return out


def pack(grads):
    return np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])


def _taps(x):
    """x [n][c][h][w] -> [9][n][c][h][w], tap t = 3*ky + kx reads x[y+ky-1][x+kx-1] (zero padded)."""
    n, c, h, w = x.shape
    xp = np.zeros((n, c, h + 2, w + 2), x.dtype)
    xp[:, :, 1:-1, 1:-1] = x
    return np.stack([xp[:, :, ky:ky + h, kx:kx + w] for ky in range(3) for kx in range(3)])


def conv3x3(x, w, b):
    """y[n][co][y][x] = b[co] + sum_{ci,ky,kx} w[co][ci][ky][kx] x[n][ci][y+ky-1][x+kx-1]."""
    t = _taps(x)
    return np.einsum("oik,knihw->nohw", w.reshape(w.shape[0], w.shape[1], 9), t) + b[None, :, None, None]


def conv3x3_backward(x, w, dy):
    """Gradients of conv3x3 given dL/dy: (dx, dw, db)."""
    co, ci = w.shape[:2]
    t = _taps(x)
    dw = np.einsum("nohw,knihw->oik", dy, t).reshape(co, ci, 3, 3)
    db = dy.sum(axis=(0, 2, 3))
    # dx[r] = sum_t w[t] dy[r - o(t)] = correlation of dy with the flipped kernel
    wf = w[:, :, ::-1, ::-1]
    td = _taps(dy)
    dx = np.einsum("oik,knohw->nihw", wf.reshape(co, ci, 9), td)
    return dx, dw, db


def pool(x):
    return 0.25 * (x[:, :, 0::2, 0::2] + x[:, :, 1::2, 0::2] + x[:, :, 0::2, 1::2] + x[:, :, 1::2, 1::2])


def pool_backward(d):
    return 0.25 * np.repeat(np.repeat(d, 2, axis=2), 2, axis=3)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up_backward(d):
    return d[:, :, 0::2, 0::2] + d[:, :, 1::2, 0::2] + d[:, :, 0::2, 1::2] + d[:, :, 1::2, 1::2]


class UNetOracle:
    def __init__(self, channels, params, bf16=False):
        self.c = tuple(channels)
        self.specs = layer_specs(self.c)
        self.layers = unpack(np.asarray(params, np.float64), self.c)
        self.bf16 = bf16

    def _q(self, a):
        return bf16_round(a) if self.bf16 else a

    def _conv(self, li, x, cache):
        w, b = self.layers[li]
        xq, wq = self._q(x), self._q(w)
        y = conv3x3(xq, wq, b)
        if self.specs[li][2]:
            y = np.maximum(y, 0.0)
        cache[li] = (xq, wq, y)
        return y

    def forward(self, x, cache=None):
        cache = {} if cache is None else cache
        x = np.asarray(x, np.float64)
        cv = lambda li, a: self._conv(li, a, cache)
        e1 = cv(1, cv(0, x))
        e2 = cv(3, cv(2, pool(e1)))
        e3 = cv(5, cv(4, pool(e2)))
        bb = cv(7, cv(6, pool(e3)))
        d3 = cv(9, cv(8, np.concatenate([up(bb), e3], axis=1)))
        d2 = cv(11, cv(10, np.concatenate([up(d3), e2], axis=1)))
        d1 = cv(13, cv(12, np.concatenate([up(d2), e1], axis=1)))
        return cv(14, d1) + x

    def _conv_bwd(self, li, dy, cache, grads):
        xq, wq, y = cache[li]
        if self.specs[li][2]:
            dy = np.where(y > 0, dy, 0.0)
        dyq = self._q(dy)
        dx, dw, db = conv3x3_backward(xq, wq, dyq)
        grads[li] = (dw, db)
        return dx

    def loss_grad(self, x, t):
        """loss = mean((net(x) - t)^2) and d loss / d params (packed)."""
        cache = {}
        y = self.forward(x, cache)
        t = np.asarray(t, np.float64)
        loss = np.mean((y - t) ** 2)
        g = {}
        bw = lambda li, d: self._conv_bwd(li, d, cache, g)
        c1, c2, c3, c4 = self.c
        dy = 2.0 * (y - t) / y.size
        dd1 = bw(13, bw(14, dy))
        dc1 = bw(12, dd1)
        dd2, de1 = up_backward(dc1[:, :c2]), dc1[:, c2:]
        dc2 = bw(10, bw(11, dd2))
        dd3, de2 = up_backward(dc2[:, :c3]), dc2[:, c3:]
        dc3 = bw(8, bw(9, dd3))
        dbb, de3 = up_backward(dc3[:, :c4]), dc3[:, c4:]
        dp3 = bw(6, bw(7, dbb))
        de3 = de3 + pool_backward(dp3)
        dp2 = bw(4, bw(5, de3))
        de2 = de2 + pool_backward(dp2)
        dp1 = bw(2, bw(3, de2))
        de1 = de1 + pool_backward(dp1)
        bw(0, bw(1, de1))
        return loss, pack([g[i] for i in range(len(self.specs))])


def init_params(channels, seed):
    """He-normal weights, zero biases, output layer at 1% scale (same recipe as
    the library; values differ, the generators are not the same)."""
    rng = np.random.default_rng(seed)
    parts = []
    specs = layer_specs(channels)
    for li, (ci, co, _) in enumerate(specs):
        sd = np.sqrt(2.0 / (ci * 9.0)) * (0.01 if li == len(specs) - 1 else 1.0)
        parts += [rng.normal(0, sd, ci * co * 9), np.zeros(co)]
    return np.concatenate(parts).astype(np.float32)


def adam(p0, grads, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """Parameters after one Adam step per gradient in `grads`, starting from
    p0 with zero moments (Kingma & Ba 2015, Algorithm 1, bias-corrected
    moments; the constants of erm_unet_train_step), in FP64."""
    p = np.asarray(p0, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t, g in enumerate(grads, 1):
        g = np.asarray(g, np.float64)
        m = beta1 * m + (1.0 - beta1) * g
        v = beta2 * v + (1.0 - beta2) * g * g
        p -= lr * (m / (1.0 - beta1 ** t)) / (np.sqrt(v / (1.0 - beta2 ** t)) + eps)
    return p
