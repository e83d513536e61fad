"""Pinning of the FP64 U-Net oracle (oracle/unet.py) on the CPU.

- the 3x3 convolution against direct loops;
- the whole-network gradient against central finite differences (FP64);
- loss and gradient against torch.autograd in float64 (an independent
  implementation of conv2d / avg_pool2d / nearest upsampling);
- bfloat16 rounding against torch's float32 -> bfloat16 conversion (RNE) and
  hand-computed ties.
A dropped skip connection, a flipped kernel in the data gradient or a
missing ReLU mask fails the finite-difference and autograd checks.
"""
import numpy as np
import pytest

import unet as U


def test_conv3x3_direct_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 6))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = U.conv3x3(x, w, b)
    ref = np.zeros((2, 4, 5, 6))
    for n in range(2):
        for o in range(4):
            for i in range(5):
                for j in range(6):
                    s = b[o]
                    for c in range(3):
                        for ky in range(3):
                            for kx in range(3):
                                ii, jj = i + ky - 1, j + kx - 1
                                if 0 <= ii < 5 and 0 <= jj < 6:
                                    s += w[o, c, ky, kx] * x[n, c, ii, jj]
                    ref[n, o, i, j] = s
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv_backward_is_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 6, 5))
    w = rng.standard_normal((4, 3, 3, 3))
    dy = rng.standard_normal((2, 4, 6, 5))
    dx, dw, db = U.conv3x3_backward(x, w, dy)
    # <conv(x), dy> = <x, dx> (bias 0) and d<conv(x),dy>/dw = dw
    y = U.conv3x3(x, w, np.zeros(4))
    assert abs(np.sum(y * dy) - np.sum(x * dx)) < 1e-10 * np.abs(np.sum(y * dy))
    assert abs(np.sum(y * dy) - np.sum(w * dw)) < 1e-10 * np.abs(np.sum(y * dy))
    np.testing.assert_allclose(db, dy.sum(axis=(0, 2, 3)))


def _tiny_net(seed=2, c=(3, 4, 5, 3)):
    p = U.init_params(c, seed).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    # larger output layer and non-zero biases so every path matters
    specs = U.layer_specs(c)
    off = 0
    for li, (ci, co, _) in enumerate(specs):
        if li == len(specs) - 1:
            p[off:off + ci * co * 9] *= 100.0
        off += ci * co * 9
        p[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    return c, p


def test_gradient_finite_differences():
    c, p = _tiny_net()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 2, 8, 8))
    t = rng.standard_normal((2, 2, 8, 8))
    loss, g = U.UNetOracle(c, p).loss_grad(x, t)
    idx = rng.choice(len(p), 60, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = U.UNetOracle(c, pp).loss_grad(x, t)[0]
        lm = U.UNetOracle(c, pm).loss_grad(x, t)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g).max()), (i, fd, g[i])


def test_against_torch_autograd():
    torch = pytest.importorskip("torch")
    import torch.nn.functional as F
    c, p = _tiny_net(seed=7, c=(4, 6, 8, 8))
    rng = np.random.default_rng(8)
    x = rng.standard_normal((3, 2, 16, 24))
    t = rng.standard_normal((3, 2, 16, 24))
    loss, g = U.UNetOracle(c, p).loss_grad(x, t)

    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    layers, off = [], 0
    for ci, co, relu in U.layer_specs(c):
        w = pt[off:off + ci * co * 9].view(co, ci, 3, 3)
        off += ci * co * 9
        b = pt[off:off + co]
        off += co
        layers.append((w, b, relu))

    def cv(i, a):
        w, b, relu = layers[i]
        y = F.conv2d(a, w, b, padding=1)
        return torch.relu(y) if relu else y

    up = lambda a: F.interpolate(a, scale_factor=2, mode="nearest")
    xt = torch.tensor(x)
    e1 = cv(1, cv(0, xt))
    e2 = cv(3, cv(2, F.avg_pool2d(e1, 2)))
    e3 = cv(5, cv(4, F.avg_pool2d(e2, 2)))
    bb = cv(7, cv(6, F.avg_pool2d(e3, 2)))
    d3 = cv(9, cv(8, torch.cat([up(bb), e3], 1)))
    d2 = cv(11, cv(10, torch.cat([up(d3), e2], 1)))
    d1 = cv(13, cv(12, torch.cat([up(d2), e1], 1)))
    y = cv(14, d1) + xt
    lt = torch.mean((y - torch.tensor(t)) ** 2)
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * loss
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-12 * np.abs(g).max())


def test_bf16_rounding():
    torch = pytest.importorskip("torch")
    e = 2.0 ** -8
    vals = np.array([1.0, 1 + e, 1 + 3 * e, 1 + 1.01 * e, -(1 + 3 * e), 3.14159, 1e-30, -7.5e5], np.float32)
    q = U.bf16_round(vals)
    assert q[0] == 1.0 and q[1] == 1.0 and q[2] == 1 + 4 * e and q[3] == 1 + 2 * e and q[4] == -(1 + 4 * e)
    rng = np.random.default_rng(0)
    r = (rng.standard_normal(10000) * 10.0 ** rng.integers(-20, 20, 10000)).astype(np.float32)
    tq = torch.tensor(r).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(U.bf16_round(r).astype(np.float32), tq)


def test_param_count_matches_layout():
    c = (16, 32, 64, 64)
    n = U.num_params(c)
    assert n == len(U.init_params(c, 0))
    assert len(U.pack([(w, b) for w, b in U.unpack(np.zeros(n), c)])) == n


def test_adam_closed_forms():
    """Bias-corrected Adam: with a constant gradient the corrected moments are
    g and g^2 at every step, so each step moves a parameter by
    lr*g/(|g| + eps) (an uncorrected first step would move it by
    0.1/sqrt(0.001) = 3.16 times that); with g then -g the second corrected
    first moment is (0.09 - 0.1) g / 0.19 = -g/19 and the second moment g^2,
    so the second step moves it back by lr*g/(19 (|g| + eps)) (a wrong beta1
    changes the 19, a wrong beta2 the second moment)."""
    rng = np.random.default_rng(0)
    p0 = rng.standard_normal(64)
    g = rng.standard_normal(64)
    lr, eps = 1e-3, 1e-8
    step = lr * g / (np.abs(g) + eps)
    for steps in (1, 2, 5):
        p = U.adam(p0, [g] * steps, lr)
        np.testing.assert_allclose(p, p0 - steps * step, rtol=0, atol=1e-14)
    p = U.adam(p0, [g, -g], lr)
    np.testing.assert_allclose(p, p0 - step + step / 19.0, rtol=0, atol=1e-14)
    assert np.array_equal(U.adam(p0, [np.zeros(64)], lr), p0)
