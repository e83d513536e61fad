"""U-Net preconditioner on the GPU (BF16 MFMA convolutions) against the FP64
oracle (oracle/unet.py) run with the same bfloat16 operand rounding.

Tolerances: the oracle rounds the same operands to bf16, so forward outputs
differ only by FP32 accumulation order until an operand's bf16 rounding flips
after a 1e-7 upstream difference; such flips cascade, so the forward bound is
"no further from FP64 than bf16 arithmetic is" and 1e-2 relative L2 to the
bf16 oracle (network output, residual removed).  Gradients: relative L2 <=
1e-2 per layer.
"""
import numpy as np
import pytest

import unet as U

pytestmark = pytest.mark.gpu


def _params(c, seed):
    """Oracle-side initialisation with a full-scale output layer and small
    random biases, so that a wrong convolution cannot hide behind the residual."""
    p = U.init_params(c, seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    off = 0
    specs = U.layer_specs(c)
    for li, (ci, co, _) in enumerate(specs):
        if li == len(specs) - 1:
            p[off:off + ci * co * 9] *= 100.0
        off += ci * co * 9
        p[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    return p.astype(np.float32)


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("c,shape", [((16, 32, 64, 64), (2, 2, 32, 48)), ((8, 12, 20, 24), (3, 2, 24, 16))])
def test_forward_matches_oracle(erm, c, shape):
    p = _params(c, 3)
    x = np.random.default_rng(4).standard_normal(shape).astype(np.float32)
    with erm.UNet(c, seed=1) as net:
        assert net.num_params == U.num_params(c)
        net.set_params(p)
        np.testing.assert_array_equal(net.get_params(), p)
        y = net.forward(x)
    yo = U.UNetOracle(c, p, bf16=True).forward(x)
    yf = U.UNetOracle(c, p, bf16=False).forward(x)
    e = _rel(y - x, yo - x)
    e_bf16 = _rel(yo - x, yf - x)
    e_gpu = _rel(y - x, yf - x)
    print(f"forward rel L2: vs bf16 oracle {e:.2e}, vs fp64 oracle {e_gpu:.2e} (bf16 oracle vs fp64 {e_bf16:.2e})")
    # Identical bf16 roundings give FP32-accumulation agreement; once one
    # activation lands on the other side of a bf16 rounding boundary the
    # difference cascades through later layers up to the bf16 noise level of
    # the network itself, so the GPU must be no further from FP64 than
    # bf16 arithmetic is, and within 1e-2 of the bf16 oracle.
    assert e < 1e-2
    assert e_gpu <= 1.5 * e_bf16


@pytest.mark.parametrize("c,shape", [((16, 32, 64, 64), (2, 2, 32, 32)), ((8, 12, 20, 24), (2, 2, 16, 24))])
def test_gradient_matches_oracle(erm, c, shape):
    p = _params(c, 5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(shape).astype(np.float32)
    t = rng.standard_normal(shape).astype(np.float32)
    with erm.UNet(c) as net:
        net.set_params(p)
        loss, g = net.loss_grad(x, t)
    lo, go = U.UNetOracle(c, p, bf16=True).loss_grad(x, t)
    assert abs(loss - lo) < 2e-3 * lo
    off = 0
    for li, (ci, co, _) in enumerate(U.layer_specs(c)):
        n = ci * co * 9 + co
        e = _rel(g[off:off + n], go[off:off + n])
        assert e < 1e-2, (li, e)
        off += n


def test_training_reduces_loss(erm):
    """Adam on a fixed smoothing target: the loss falls by 5x within 150 steps."""
    c = (16, 32, 64, 64)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((8, 2, 32, 32)).astype(np.float32)
    k = np.ones(5) / 5.0
    t = np.apply_along_axis(lambda r: np.convolve(r, k, "same"), 3, x)
    t = (0.5 * np.apply_along_axis(lambda r: np.convolve(r, k, "same"), 2, t)).astype(np.float32)
    with erm.UNet(c, seed=2) as net:
        l0 = net.loss_grad(x, t)[0]
        losses = net.train(x, t, batch=4, epochs=75, lr=2e-3, seed=1)
        ms, flops = net.last_stats()
    print(f"loss {l0:.4f} -> {losses[-1]:.4f}; {ms:.1f} ms, {flops / ms / 1e9:.2f} TFLOP/s")
    assert losses[-1] < l0 / 5


def test_bad_shape_raises(erm):
    with erm.UNet((8, 8, 8, 8)) as net:
        with pytest.raises(erm.ErmError):
            net.forward(np.zeros((1, 2, 12, 16), np.float32))


def test_adam_steps_match_oracle(erm):
    """Two train steps on the GPU against the FP64 Adam restatement fed with
    the GPU's own gradients at p0 and p1.  Tolerance: 2e-6 absolute on
    parameters of magnitude <= ~1 (0.2% of a 1e-3 step; FP32 rounding of p
    is ~1e-7), on parameters whose gradients exceed 1e-6 at both points
    (100x the FP32 rounding of the largest partial sums, ~1e-8), so that
    the atomic reordering between the two gradient evaluations cannot flip a
    step's sign; that keeps 60-80% of the parameters (the rest have zero or
    near-zero gradients behind the 1%-scale output layer).  The bf16 copy the Adam kernel
    refreshes must equal a fresh conversion of the same FP32 weights (the
    forward pass is deterministic: it has no atomics)."""
    c = (8, 8, 8, 8)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 2, 16, 16)).astype(np.float32)
    t = (0.5 * x).astype(np.float32)
    lr = 1e-3
    with erm.UNet(c, seed=4) as net:
        p0 = net.get_params()
        _, g0 = net.loss_grad(x, t)
        net.train_step(x, t, lr)
        p1 = net.get_params()
        _, g1 = net.loss_grad(x, t)
        net.train_step(x, t, lr)
        p2 = net.get_params()
        y_adam = net.forward(x)
        net.set_params(p2)
        y_fresh = net.forward(x)
    r1 = U.adam(p0, [g0], lr)
    r2 = U.adam(p0, [g0, g1], lr)
    keep = (np.abs(g0) > 1e-6) & (np.abs(g1) > 1e-6)
    assert keep.mean() > 0.5, keep.mean()
    e1 = np.abs(p1 - r1)[keep].max()
    e2 = np.abs(p2 - r2)[keep].max()
    moved = np.abs(p1 - p0)[keep]
    print(f"adam: {keep.sum()} of {keep.size} params compared; max |dp| error {e1:.2e} / {e2:.2e};"
          f" median step {np.median(moved):.2e}")
    assert np.median(moved) > 0.9 * lr
    assert e1 < 2e-6 and e2 < 2e-6
    np.testing.assert_allclose(y_adam, y_fresh, rtol=0, atol=1e-6 * np.abs(y_fresh).max())
