"""GPU kernels of the learned corrector against the FP64 reference network
(oracle/cnn_oracle.py) evaluated with the device's roundings and GELU form (gelu_fit): half
precision for the 32 -> 32 layers' weights and activations in the forward pass; bfloat16 for
those weights, the pre-activation gradients and the activations entering the weight
gradients in the backward pass.  The remaining differences are FP32 accumulation order and
single-precision GELU arithmetic."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cnn():
    spec = importlib.util.spec_from_file_location("cnn_oracle", os.path.join(ROOT, "oracle", "cnn_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _params(cnn, seed):
    p = cnn.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p[288:320] = rng.uniform(-0.2, 0.2, 32)  # non-zero biases
    p[-289:-1] *= 5.0
    p[-1] = 0.05
    return p.astype(np.float32).astype(np.float64)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_matches_reference(gc, cnn, ctx, dtype):
    B, H, W = 2, 16, 64
    p = _params(cnn, 3)
    c = np.random.default_rng(0).uniform(-1, 1, (B, H, W))
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(p, 0.1)
    d_c = gc.DeviceArray.from_numpy(ctx, c.astype(dtype))
    got = net.forward(d_c, B).astype(np.float64)
    want = cnn.forward(p, c.astype(dtype).astype(np.float64), q=cnn.fp16_round,
                       gelu_fn=cnn.gelu_fit)
    assert _rel(got, want) < 3e-3, _rel(got, want)
    net.close()


def test_gradients_match_reference(gc, cnn, ctx):
    B, H, W = 2, 16, 64
    p = _params(cnn, 7)
    rng = np.random.default_rng(1)
    c = rng.uniform(-1, 1, (B, H, W))
    target = c + 0.05 * rng.normal(size=(B, H, W))
    scale = 0.1
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(p, scale)
    d_c = gc.DeviceArray.from_numpy(ctx, c)
    d_t = gc.DeviceArray.from_numpy(ctx, target)
    loss = net.loss_grad(d_c, d_t, B=B)
    g = net.grads().astype(np.float64)
    out, cache = cnn.forward(p, c, keep=True, q=cnn.fp16_round, gelu_fn=cnn.gelu_fit)
    d = out - (target - c) / scale
    assert abs(loss - np.mean(d * d)) < 3e-3 * np.mean(d * d)
    want = cnn.backward(p, cache, 2 * d / d.size, q=cnn.bf16_round, qa=cnn.bf16_round)
    # per parameter block: W1 b1 (W b)x4 W6 b6
    bounds = [0, 288, 320]
    for l in range(4):
        bounds += [bounds[-1] + 9216, bounds[-1] + 9216 + 32]
    bounds += [bounds[-1] + 288, bounds[-1] + 289]
    for lo, hi in zip(bounds, bounds[1:]):
        assert _rel(g[lo:hi], want[lo:hi]) < 2e-2, (lo, hi, _rel(g[lo:hi], want[lo:hi]))
    net.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_chain_gradients_match_reference(gc, cnn, ctx, dtype):
    """gc_cnn_loss_grad_chain through three residual applications against the oracle's
    back-propagation through the same chain (same bfloat16 roundings and GELU form): chains
    start at images 1 and 3 of a 7-image array, targets are the three images after each."""
    B, H, W, K = 2, 16, 64, 3
    p = _params(cnn, 17)
    rng = np.random.default_rng(5)
    data = (rng.uniform(-1, 1, (1, H, W)) + 0.05 * rng.normal(size=(7, H, W)).cumsum(0)).astype(dtype)
    idx = np.array([1, 3])
    scale, weights = 0.1, (0.25, 0.25, 0.5)
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(p, scale)
    d = gc.DeviceArray.from_numpy(ctx, data)
    loss = net.chain_loss_grad(d, idx, K, weights)
    g = net.grads().astype(np.float64)
    x0 = data[idx].astype(np.float64)
    targets = np.stack([data[idx + k] for k in range(1, K + 1)]).astype(np.float64)
    L, want = cnn.chain_loss_grad(p, x0, targets, scale, weights, q=cnn.fp16_round,
                                  gelu_fn=cnn.gelu_fit, qb=cnn.bf16_round, qa=cnn.bf16_round)
    assert abs(loss - L) < 3e-3 * L, (loss, L)
    bounds = [0, 288, 320]
    for l in range(4):
        bounds += [bounds[-1] + 9216, bounds[-1] + 9216 + 32]
    bounds += [bounds[-1] + 288]
    for lo, hi in zip(bounds, bounds[1:]):
        assert _rel(g[lo:hi], want[lo:hi]) < 2e-2, (lo, hi, _rel(g[lo:hi], want[lo:hi]))
    # the output bias only shifts each image's mean, which every application removes
    assert abs(g[-1]) <= 1e-4 * np.abs(g).max()
    d.free()
    net.close()


def test_chain_rejects_bad_arguments(gc, cnn, ctx):
    H, W = 8, 32
    net = gc.Corrector(ctx, H, W, 2)
    net.set_params(_params(cnn, 3), 0.1)
    d = gc.DeviceArray.from_numpy(ctx, np.zeros((5, H, W), np.float32))
    for idx, K, w in [([1, 2], 3, None),   # chain from image 2 needs images 3..5
                      ([0, 1], 0, None), ([0, 0], 9, None), ([-1, 0], 1, None),
                      ([0, 1], 2, (1.0, -0.5)), ([0, 1], 2, (1.0, np.nan))]:
        with pytest.raises(gc.GridcalcError):
            net.chain_loss_grad(d, idx, K, w)
    # a constant state gives a constant output, which the mean removal cancels exactly
    assert net.chain_loss_grad(d, [0, 1], 3, None) == 0.0
    d.free()
    net.close()


def test_predict_conserves_mass(gc, cnn, ctx):
    B, H, W = 1, 32, 32
    p = _params(cnn, 11)
    c = np.random.default_rng(2).uniform(-0.5, 0.5, (B, H, W))
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(p, 0.1)
    d_c = gc.DeviceArray.from_numpy(ctx, c)
    d_n = gc.DeviceArray((B, H, W), np.float64)
    net.predict(d_c, d_n, B)
    cn = d_n.numpy(ctx)
    out = net.forward(d_c, B).astype(np.float64)
    assert abs(cn.sum() - c.sum()) < 1e-12 * np.abs(c).sum()
    assert np.abs((cn - c) - 0.1 * (out - out.mean())).max() < 1e-9
    net.close()


def test_adam_matches_numpy(gc, cnn, ctx):
    B, H, W = 1, 8, 32
    p = _params(cnn, 13)
    rng = np.random.default_rng(3)
    c = rng.uniform(-1, 1, (B, H, W))
    t = c + 0.01 * rng.normal(size=(B, H, W))
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(p, 0.1)
    d_c, d_t = gc.DeviceArray.from_numpy(ctx, c), gc.DeviceArray.from_numpy(ctx, t)
    pw, m, v = p.astype(np.float32), np.zeros(p.size, np.float32), np.zeros(p.size, np.float32)
    for step in range(1, 4):
        net.loss_grad(d_c, d_t, B=B)
        g = net.grads()
        net.adam(1e-3)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        pw = pw - 1e-3 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
        assert np.abs(net.params() - pw).max() < 1e-6
    net.close()


def test_training_reduces_loss(gc, cnn, ctx):
    """A few hundred Adam steps on one batch of (c, c + small smooth increment) pairs
    lower the loss by an order of magnitude."""
    B, H, W = 4, 32, 64
    rng = np.random.default_rng(4)
    c = rng.uniform(-0.5, 0.5, (B, H, W))
    lap = (np.roll(c, 1, 1) + np.roll(c, -1, 1) + np.roll(c, 1, 2) + np.roll(c, -1, 2) - 4 * c)
    target = c + 0.02 * lap
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(cnn.init_params(0), 0.1)
    d_c, d_t = gc.DeviceArray.from_numpy(ctx, c), gc.DeviceArray.from_numpy(ctx, target)
    l0 = net.loss_grad(d_c, d_t, B=B)
    for _ in range(300):
        net.loss_grad(d_c, d_t, B=B)
        net.adam(2e-3)
    l1 = net.loss_grad(d_c, d_t, B=B)
    assert l1 < 0.1 * l0, (l0, l1)
    net.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(2, 16, 64), (1, 256, 256)])
def test_fused_inference_equals_layer_kernels(gc, cnn, ctx, dtype, shape):
    """The one-kernel inference path (all six layers per 8 x 32 tile, halos recomputed) gives
    the per-layer kernels' output bit for bit."""
    B, H, W = shape
    p = _params(cnn, 13)
    c = np.random.default_rng(4).uniform(-1, 1, (B, H, W))
    net = gc.Corrector(ctx, H, W, B)
    net.set_params(p, 0.1)
    d_c = gc.DeviceArray.from_numpy(ctx, c.astype(dtype))
    fused = net.forward(d_c, B)
    layers = net.forward(d_c, B, layers=True)
    assert np.array_equal(fused, layers)
    net.close()


def test_finetune_with_pushforward_runs(gc, ctx):
    """corrector.finetune on a small configuration (4 trajectories of 64 steps), one stage
    with pushforward iterations (every second iteration trains on the network's own one-step
    prediction): returns finite parameters that moved, and one loss per batch."""
    from gridcalc_amd import corrector as cr
    cfg = dict(cr.CONFIG, batch_traj=4, n_steps=64)
    p0 = cr.init_params(5)
    p, hist = cr.finetune(p0, 0.01, 0, 3, ((2, 4, 1e-4, 1e-5, 11, 2),), cfg, log=lambda *a: None)
    assert p.shape == p0.shape and np.isfinite(p).all()
    assert not np.array_equal(p, p0)
    assert len(hist) == 2 and all(np.isfinite(h) and h > 0 for h in hist)


def test_finetune_with_chains_runs(gc, ctx):
    """corrector.finetune with a chain stage (three applications per training chain) on a
    small configuration: finite parameters that moved, one positive loss per batch."""
    from gridcalc_amd import corrector as cr
    cfg = dict(cr.CONFIG, batch_traj=4, n_steps=80)
    p0 = cr.init_params(6)
    p, hist = cr.finetune(p0, 0.01, 0, 5, ((2, 3, 1e-4, 1e-5, 12, 0, 3),), cfg, log=lambda *a: None)
    assert p.shape == p0.shape and np.isfinite(p).all()
    assert not np.array_equal(p, p0)
    assert len(hist) == 2 and all(np.isfinite(h) and h > 0 for h in hist)
