"""Pins the FP64 reference of the corrector network (oracle/cnn_oracle.py): its forward pass
against torch's CPU conv2d with circular padding, and its gradients against central finite
differences.  CPU only."""
import importlib.util
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cnn():
    spec = importlib.util.spec_from_file_location("cnn_oracle", os.path.join(ROOT, "oracle", "cnn_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_param_count(cnn):
    assert cnn.N_PARAMS == 37601


def test_forward_matches_torch_conv2d(cnn):
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    p = cnn.init_params(3)
    p[-1] = 0.2
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, size=(2, 7, 9))
    W1, b1, mids, W6, b6 = cnn.split(p)
    t = torch.tensor(x, dtype=torch.float64)[:, None]

    def conv(inp, w, b):  # w: (co, ci, 3, 3)
        return F.conv2d(F.pad(inp, (1, 1, 1, 1), mode="circular"), w, b)

    w1 = torch.tensor(W1.reshape(32, 1, 3, 3))
    a = F.gelu(conv(t, w1, torch.tensor(b1)))
    for W, b in mids:  # W[co][tap][ci] -> (co, ci, 3, 3)
        w = torch.tensor(np.transpose(W.reshape(32, 3, 3, 32), (0, 3, 1, 2)).copy())
        a = F.gelu(conv(a, w, torch.tensor(b)))
    w6 = torch.tensor(np.transpose(W6.reshape(3, 3, 32), (2, 0, 1))[None].copy())
    out_t = conv(a, w6, torch.tensor(b6))[:, 0].numpy()
    out = cnn.forward(p, x)
    assert np.abs(out - out_t).max() < 1e-12


def test_gradients_match_finite_differences(cnn):
    p = cnn.init_params(5)
    p[-289:-1] *= 10.0  # make the last layer matter
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(1, 5, 6))
    w = rng.normal(size=(1, 5, 6))
    out, cache = cnn.forward(p, x, keep=True)
    g = cnn.backward(p, cache, w)
    idx = list(range(0, 288, 37)) + [300, 320 + 17, 320 + 9216 + 5, 4 * 9248 + 320 - 3,
                                    4 * 9248 + 320 + 7, cnn.N_PARAMS - 1]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = ((w * cnn.forward(pp, x)).sum() - (w * cnn.forward(pm, x)).sum()) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_input_gradient_matches_finite_differences(cnn):
    p = cnn.init_params(6)
    p[-289:-1] *= 10.0
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, size=(2, 5, 6))
    w = rng.normal(size=(2, 5, 6))
    _, cache = cnn.forward(p, x, keep=True)
    _, dx = cnn.backward(p, cache, w, want_input=True)
    eps = 1e-6
    for b, y, xx in [(0, 0, 0), (0, 4, 5), (1, 2, 3), (1, 0, 5), (0, 3, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[b, y, xx] += eps
        xm[b, y, xx] -= eps
        fd = ((w * cnn.forward(p, xp)).sum() - (w * cnn.forward(p, xm)).sum()) / (2 * eps)
        assert abs(fd - dx[b, y, xx]) <= 1e-6 * max(1.0, abs(fd)), (b, y, xx, fd, dx[b, y, xx])


def test_chain_gradient_matches_finite_differences(cnn):
    """Back-propagation through three residual applications (per-image mean removal, loss
    on every step with unequal weights) against central differences of the chain loss."""
    p = cnn.init_params(8)
    p[-289:-1] *= 10.0
    p[-1] = 0.3  # a mean the update must remove
    rng = np.random.default_rng(4)
    x0 = rng.uniform(-1, 1, size=(2, 5, 6))
    targets = x0 + 0.05 * rng.normal(size=(3, 2, 5, 6)).cumsum(axis=0)
    scale, weights = 0.3, (0.5, 0.2, 1.0)
    L, g = cnn.chain_loss_grad(p, x0, targets, scale, weights)
    assert abs(L - cnn.chain_loss(p, x0, targets, scale, weights)) <= 1e-14 * L
    idx = list(range(0, 288, 41)) + [290, 320 + 11, 320 + 9216 + 3, 2 * 9248 + 320 + 500,
                                    4 * 9248 + 320 + 9, cnn.N_PARAMS - 2]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (cnn.chain_loss(pp, x0, targets, scale, weights)
              - cnn.chain_loss(pm, x0, targets, scale, weights)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the output bias only shifts every image's mean, which the update removes
    assert abs(g[-1]) <= 1e-12 * np.abs(g).max()


def test_gelu_fit_error_bound(cnn):
    """The device's GELU form against z Phi(z): the stated maximum error, the limits z and 0
    far out, and z / 2 slope at the origin (a wrong coefficient or sign breaks one)."""
    z = np.linspace(-40.0, 40.0, 800001)
    err = np.abs(cnn.gelu_fit(z) - cnn.gelu(z))
    assert err.max() < 2.6e-5
    assert err[np.abs(z) <= 9].max() > 2.0e-5  # the fit's equioscillation level, not smaller
    assert np.array_equal(cnn.gelu_fit(np.array([0.0])), [0.0])
    assert abs(cnn.gelu_fit(50.0) - 50.0) < 1e-9 and abs(cnn.gelu_fit(-50.0)) < 1e-9
    h = 1e-6
    assert abs(cnn.gelu_fit(h) / h - 0.5) < 1e-5


def test_gelu_derivative(cnn):
    z = np.linspace(-4, 4, 41)
    fd = (cnn.gelu(z + 1e-6) - cnn.gelu(z - 1e-6)) / 2e-6
    assert np.abs(fd - cnn.gelu_grad(z)).max() < 1e-8
    assert abs(cnn.gelu(1.0) - 0.8413447460685429) < 1e-15  # 1 * Phi(1)


def test_bf16_round(cnn):
    assert cnn.bf16_round(1.0 + 2 ** -9) == 1.0           # ties to even
    assert cnn.bf16_round(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert cnn.bf16_round(-3.140625) == -3.140625
    assert math.isclose(cnn.bf16_round(0.1), 0.10009765625)


def test_fp16_round(cnn):
    assert cnn.fp16_round(1.0 + 2 ** -11) == 1.0          # ties to even
    assert cnn.fp16_round(1.0 + 3 * 2 ** -11) == 1.0 + 2 ** -9
    assert cnn.fp16_round(65504.0) == 65504.0              # largest finite half
    assert cnn.fp16_round(2.0 ** -24) == 2.0 ** -24        # smallest subnormal
    assert cnn.fp16_round(2.0 ** -26) == 0.0
    assert math.isclose(cnn.fp16_round(0.1), 0.0999755859375)


def test_backward_activation_rounding_applies_to_mid_weight_gradients_only(cnn):
    """qa rounds the activations entering the four 32 -> 32 weight gradients; the first and
    last layers' gradients and the input gradient do not see it."""
    p = cnn.init_params(9)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(1, 4, 5))
    w = rng.normal(size=(1, 4, 5))
    _, cache = cnn.forward(p, x, keep=True)
    g0, dx0 = cnn.backward(p, cache, w, want_input=True)
    g1, dx1 = cnn.backward(p, cache, w, want_input=True, qa=lambda a: 2.0 * a)
    mid = slice(320, 320 + 4 * 9248)
    assert np.array_equal(g0[:320], g1[:320]) and np.array_equal(g0[-289:], g1[-289:])
    assert np.array_equal(dx0, dx1)
    for l in range(4):  # weights doubled, biases unchanged
        o = 320 + l * 9248
        assert np.allclose(g1[o:o + 9216], 2.0 * g0[o:o + 9216], rtol=1e-12, atol=0)
        assert np.array_equal(g1[o + 9216:o + 9248], g0[o + 9216:o + 9248])
    assert not np.array_equal(g0[mid], g1[mid])
