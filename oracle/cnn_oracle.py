"""FP64 numpy reference of the learned Cahn-Hilliard corrector network (test infrastructure:
used only by tests/ to check the GPU kernels' forward pass and gradients).

Network (6 layers, 3x3 periodic convolutions, 32 channels, GELU, residual):
    a1 = gelu(conv(x, W1) + b1)            1 -> 32
    a_l = gelu(conv(a_{l-1}, W_l) + b_l)   32 -> 32, l = 2..5
    out = conv(a5, W6) + b6                32 -> 1
    prediction = c + scale * out
Parameters are one flat float vector in this order (37601 values):
    W1[co][tap] (32*9), b1 (32), {W_l[co][tap][ci] (32*9*32), b_l (32)} for l = 2..5,
    W6[tap][ci] (9*32), b6 (1)
with tap = 3*(dy+1) + (dx+1) for the offset (dy, dx) in {-1,0,1}^2 and conv(x, W)[y, x] =
sum_tap W[tap] * x[(y+dy) mod H, (x+dx) mod W] (cross-correlation, periodic padding).
Arrays are (B, H, W) for single-channel fields and (B, H, W, C) for 32-channel activations.

No code of the reference is involved: the reference has no network; this module defines it.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

C = 32
TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
N_PARAMS = 32 * 9 + 32 + 4 * (32 * 9 * 32 + 32) + 9 * 32 + 1


def split(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.size == N_PARAMS
    o = 0
    W1 = p[o:o + 288].reshape(32, 9); o += 288
    b1 = p[o:o + 32]; o += 32
    mids = []
    for _ in range(4):
        W = p[o:o + 9216].reshape(32, 9, 32); o += 9216
        b = p[o:o + 32]; o += 32
        mids.append((W, b))
    W6 = p[o:o + 288].reshape(9, 32); o += 288
    b6 = p[o:o + 1]; o += 1
    return W1, b1, mids, W6, b6


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


# The device evaluates GELU as z sigma(z P(z^2)), sigma(u) = 1 / (1 + exp(-u)), with this P
# (least maximum error against gelu on |z| <= 9; z clamped to [-9, 9] inside P).
GELU_FIT = (1.5950206, 0.074006644, -7.0221731e-4)


def gelu_fit(z):
    """The device's GELU form (|gelu_fit - gelu| <= 2.6e-5)."""
    zc = np.clip(z, -9.0, 9.0)
    t = zc * zc
    P = GELU_FIT[0] + t * (GELU_FIT[1] + t * GELU_FIT[2])
    return z / (1.0 + np.exp(-zc * P))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def shift(x, dy, dx):
    """x[(y+dy) mod H, (x+dx) mod W] over axes 1, 2."""
    return np.roll(np.roll(x, -dy, axis=1), -dx, axis=2)


def forward(p, x, keep=False, q=None, gelu_fn=None):
    """x: (B, H, W) network input.  Returns out (B, H, W) and, with keep, the cache.
    q, if given, is applied to the weights of the four 32 -> 32 layers and to every
    activation (e.g. fp16_round, to model a device that stores those in half precision and
    keeps the input, W1 and W6 in float32); gelu_fn replaces gelu (e.g. gelu_fit)."""
    gelu_ = gelu_fn if gelu_fn is not None else gelu
    W1, b1, mids, W6, b6 = split(p)
    if q is not None:
        mids = [(q(W), b) for W, b in mids]
    act = (lambda z: q(gelu_(z))) if q is not None else gelu_
    z1 = np.zeros(x.shape + (C,))
    for t, (dy, dx) in enumerate(TAPS):
        z1 += shift(x, dy, dx)[..., None] * W1[:, t]
    z1 += b1
    zs, acts = [z1], [act(z1)]
    for W, b in mids:
        a = acts[-1]
        z = np.zeros_like(a)
        for t, (dy, dx) in enumerate(TAPS):
            z += shift(a, dy, dx) @ W[:, t, :].T
        z += b
        zs.append(z)
        acts.append(act(z))
    out = np.zeros(x.shape)
    for t, (dy, dx) in enumerate(TAPS):
        out += shift(acts[-1], dy, dx) @ W6[t]
    out += b6[0]
    if keep:
        return out, (x, zs, acts)
    return out


def backward(p, cache, dout, q=None, want_input=False, qa=None):
    """Gradient of sum(dout * out) with respect to the flat parameters.  q, if given, is
    applied to the weights of the four 32 -> 32 layers and to every pre-activation gradient,
    qa to the cached activations that enter those four layers' weight gradients (e.g. q = qa
    = bf16_round for a device whose backward pass runs in bfloat16).  With want_input also
    the gradient with respect to the network input x, returned second."""
    W1, b1, mids, W6, b6 = split(p)
    if q is not None:
        mids = [(q(W), b) for W, b in mids]
    qg = q if q is not None else (lambda a: a)
    qa_ = qa if qa is not None else (lambda a: a)
    x, zs, acts = cache
    g = []
    # last layer
    a5 = acts[-1]
    dW6 = np.zeros_like(W6)
    da = np.zeros_like(a5)
    for t, (dy, dx) in enumerate(TAPS):
        dW6[t] = np.einsum("bhw,bhwc->c", dout, shift(a5, dy, dx))
        da += shift(dout, -dy, -dx)[..., None] * W6[t]
    db6 = np.array([dout.sum()])
    mid_grads = []
    for li in range(3, -1, -1):
        W, b = mids[li]
        dz = qg(da * gelu_grad(zs[li + 1]))
        a_in = qa_(acts[li])
        dW = np.zeros_like(W)
        da = np.zeros_like(a_in)
        for t, (dy, dx) in enumerate(TAPS):
            dW[:, t, :] = np.einsum("bhwo,bhwi->oi", dz, shift(a_in, dy, dx))
            da += shift(dz, -dy, -dx) @ W[:, t, :]
        mid_grads.insert(0, (dW, dz.sum(axis=(0, 1, 2))))
    dz1 = qg(da * gelu_grad(zs[0]))
    dW1 = np.zeros_like(W1)
    for t, (dy, dx) in enumerate(TAPS):
        dW1[:, t] = np.einsum("bhwc,bhw->c", dz1, shift(x, dy, dx))
    db1 = dz1.sum(axis=(0, 1, 2))
    if want_input:
        dx = np.zeros_like(x)
        for t, (dy, dx_) in enumerate(TAPS):
            dx += shift(dz1, -dy, -dx_) @ W1[:, t]
    g.append(dW1.ravel())
    g.append(db1)
    for dW, db in mid_grads:
        g.append(dW.ravel())
        g.append(db)
    g.append(dW6.ravel())
    g.append(db6)
    if want_input:
        return np.concatenate(g), dx
    return np.concatenate(g)


def chain(p, x0, K, scale, q=None, gelu_fn=None):
    """K applications of the residual update c_k = c_{k-1} + scale (out(c_{k-1}) - mean(out)),
    the mean taken per image: returns [c_0, ..., c_K] and the K forward caches."""
    cs, caches = [np.asarray(x0, dtype=np.float64)], []
    for _ in range(K):
        out, cache = forward(p, cs[-1], keep=True, q=q, gelu_fn=gelu_fn)
        caches.append(cache)
        cs.append(cs[-1] + scale * (out - out.mean(axis=(1, 2), keepdims=True)))
    return cs, caches


def chain_loss(p, x0, targets, scale, weights, q=None, gelu_fn=None):
    """L = sum_k weights[k-1] mean(((c_k - targets[k-1]) / scale)^2), k = 1..K, the mean over
    all B H W sites; targets (K, B, H, W)."""
    cs, _ = chain(p, x0, len(targets), scale, q, gelu_fn)
    return sum(w * np.mean(((c - t) / scale) ** 2) for w, c, t in zip(weights, cs[1:], targets))


def chain_loss_grad(p, x0, targets, scale, weights, q=None, gelu_fn=None, qb=None, qa=None):
    """chain_loss and its gradient with respect to the flat parameters, through every
    application (back-propagation through the unrolled chain).  With G_k = dL/dc_k:
    G_K = 2 w_K (c_K - t_K) / (scale^2 N); application k's output gradient is
    scale (G_k - mean(G_k)) (the mean removal is a projection); G_{k-1} = G_k (the residual
    path) + the network's input gradient + 2 w_{k-1} (c_{k-1} - t_{k-1}) / (scale^2 N).
    q and gelu_fn as in forward; qb and qa as backward's q and qa."""
    K = len(targets)
    cs, caches = chain(p, x0, K, scale, q, gelu_fn)
    N = cs[0].size
    L = sum(w * np.mean(((c - t) / scale) ** 2) for w, c, t in zip(weights, cs[1:], targets))
    G = 2.0 * weights[K - 1] * (cs[K] - targets[K - 1]) / (scale * scale * N)
    grad = np.zeros(N_PARAMS)
    for k in range(K, 0, -1):
        dout = scale * (G - G.mean(axis=(1, 2), keepdims=True))
        gk, dx = backward(p, caches[k - 1], dout, q=qb, want_input=True, qa=qa)
        grad += gk
        if k > 1:
            G = G + dx + 2.0 * weights[k - 2] * (cs[k - 1] - targets[k - 2]) / (scale * scale * N)
    return L, grad


def init_params(seed=0):
    """He-style initialisation (uniform), small last layer."""
    rng = np.random.default_rng(seed)
    parts = []
    parts.append(rng.uniform(-1, 1, 288) * math.sqrt(6.0 / 9.0))
    parts.append(np.zeros(32))
    for _ in range(4):
        parts.append(rng.uniform(-1, 1, 9216) * math.sqrt(6.0 / 288.0))
        parts.append(np.zeros(32))
    parts.append(rng.uniform(-1, 1, 288) * math.sqrt(6.0 / 288.0) * 0.1)
    parts.append(np.zeros(1))
    return np.concatenate(parts)


def fp16_round(a):
    """Round-to-nearest-even to IEEE half precision, returned as float64."""
    return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)
