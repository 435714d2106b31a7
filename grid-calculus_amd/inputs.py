"""Seeded synthetic inputs for the benchmark configurations and the tests.

Values come from a counter-based generator (splitmix64 of seed + (index+1)*golden gamma,
top 53 bits as a uniform double in [0, 1)), so any element can be produced independently
and the same array is fed to the GPU path and to the CPU oracle.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform(seed, n, offset=0):
    """n uniforms in [0, 1) for counters offset .. offset+n-1 of stream `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def symmetric(seed, n, amp=1.0, offset=0):
    """amp * U(-1, 1)."""
    return amp * (2.0 * uniform(seed, n, offset) - 1.0)


def normal(seed, n):
    """Standard normals by Box-Muller over two uniform streams."""
    u1 = uniform(seed, n, 0)
    u2 = uniform(seed ^ 0x5DEECE66D, n, 0)
    u1 = np.where(u1 <= 0.0, 1e-300, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def ch_initial(nx, ny, nz, amp, seed, mean=0.0):
    """Cahn-Hilliard initial condition mean + amp*U(-1,1), shape (nz, ny, nx)."""
    return (mean + symmetric(seed, nx * ny * nz, amp)).reshape(nz, ny, nx)


def sweep_field(n, h, seed, n_modes=8, kmax=16):
    """u ~ U(-1,1) plus a sum of n_modes sines with integer wavevectors in [-kmax, kmax]^3
    and random phases on an n^3 periodic grid of spacing h (domain length n*h)."""
    rng = np.random.default_rng(seed)
    L = n * h
    x = np.arange(n) * h
    u = symmetric(seed, n ** 3).reshape(n, n, n)
    for _ in range(n_modes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        phi = rng.uniform(0.0, 2.0 * np.pi)
        X = 2 * np.pi * k[0] * x / L + phi
        Y = 2 * np.pi * k[1] * x / L
        Z = 2 * np.pi * k[2] * x / L
        sx, cx = np.sin(X)[None, None, :], np.cos(X)[None, None, :]
        sy, cy = np.sin(Y)[None, :, None], np.cos(Y)[None, :, None]
        sz, cz = np.sin(Z)[:, None, None], np.cos(Z)[:, None, None]
        # sin(X+Y+Z) = cos Z sin(X+Y) + sin Z cos(X+Y); one n^3 temporary at a time
        u += cz * (sx * cy + cx * sy)
        u += sz * (cx * cy - sx * sy)
    return u


def lognormal(seed, n):
    return np.exp(normal(seed, n))
