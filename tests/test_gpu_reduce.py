"""GPU parity of the reduction path: integrate() (pairwise reproduced exactly, Kahan to
within compensated-summation error) and the fused free / gradient energy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 7, 63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 1000, 4095, 4096, 4097,
         65536, 65537, 100003, 1 << 20, 3 * (1 << 20) + 17]


def _wild(gc, n, seed):
    """Values spanning 16 decades, so every change of summation order shows."""
    mag = 10.0 ** (16.0 * gc.inputs.uniform(seed + 1, n) - 8.0)
    return gc.inputs.symmetric(seed, n) * mag


@pytest.mark.parametrize("n", SIZES)
def test_pairwise_bit_exact(gc, orc, ctx, n):
    x = _wild(gc, n, seed=n)
    g = gc.Grid.make(n, 1, 1, (0.5, 1.0, 2.0))
    og = orc.Grid.make(n, 1, 1, (0.5, 1.0, 2.0))
    got = gc.integrate(ctx, g, gc.DeviceArray.from_numpy(ctx, x), gc.PAIRWISE)
    assert got == orc.integrate(og, x, orc.PAIRWISE)


@pytest.mark.parametrize("n", [1, 64, 65, 1000, 65537, 1 << 20])
def test_pairwise_f32_input(gc, orc, ctx, n):
    x = gc.inputs.symmetric(n, n).astype(np.float32)
    g = gc.Grid.make(n, 1, 1)
    got = gc.integrate(ctx, g, gc.DeviceArray.from_numpy(ctx, x), gc.PAIRWISE)
    assert got == orc.integrate(orc.Grid.make(n, 1, 1), x.astype(np.float64), orc.PAIRWISE)


@pytest.mark.parametrize("n", SIZES)
def test_kahan_close(gc, orc, ctx, n):
    x = _wild(gc, n, seed=3 * n)
    g = gc.Grid.make(n, 1, 1)
    got = gc.integrate(ctx, g, gc.DeviceArray.from_numpy(ctx, x), gc.KAHAN)
    want = orc.integrate(orc.Grid.make(n, 1, 1), x, orc.KAHAN)
    # both are compensated sums: they agree to a few ulp of sum |x|
    assert abs(got - want) <= 4e-16 * np.abs(x).sum() + 1e-300


def test_integrate_known_answers(gc, ctx):
    # integrate_test.cpp:31-48: unit field over an anisotropic grid recovers the volume
    g = gc.Grid.make(8, 16, 4, (0.25, 0.125, 0.5))
    ones = gc.DeviceArray.from_numpy(ctx, np.ones(g.shape))
    vol = (8 * 0.25) * (16 * 0.125) * (4 * 0.5)
    assert gc.integrate(ctx, g, ones, gc.PAIRWISE) == vol
    assert gc.integrate(ctx, g, ones, gc.KAHAN) == vol


def test_integrate_deterministic(gc, ctx):
    n = 32
    h = 2 * np.pi / n
    x = np.arange(n) * h
    f = np.sin(x)[None, None, :] * np.sin(x)[None, :, None] * np.sin(x)[:, None, None]
    g = gc.Grid.make(n, n, n, h)
    d = gc.DeviceArray.from_numpy(ctx, f)
    for m in (gc.PAIRWISE, gc.KAHAN):
        a, b = gc.integrate(ctx, g, d, m), gc.integrate(ctx, g, d, m)
        assert np.float64(a).tobytes() == np.float64(b).tobytes()
        assert abs(a) < 1e-12


@pytest.mark.parametrize("shape", [(64, 64, 1), (128, 128, 1), (16, 12, 10), (33, 17, 5)])
@pytest.mark.parametrize("method", [0, 1])
def test_free_energy(gc, orc, ctx, shape, method):
    nx, ny, nz = shape
    g, og = gc.Grid.make(nx, ny, nz, (1.0, 0.7, 1.3)), orc.Grid.make(nx, ny, nz, (1.0, 0.7, 1.3))
    psi = gc.inputs.ch_initial(nx, ny, nz, 0.6, seed=21)
    d = gc.DeviceArray.from_numpy(ctx, psi)
    F = gc.ch_free_energy(ctx, g, d, 0.8, method)
    E = gc.gradient_energy(ctx, g, d, method)
    F_ref = orc.free_energy(og, psi, 0.8, method)
    E_ref = orc.gradient_energy(og, psi, method)
    if method == 0:
        assert F == F_ref and E == E_ref
    else:
        assert abs(F - F_ref) <= 1e-15 * abs(F_ref) + 1e-13
        assert abs(E - E_ref) <= 1e-15 * abs(E_ref) + 1e-13


def test_free_energy_f32_in_f64(gc, orc, ctx):
    g, og = gc.Grid.make(96, 80, 1), orc.Grid.make(96, 80, 1)
    psi = gc.inputs.ch_initial(96, 80, 1, 0.5, seed=2).astype(np.float32)
    F = gc.ch_free_energy(ctx, g, gc.DeviceArray.from_numpy(ctx, psi), 1.0, gc.PAIRWISE)
    assert F == orc.free_energy(og, psi.astype(np.float64), 1.0, orc.PAIRWISE)
