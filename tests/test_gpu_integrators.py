"""GPU parity of the time-stepping path: the fused Cahn-Hilliard RK4/Euler pipeline, the
diffusion drivers and the pointwise integrator steps, against the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grids(gc, orc, nx, ny, nz, h=1.0):
    return gc.Grid.make(nx, ny, nz, h), orc.Grid.make(nx, ny, nz, h)


@pytest.mark.parametrize("shape", [(32, 32, 1), (24, 20, 1), (12, 10, 9), (16, 16, 16), (5, 7, 3)])
def test_ch_rhs_bit_exact(gc, orc, ctx, shape):
    g, og = _grids(gc, orc, *shape, h=(1.0, 0.8, 1.2))
    psi = gc.inputs.ch_initial(*shape, 0.7, seed=5)
    got = gc.ch_rhs(ctx, g, gc.DeviceArray.from_numpy(ctx, psi), M=1.3, kappa=0.6).numpy(ctx)
    assert np.array_equal(got, orc.ch_rhs(og, psi, M=1.3, kappa=0.6))


@pytest.mark.parametrize("integ", ["rk4", "euler"])
@pytest.mark.parametrize("shape,steps", [((64, 64, 1), 200), ((16, 12, 10), 40), ((9, 1, 1), 30)])
def test_ch_integrate_bit_exact(gc, orc, ctx, integ, shape, steps):
    g, og = _grids(gc, orc, *shape)
    psi = gc.inputs.ch_initial(*shape, 0.05, seed=17)
    d = gc.DeviceArray.from_numpy(ctx, psi)
    method = gc.RK4 if integ == "rk4" else gc.EULER
    gc.ch_integrate(ctx, g, d, 1.0, 0.5, 0.01, steps, method)
    want = orc.ch_integrate(og, psi, 1.0, 0.5, 0.01, steps, orc.RK4 if integ == "rk4" else orc.EULER)
    assert np.array_equal(d.numpy(ctx), want)


def test_config1_ch_128_fp64(gc, orc, ctx):
    """Plain 2-D CH, FP64, 128x128, dx=1, M=kappa=1, c0 = 0.05*U(-1,1), RK4 dt=0.01,
    1000 steps: the GPU field equals the oracle's bit for bit, mass and F agree."""
    g, og = _grids(gc, orc, 128, 128, 1)
    c0 = gc.inputs.ch_initial(128, 128, 1, 0.05, seed=2024)
    d = gc.DeviceArray.from_numpy(ctx, c0)
    gc.ch_integrate(ctx, g, d, 1.0, 1.0, 0.01, 1000, gc.RK4)
    want = orc.ch_integrate(og, c0, 1.0, 1.0, 0.01, 1000, orc.RK4)
    got = d.numpy(ctx)
    assert np.array_equal(got, want)
    assert gc.integrate(ctx, g, d) == orc.integrate(og, want)
    assert gc.ch_free_energy(ctx, g, d, 1.0) == orc.free_energy(og, want, 1.0)
    # conservation: mass drift stays at rounding level
    m0 = orc.integrate(og, c0)
    assert abs(gc.integrate(ctx, g, d) - m0) < 1e-10 * g.size


def test_ch_fp32_vs_fp64(gc, orc, ctx):
    """FP32 stages against the FP64 oracle: L2 and Linf error reported and bounded."""
    g, og = _grids(gc, orc, 64, 64, 1)
    c0 = gc.inputs.ch_initial(64, 64, 1, 0.1, seed=9)
    d = gc.DeviceArray.from_numpy(ctx, c0.astype(np.float32))
    gc.ch_integrate(ctx, g, d, 1.0, 1.0, 0.01, 300, gc.RK4)
    want = orc.ch_integrate(og, c0.astype(np.float32).astype(np.float64), 1.0, 1.0, 0.01, 300, orc.RK4)
    got = d.numpy(ctx).astype(np.float64)
    l2 = np.linalg.norm(got - want) / np.linalg.norm(want)
    linf = np.abs(got - want).max() / np.abs(want).max()
    assert l2 < 1e-5 and linf < 1e-4, (l2, linf)


def test_ch_zero_rhs_constant_field(gc, ctx):
    # cahn_hilliard_test.cpp:47-58
    g = gc.Grid.make(8, 8, 8)
    rhs = gc.ch_rhs(ctx, g, gc.DeviceArray.from_numpy(ctx, np.full(g.shape, 0.42))).numpy(ctx)
    assert np.abs(rhs).max() < 1e-12


@pytest.mark.parametrize("integ", [0, 1])
@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 24, 1), (70, 9, 40), (8, 12, 5)])
def test_diffuse_const_bit_exact(gc, orc, ctx, integ, shape):
    h = 2 * np.pi / shape[0]
    g, og = _grids(gc, orc, *shape, h=h)
    psi = gc.inputs.symmetric(31, g.size).reshape(g.shape)
    d = gc.DeviceArray.from_numpy(ctx, psi)
    D, dt = 0.1, 0.2 * h * h
    gc.diffuse_const(ctx, g, d, D, dt, 25, integ)
    assert np.array_equal(d.numpy(ctx), orc.diffuse_const(og, psi, D, dt, 25, integ))


@pytest.mark.parametrize("integ", [0, 1])
@pytest.mark.parametrize("shape", [(16, 16, 16), (40, 8, 1), (70, 9, 40), (8, 12, 5)])
def test_diffuse_field_bit_exact(gc, orc, ctx, integ, shape):
    h = 2 * np.pi / shape[0]
    g, og = _grids(gc, orc, *shape, h=h)
    psi = gc.inputs.symmetric(41, g.size).reshape(g.shape)
    D = 0.05 + 0.1 * gc.inputs.uniform(42, g.size).reshape(g.shape)
    d = gc.DeviceArray.from_numpy(ctx, psi)
    dt = 0.3 * h * h
    gc.diffuse_field(ctx, g, d, gc.DeviceArray.from_numpy(ctx, D), dt, 20, integ)
    assert np.array_equal(d.numpy(ctx), orc.diffuse_field(og, psi, D, dt, 20, integ))


def test_diffuse_errors(gc, ctx):
    g = gc.Grid.make(16, 16, 16)
    d = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    with pytest.raises(gc.GridcalcError) as e:
        gc.diffuse_const(ctx, g, d, 1.0, 1.0, 1, gc.EULER)  # CFL 3 > 0.5
    assert e.value.code == gc.ERR_CFL and "solve::diffuse" in str(e.value)
    with pytest.raises(gc.GridcalcError) as e:
        gc.diffuse_const(ctx, g, d, -1.0, 0.1, 1)
    assert "D must be >= 0" in str(e.value)
    with pytest.raises(gc.GridcalcError) as e:
        gc.ch_integrate(ctx, g, d, 1.0, 1.0, 0.01, -1)
    assert "n_steps must be >= 0" in str(e.value) and "RK4" in str(e.value)
    with pytest.raises(gc.GridcalcError) as e:
        gc.ch_integrate(ctx, g, d, 1.0, 1.0, -0.01, 1, gc.EULER)
    assert "dt must be >= 0" in str(e.value) and "ExplicitEuler" in str(e.value)
    Dbad = np.ones(g.shape)
    Dbad.ravel()[5] = 0.0
    with pytest.raises(gc.GridcalcError) as e:
        gc.diffuse_field(ctx, g, d, gc.DeviceArray.from_numpy(ctx, Dbad), 0.01, 1)
    assert e.value.code == gc.ERR_NONPOS_D and "flat index 5" in str(e.value)
    # zero steps: no right-hand-side evaluation, hence no D check (Diffusion.hpp:210-215)
    gc.diffuse_field(ctx, g, d, gc.DeviceArray.from_numpy(ctx, Dbad), 0.01, 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pointwise_steps(gc, ctx, dtype):
    n = 100003
    a, b, c, e, f = (gc.inputs.symmetric(s, n).astype(dtype) for s in range(5))
    da, db, dc, de, df = (gc.DeviceArray.from_numpy(ctx, x) for x in (a, b, c, e, f))
    out = gc.DeviceArray((n,), dtype)
    s = dtype(0.37)
    gc.axpy(ctx, da, float(s), db, out)
    assert np.array_equal(out.numpy(ctx), a + s * b)
    gc.euler_update(ctx, da, float(s), db)
    a1 = a + s * b
    assert np.array_equal(da.numpy(ctx), a1)
    dt = 0.01
    gc.rk4_combine(ctx, da, dt, db, dc, de, df)
    sixth, third = dtype(dt / 6.0), dtype(dt / 3.0)
    assert np.array_equal(da.numpy(ctx), a1 + (sixth * (b + f) + third * (c + e)))


def test_snapshot_capture_while_stepping(gc, orc, ctx):
    """gc_snapshot: each capture holds the state at the time of the call, while the stepping
    queued right after it (on the context's stream) changes the field."""
    g = gc.Grid.make(40, 24, 16)
    og = orc.Grid.make(40, 24, 16)
    c0 = gc.inputs.ch_initial(40, 24, 16, 0.1, seed=8)
    d = gc.DeviceArray.from_numpy(ctx, c0)
    snaps = [gc.Snapshot(ctx, g.shape) for _ in range(3)]
    for s in snaps:
        s.capture(d)
        gc.ch_integrate(ctx, g, d, 1.0, 0.5, 0.01, 5, gc.RK4)
    want = c0
    for s in snaps:
        assert np.array_equal(s.numpy(), want)
        want = orc.ch_integrate(og, want, 1.0, 0.5, 0.01, 5, orc.RK4)
    assert np.array_equal(d.numpy(ctx), want)
    for s in snaps:
        s.close()
