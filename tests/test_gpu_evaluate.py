"""func::evaluate on the device (gc_evaluate): every recognised density against the oracle's
materialise-then-integrate restatement of func/Functional.hpp:88-220 -- pairwise bit for bit
in FP64 and on FP32 fields (the densities are formed in double from the stored values), Kahan
within its error bound -- on cubic, anisotropic, 2-D and ragged multi-block grids; the extent
checks of the operators the reference materialises."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [((16, 16, 16), (2 * np.pi / 16,) * 3), ((12, 10, 9), (0.8, 1.1, 1.3)),
          ((40, 24, 1), (1.0, 0.5, 1.0)), ((70, 9, 33), (1.0, 1.0, 1.0)),
          ((130, 67, 41), (0.3, 0.2, 0.5))]
SCALAR = [("value", 0.0, 0.0), ("gradient_energy", 0.0, 0.0), ("cahn_hilliard", 0.7, 0.0),
          ("ginzburg_landau", 1.3, 0.0), ("pfc", 1.1, 0.1)]


def _ids(x):
    return "x".join(map(str, x[0]))


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_scalar_densities_match_oracle(gc, orc, ctx, shape, dtype):
    (nx, ny, nz), h = shape
    g, og = gc.Grid.make(nx, ny, nz, h), orc.Grid.make(nx, ny, nz, h)
    f = (0.8 * gc.inputs.symmetric(21, g.size)).reshape(g.shape).astype(dtype)
    d = gc.DeviceArray.from_numpy(ctx, f)
    f64 = f.astype(np.float64)
    for kind, a, b in SCALAR:
        if kind == "pfc" and min(n for n in (nx, ny, nz) if n > 1) < 5:
            continue
        got = gc.evaluate(ctx, g, kind, d, a, b, gc.PAIRWISE)
        want = orc.evaluate(og, kind, f64, a, b, orc.PAIRWISE)
        assert got == want, (kind, got, want)
        gk = gc.evaluate(ctx, g, kind, d, a, b, gc.KAHAN)
        wk = orc.evaluate(og, kind, f64, a, b, orc.KAHAN)
        assert abs(gk - wk) <= 1e-13 * max(1.0, abs(wk)), (kind, gk, wk)
    d.free()


@pytest.mark.parametrize("shape", SHAPES[:4], ids=_ids)
def test_linear_elastic_matches_oracle(gc, orc, ctx, shape):
    (nx, ny, nz), h = shape
    g, og = gc.Grid.make(nx, ny, nz, h), orc.Grid.make(nx, ny, nz, h)
    u = (0.01 * gc.inputs.symmetric(31, 3 * g.size)).reshape(g.shape + (3,))
    d = gc.DeviceArray.from_numpy(ctx, u)
    for method in (gc.PAIRWISE, gc.KAHAN):
        got = gc.evaluate(ctx, g, "linear_elastic", d, 1.7, 0.6, method)
        want = orc.evaluate(og, "linear_elastic", u, 1.7, 0.6, method)
        if method == gc.PAIRWISE:
            assert got == want
        else:
            assert abs(got - want) <= 1e-13 * abs(want)


def test_free_energy_entry_points_agree(gc, ctx):
    """gc_ch_free_energy and gc_gradient_energy are the evaluate densities they name."""
    g = gc.Grid.make(33, 20, 14, (1.0, 0.9, 1.1))
    d = gc.DeviceArray.from_numpy(ctx, gc.inputs.symmetric(5, g.size).reshape(g.shape))
    for m in (gc.PAIRWISE, gc.KAHAN):
        assert gc.ch_free_energy(ctx, g, d, 0.6, m) == gc.evaluate(ctx, g, "cahn_hilliard", d, 0.6, 0, m)
        assert gc.gradient_energy(ctx, g, d, m) == gc.evaluate(ctx, g, "gradient_energy", d, 0, 0, m)


def test_closed_form_on_device(gc, ctx):
    """The exact discrete closed form of the phase-field-crystal density on
    sin x sin y sin z (the oracle test's formula), FP64 to 1e-12."""
    N = 16
    h = 2 * np.pi / N
    x = np.arange(N) * h
    s = np.sin(x)
    psi = s[None, None, :] * s[None, :, None] * s[:, None, None]
    g = gc.Grid.make(N, N, N, h)
    S2, S4 = (N / 2) ** 3, (3 * N / 8) ** 3
    s2 = (2 * np.cos(h) - 2) / h ** 2
    s4 = (2 * np.cos(2 * h) - 8 * np.cos(h) + 6) / h ** 4
    q0, eps = 1.1, 0.1
    want = (0.5 * (q0 ** 4 - eps) * S2 + q0 ** 2 * 3 * s2 * S2 + 0.5 * (3 * s4 + 6 * s2 * s2) * S2
            + 0.25 * S4) * h ** 3
    got = gc.evaluate(ctx, g, "pfc", gc.DeviceArray.from_numpy(ctx, psi), q0, eps)
    assert abs(got - want) <= 1e-12 * abs(want)


def test_extent_checks(gc, ctx):
    d = gc.DeviceArray.from_numpy(ctx, np.zeros(4 * 8 * 8))
    g = gc.Grid.make(4, 8, 8)  # radius 1 fits, the biharmonic's radius 2 does not
    gc.evaluate(ctx, g, "ginzburg_landau", d, 1.0)
    with pytest.raises(gc.GridcalcError) as e:
        gc.evaluate(ctx, g, "pfc", d, 1.0, 0.1)
    assert e.value.code == gc.ERR_EXTENT and "axis x" in str(e.value) and "radius 2" in str(e.value)
    g2 = gc.Grid.make(8, 2, 16)
    with pytest.raises(gc.GridcalcError) as e:
        gc.evaluate(ctx, g2, "cahn_hilliard", d, 1.0)
    assert "axis y" in str(e.value)
    gc.evaluate(ctx, g2, "value", d)  # arity 1 reads no neighbours
    D = gc.Density(99, 0.0, 0.0)
    r = C.c_double()
    with pytest.raises(gc.GridcalcError, match="unknown density kind"):
        gc._check(gc.lib().gc_evaluate(ctx.handle, C.byref(g), gc.F64, C.byref(D), 0, d.ptr,
                                       C.byref(r)))
    d.free()
