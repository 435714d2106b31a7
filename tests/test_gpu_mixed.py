"""Mixed partial derivatives on the device (gc_mixed_partial) against the oracle's restatement
of diff/detail/MultiIndexDerivative.hpp:96-136: every multi-index with 1 <= nx+ny+nz <= 4 at
Order 2 and 4, FP64 bit for bit on anisotropic, multi-tile/multi-chunk and 2-D grids, FP32
within a stated bound; the per-axis extent rule."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MULTI = [(a, b, c) for a in range(5) for b in range(5) for c in range(5) if 1 <= a + b + c <= 4]
SHAPES = [((14, 11, 9), (0.3, 0.2, 0.5)), ((70, 9, 33), (1.0, 1.0, 1.0)),
          ((40, 24, 1), (0.5, 0.25, 1.0))]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s[0])))
@pytest.mark.parametrize("order", [2, 4])
def test_every_multi_index_bit_exact(gc, orc, ctx, shape, order):
    (nx, ny, nz), h = shape
    g, og = gc.Grid.make(nx, ny, nz, h), orc.Grid.make(nx, ny, nz, h)
    f = gc.inputs.symmetric(13, g.size).reshape(g.shape)
    d64 = gc.DeviceArray.from_numpy(ctx, f)
    f32 = f.astype(np.float32)
    d32 = gc.DeviceArray.from_numpy(ctx, f32)
    for m in MULTI:
        try:
            want = orc.mixed_partial(og, f, *m, order)
        except orc.OracleError:  # an axis too short for its radius: the device must refuse too
            with pytest.raises(gc.GridcalcError) as e:
                gc.mixed_partial(ctx, g, d64, *m, order)
            assert e.value.code == gc.ERR_EXTENT
            continue
        got = gc.mixed_partial(ctx, g, d64, *m, order).numpy(ctx)
        assert np.array_equal(got, want), (m, order)
        # FP32: rounding of the stencil's terms, bounded by their magnitude
        # max|f| * prod_a sum_m |w_a,m| / h_a^n_a (the result itself may cancel to 0)
        w32 = orc.mixed_partial(og, f32.astype(np.float64), *m, order)
        g32 = gc.mixed_partial(ctx, g, d32, *m, order).numpy(ctx).astype(np.float64)
        terms = np.abs(f32).max()
        for a in range(3):
            if m[a]:
                terms *= sum(abs(w) for w in orc.stencil_table(m[a], order)[2]) / h[a] ** m[a]
        assert np.abs(g32 - w32).max() <= 1e-6 * terms, (m, order)


def test_argument_checks(gc, ctx):
    g = gc.Grid.make(8, 8, 8)
    d = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    for bad in [(0, 0, 0), (3, 2, 0), (-1, 1, 0)]:
        with pytest.raises(gc.GridcalcError, match="total of 1 to 4"):
            gc.mixed_partial(ctx, g, d, *bad)
    with pytest.raises(gc.GridcalcError, match="Order must be 2 or 4"):
        gc.mixed_partial(ctx, g, d, 1, 1, 0, 3)
    g2 = gc.Grid.make(5, 2, 16)
    d2 = gc.DeviceArray.from_numpy(ctx, np.zeros(g2.shape))
    gc.mixed_partial(ctx, g2, d2, 2, 0, 2)  # y has order 0: its extent of 2 is fine
    with pytest.raises(gc.GridcalcError) as e:
        gc.mixed_partial(ctx, g2, d2, 1, 1, 0)
    assert e.value.code == gc.ERR_EXTENT and "axis y" in str(e.value)
