"""GPU parity of the stencil engine against the CPU oracle.

FP64 results must be bit-identical to the oracle (same arithmetic order, no contraction).
FP32 results are compared with the FP64 oracle evaluated on the same (float-rounded)
input; the tolerance is stated per test.
"""
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (nx, ny, nz, h): cubic, anisotropic/ragged, 2-D, 1-D strip, degenerate x, multi-tile
SHAPES = [
    (16, 16, 16, (1.0, 1.0, 1.0)),
    (17, 12, 9, (0.3, 0.17, 0.5)),
    (40, 24, 1, (0.25, 0.5, 1.0)),
    (70, 1, 1, (0.1, 1.0, 1.0)),
    (1, 13, 11, (1.0, 0.2, 0.3)),
    (130, 7, 8, (1.0 / 130, 1.0 / 7, 1.0 / 8)),
    (72, 20, 70, (0.5, 0.25, 0.125)),  # several xy tiles and z chunks, ragged last chunk
]


def _field(gc, shape, seed, comps=1):
    nx, ny, nz = shape
    n = nx * ny * nz * comps
    a = gc.inputs.symmetric(seed, n)
    return a.reshape((nz, ny, nx) + ((comps,) if comps > 1 else ()))


def _grids(gc, orc, s):
    nx, ny, nz, h = s
    return gc.Grid.make(nx, ny, nz, h), orc.Grid.make(nx, ny, nz, h)


def _fits(s, radius):
    return all(n == 1 or n >= 2 * radius + 1 for n in s[:3])


OPS = [
    # name, gpu fn, oracle fn, in comps, radius(order)
    ("laplacian", "laplacian", "laplacian", 1, {2: 1, 4: 2}),
    ("gradient", "gradient", "gradient", 1, {2: 1, 4: 2}),
    ("gradient_vec", "gradient_vec", "gradient_vec", 3, {2: 1, 4: 2}),
    ("divergence", "divergence", "divergence", 3, {2: 1, 4: 2}),
    ("divergence_mat", "divergence_mat", "divergence_mat", 9, {2: 1, 4: 2}),
    ("biharmonic", "biharmonic", "biharmonic", 1, {2: 2, 4: 3}),
]


@pytest.mark.parametrize("op", OPS, ids=[o[0] for o in OPS])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}")
def test_operator_f64_bit_exact(gc, orc, ctx, op, order, shape):
    name, gfn, ofn, comps, radius = op
    if not _fits(shape, radius[order]):
        pytest.skip("extent aliases this stencil")
    g, og = _grids(gc, orc, shape)
    f = _field(gc, shape[:3], seed=zlib.crc32(f"{name}{order}".encode()) & 0xFFFF, comps=comps)
    d = gc.DeviceArray.from_numpy(ctx, f)
    got = getattr(gc, gfn)(ctx, g, d, order=order).numpy(ctx)
    want = getattr(orc, ofn)(og, f, order=order)
    assert got.shape == want.shape
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"


@pytest.mark.parametrize("op", OPS, ids=[o[0] for o in OPS])
@pytest.mark.parametrize("order", [2, 4])
def test_operator_f32_close(gc, orc, ctx, op, order):
    name, gfn, ofn, comps, radius = op
    shape = (24, 20, 18, (0.1, 0.2, 0.3))
    g, og = _grids(gc, orc, shape)
    f32 = _field(gc, shape[:3], seed=5, comps=comps).astype(np.float32)
    d = gc.DeviceArray.from_numpy(ctx, f32)
    got = getattr(gc, gfn)(ctx, g, d, order=order).numpy(ctx).astype(np.float64)
    want = getattr(orc, ofn)(og, f32.astype(np.float64), order=order)
    # single-precision rounding of ~(2r+1)^2 terms of size |w| |f| / h^p
    rel = np.abs(got - want).max() / np.abs(want).max()
    assert rel < 2e-5, rel


@pytest.mark.parametrize("order", [2, 4])
def test_biharmonic_f64_bit_exact_long_chunks(gc, orc, ctx, order):
    # large enough that the FP64 biharmonic marches 64-plane chunks (the small shapes above
    # run 8- to 32-plane chunks); ragged x and y tiles, a ragged last chunk of 8 planes
    shape = (250, 232, 520, (1.0 / 250, 1.0 / 232, 1.0 / 520))
    g, og = _grids(gc, orc, shape)
    f = _field(gc, shape[:3], seed=77 + order)
    got = gc.biharmonic(ctx, g, gc.DeviceArray.from_numpy(ctx, f), order=order).numpy(ctx)
    want = orc.biharmonic(og, f, order=order)
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cell_laplacian(gc, orc, ctx, shape, dtype):
    g, og = _grids(gc, orc, shape)
    n = g.size
    psi = gc.inputs.symmetric(3, n).reshape(g.shape).astype(dtype)
    D = gc.inputs.lognormal(4, n).reshape(g.shape).astype(dtype)
    got = gc.cell_laplacian(ctx, g, gc.DeviceArray.from_numpy(ctx, psi),
                            gc.DeviceArray.from_numpy(ctx, D)).numpy(ctx)
    want = orc.cell_laplacian(og, psi.astype(np.float64), D.astype(np.float64))
    if dtype == np.float64:
        assert np.array_equal(got, want)
    else:
        assert np.abs(got - want).max() / np.abs(want).max() < 1e-5


def test_cell_laplacian_known_answer(gc, ctx):
    # fvm_cell_laplacian_test.cpp:90-120: D = [1,4,4,1], psi = [0,1,1,0] -> [1.6,-1.6,-1.6,1.6]
    g = gc.Grid.make(4, 1, 1)
    psi = gc.DeviceArray.from_numpy(ctx, np.array([0.0, 1.0, 1.0, 0.0]).reshape(1, 1, 4))
    D = gc.DeviceArray.from_numpy(ctx, np.array([1.0, 4.0, 4.0, 1.0]).reshape(1, 1, 4))
    out = gc.cell_laplacian(ctx, g, psi, D).numpy(ctx).ravel()
    assert out.tolist() == [1.6, -1.6, -1.6, 1.6]


def test_cell_laplacian_rejects_nonpositive_d(gc, ctx):
    g = gc.Grid.make(8, 8, 8)
    D = np.ones(g.shape)
    D.ravel()[37] = 0.0
    D.ravel()[300] = -1.0
    psi = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    with pytest.raises(gc.GridcalcError) as e:
        gc.cell_laplacian(ctx, g, psi, gc.DeviceArray.from_numpy(ctx, D))
    assert e.value.code == gc.ERR_NONPOS_D
    assert "flat index 37" in str(e.value)
    assert gc.first_nonpositive(ctx, g, gc.DeviceArray.from_numpy(ctx, D)) == 37
    assert gc.first_nonpositive(ctx, g, gc.DeviceArray.from_numpy(ctx, np.ones(g.shape))) == -1


@pytest.mark.parametrize("shape,bad", [((70, 9, 40), [70 * 9 * 33 + 5, 70 * 9 * 39 + 69]),
                                       ((64, 8, 32), [64 * 8 * 31 + 64 * 7 + 63])])
def test_cell_laplacian_reports_first_bad_d_anywhere(gc, ctx, shape, bad):
    """The D > 0 check runs inside the operator kernel: the smallest failing index is found
    wherever it lies (tile edges, last z-chunk, the very last site); NaN fails too."""
    g = gc.Grid.make(*shape)
    D = np.ones(g.shape)
    for i, b in enumerate(bad):
        D.ravel()[b] = np.nan if i == 0 else -2.0
    psi = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    with pytest.raises(gc.GridcalcError) as e:
        gc.cell_laplacian(ctx, g, psi, gc.DeviceArray.from_numpy(ctx, D))
    assert e.value.code == gc.ERR_NONPOS_D
    assert f"flat index {min(bad)}" in str(e.value)


@pytest.mark.parametrize("axis,shape", [("x", (4, 16, 16)), ("y", (16, 3, 16)), ("z", (16, 16, 2))])
def test_extent_rejected(gc, ctx, axis, shape):
    # stencil_axis_extent_test.cpp: 2 <= N <= 2r throws, message names the axis
    g = gc.Grid.make(*shape)
    f = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    with pytest.raises(gc.GridcalcError) as e:
        gc.laplacian(ctx, g, f, order=4)
    assert e.value.code == gc.ERR_EXTENT
    assert f"axis {axis}" in str(e.value)


def test_bad_order_and_grid(gc, ctx):
    g = gc.Grid.make(8, 8, 8)
    f = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    with pytest.raises(gc.GridcalcError) as e:
        gc.laplacian(ctx, g, f, order=3)
    assert e.value.code == gc.ERR_INVALID
    bad = gc.Grid.make(8, 0, 8)
    with pytest.raises(gc.GridcalcError) as e:
        gc.laplacian(ctx, bad, f)
    assert "must be > 0" in str(e.value)


def test_max_face_d(gc, orc, ctx):
    g, og = _grids(gc, orc, (17, 12, 9, (1.0, 1.0, 1.0)))
    D = gc.inputs.lognormal(9, g.size).reshape(g.shape)
    assert gc.max_face_d(ctx, g, gc.DeviceArray.from_numpy(ctx, D)) == orc.max_face_d(og, D)
