"""Every tile shape of the fused Cahn-Hilliard RK4 kernels -- one kernel per stage and, in 3-D,
two stages per kernel -- gives the oracle's result bit for bit (FP64) and the same bits as
every other shape (FP32); the autotuner picks a valid shape and keeps results unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# ragged in x and y against every tile width/height; several z-chunks with a short last one
SHAPES_3D = [(36, 20, 70), (16, 16, 8), (70, 9, 33)]


def _run(gc, ctx, g, c0, steps, dtype):
    d = gc.DeviceArray.from_numpy(ctx, c0.astype(dtype))
    gc.ch_integrate(ctx, g, d, 1.0, 0.7, 0.01, steps, gc.RK4)
    return d.numpy(ctx)


@pytest.mark.parametrize("shape", SHAPES_3D, ids=lambda s: "x".join(map(str, s)))
def test_every_3d_config_bit_exact_fp64(gc, orc, ctx, shape):
    g = gc.Grid.make(*shape, (1.0, 0.9, 1.1))
    og = orc.Grid.make(*shape, (1.0, 0.9, 1.1))
    c0 = gc.inputs.ch_initial(*shape, 0.3, seed=41)
    steps = 5  # odd: the two-stage path ends in its second state buffer and copies back
    want = orc.ch_integrate(og, c0, 1.0, 0.7, 0.01, steps, orc.RK4)
    n = gc.ch_config_count(g)
    assert n > 0
    fused = [gc.ch_config_fused(g, i) for i in range(n)]
    assert any(fused) and not all(fused)
    try:
        for i in range(n):
            gc.ch_set_config(g, gc.F64, i)
            got = _run(gc, ctx, g, c0, steps, np.float64)
            assert np.array_equal(got, want), (i, gc.ch_config_shape(g, i), fused[i])
    finally:
        gc.ch_set_config(g, gc.F64, -1)


# 2-D: ragged against every tile, and smaller than the 8-site halo of the one-kernel-per-step
# candidates (the halo wraps around the grid more than once)
@pytest.mark.parametrize("shape", [(70, 37, 1), (16, 9, 1)], ids=lambda s: "x".join(map(str, s)))
def test_every_2d_config_bit_exact_fp64(gc, orc, ctx, shape):
    g = gc.Grid.make(*shape, (1.0, 0.9, 1.1))
    og = orc.Grid.make(*shape, (1.0, 0.9, 1.1))
    c0 = gc.inputs.ch_initial(*shape, 0.3, seed=47)
    steps = 5  # odd: the one-kernel-per-step path ends in its second buffer and copies back
    want = orc.ch_integrate(og, c0, 1.0, 0.7, 0.01, steps, orc.RK4)
    n = gc.ch_config_count(g)
    try:
        for i in range(n):
            gc.ch_set_config(g, gc.F64, i)
            got = _run(gc, ctx, g, c0, steps, np.float64)
            assert np.array_equal(got, want), (i, gc.ch_config_shape(g, i), gc.ch_config_kind(g, i))
    finally:
        gc.ch_set_config(g, gc.F64, -1)


@pytest.mark.parametrize("shape", [(36, 20, 70), (64, 40, 1)], ids=lambda s: "x".join(map(str, s)))
def test_every_config_same_bits_fp32(gc, ctx, shape):
    g = gc.Grid.make(*shape)
    c0 = gc.inputs.ch_initial(*shape, 0.3, seed=43)
    outs = []
    try:
        for i in range(gc.ch_config_count(g)):
            gc.ch_set_config(g, gc.F32, i)
            outs.append(_run(gc, ctx, g, c0, 4, np.float32))
    finally:
        gc.ch_set_config(g, gc.F32, -1)
    for i, o in enumerate(outs[1:], 1):
        assert np.array_equal(o, outs[0]), i


def test_tuner_choice_keeps_results(gc, orc, ctx):
    shape = (48, 24, 16)
    g = gc.Grid.make(*shape)
    og = orc.Grid.make(*shape)
    c0 = gc.inputs.ch_initial(*shape, 0.3, seed=45)
    try:
        best, ms = gc.ch_tune(ctx, g, gc.F64)
        assert 0 <= best < gc.ch_config_count(g)
        assert len(ms) == gc.ch_config_count(g) and all(m > 0 for m in ms)
        assert gc.ch_get_config(g, gc.F64) == best
        got = _run(gc, ctx, g, c0, 3, np.float64)
        assert np.array_equal(got, orc.ch_integrate(og, c0, 1.0, 0.7, 0.01, 3, orc.RK4))
    finally:
        gc.ch_set_config(g, gc.F64, -1)


def test_config_queries(gc):
    g3 = gc.Grid.make(32, 32, 32)
    n = gc.ch_config_count(g3)
    assert gc.ch_config_count(gc.Grid.make(32, 32, 4)) == 0  # untiled
    with pytest.raises(gc.GridcalcError):
        gc.ch_config_fused(g3, n)
    with pytest.raises(gc.GridcalcError):
        gc.ch_set_config(g3, gc.F64, n)
    g2 = gc.Grid.make(64, 64, 1)
    kinds = [gc.ch_config_kind(g2, i) for i in range(gc.ch_config_count(g2))]
    assert 0 in kinds and 3 in kinds and set(kinds) <= {0, 3}  # one stage / one step per kernel
