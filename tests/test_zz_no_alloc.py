"""gc_integrate_noalloc (partial sums in a fixed per-device array, result left in device
memory, no allocation) against gc_integrate: the same bits, pairwise and Kahan, FP64 and FP32,
on two calls in a row with the second input smaller than the first; then calls queued back to
back on one stream before a single synchronisation, each reading its own result."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(256, 256, 256), (37, 29, 11)]  # a large input, then a smaller one


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("method", [0, 1], ids=["pairwise", "kahan"])
def test_noalloc_matches_integrate(gc, ctx, dtype, method):
    res = gc.DeviceArray((1,), np.float64)
    for seed, shape in enumerate(SHAPES):
        g = gc.Grid.make(*shape, (0.5, 0.25, 2.0))
        print("noalloc case:", shape, np.dtype(dtype).name, method, flush=True)
        f = gc.DeviceArray.from_numpy(ctx, gc.inputs.symmetric(100 + seed, g.size)
                                      .reshape(g.shape).astype(dtype))
        gc.integrate_noalloc(ctx, g, f, method, res)
        ctx.sync()
        got = float(res.numpy(ctx)[0])
        want = gc.integrate(ctx, g, f, method)
        assert got == want, (shape, got, want)
        f.free()
    res.free()


def test_noalloc_queued_calls(gc, ctx):
    """Calls ordered on one stream reuse the array without interfering: five integrals of
    different fields queued before one synchronisation."""
    outs = gc.DeviceArray((5,), np.float64)
    fields, wants = [], []
    for i in range(5):
        shape = (64 + 7 * i, 48, 9 + i)
        g = gc.Grid.make(*shape)
        f = gc.DeviceArray.from_numpy(ctx, gc.inputs.symmetric(7 + i, g.size).reshape(g.shape))
        fields.append((g, f))
        wants.append(gc.integrate(ctx, g, f, gc.PAIRWISE))
    for i, (g, f) in enumerate(fields):
        out_i = C.c_void_p(outs.ptr.value + 8 * i)
        gc._check(gc.lib().gc_integrate_noalloc(ctx.handle, C.byref(g), gc.F64, gc.PAIRWISE,
                                                f.ptr, out_i))
    got = outs.numpy(ctx)
    assert [float(x) for x in got] == wants
    for _, f in fields:
        f.free()
    outs.free()


def test_noalloc_rejects_bad_arguments(gc, ctx):
    g = gc.Grid.make(8, 8, 8)
    f = gc.DeviceArray.from_numpy(ctx, np.ones(g.shape))
    res = gc.DeviceArray((1,), np.float64)
    with pytest.raises(gc.GridcalcError, match="null result"):
        gc._check(gc.lib().gc_integrate_noalloc(ctx.handle, C.byref(g), gc.F64, 0, f.ptr, None))
    with pytest.raises(gc.GridcalcError, match="Tag must be"):
        gc.integrate_noalloc(ctx, g, f, 7, res)
    big = gc.Grid.make(65536, 65536, 1)  # 2^32 sites: beyond the library's grid limit
    with pytest.raises(gc.GridcalcError, match="2\\^31-1"):
        gc.integrate_noalloc(ctx, big, f, 0, res)
    f.free()
    res.free()
