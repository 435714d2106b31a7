"""The tensor-expression kernel (gc_expr_eval) against the oracle's per-site restatements of
the tensor::expr nodes (trace, symmetric part, single and double contraction, scaled sums),
bit for bit, on a multi-block grid; programs are written directly as instruction lists, as
gridcalc/tensor/Expressions.hpp lowers them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOAD, IDENT, SCALE, ADD, SUB, MUL, TRACE, SYM, MATMUL, DDOT = range(10)


def _insn(gc, op, dst, a=0, b=0, nc=1, c=0.0, ptr=None):
    return gc.ExprInsn(op, dst, a, b, nc, c, ptr)


def _run(gc, ctx, g, prog, out_slot, out_nc):
    return gc.expr_eval(ctx, g, prog, out_slot, out_nc).numpy(ctx)


def test_expression_programs_bit_exact(gc, orc, ctx):
    g = gc.Grid.make(37, 21, 19)
    n = g.size
    A = gc.inputs.symmetric(1, 9 * n).reshape(g.shape + (9,))
    B = gc.inputs.symmetric(2, 9 * n).reshape(g.shape + (9,))
    s = gc.inputs.symmetric(3, n).reshape(g.shape)
    dA, dB, ds = (gc.DeviceArray.from_numpy(ctx, x) for x in (A, B, s))
    I = lambda *a, **k: _insn(gc, *a, **k)  # noqa: E731,E741
    L = lambda slot, d, nc: I(LOAD, slot, nc=nc, ptr=d.ptr.value)  # noqa: E731
    # trace(sym(A)) * s + 0.5 * (A : B)
    prog = [L(0, dA, 9), I(SYM, 0, 0, nc=9), I(TRACE, 0, 0), L(1, ds, 1), I(MUL, 0, 0, 1),
            L(1, dA, 9), L(2, dB, 9), I(DDOT, 1, 1, 2), I(SCALE, 1, 1, c=0.5), I(ADD, 0, 0, 1)]
    want = orc.mat_trace(orc.mat_sym(A)) * s + 0.5 * orc.mat_ddot(A, B)
    assert np.array_equal(_run(gc, ctx, g, prog, 0, 1), want)
    # A B - 2 I
    prog = [L(0, dA, 9), L(1, dB, 9), I(MATMUL, 0, 0, 1, nc=9), I(IDENT, 1, nc=9),
            I(SCALE, 1, 1, nc=9, c=2.0), I(SUB, 0, 0, 1, nc=9)]
    eye = np.zeros(9)
    eye[[0, 4, 8]] = 1.0
    want = orc.mat_mul(A, B) - 2.0 * eye
    assert np.array_equal(_run(gc, ctx, g, prog, 0, 9), want)


def test_program_validation(gc, ctx):
    g = gc.Grid.make(8, 8, 8)
    d = gc.DeviceArray.from_numpy(ctx, np.zeros(g.shape))
    bad = [[_insn(gc, LOAD, 9, ptr=d.ptr.value)],        # slot out of range
           [_insn(gc, LOAD, 0, nc=2, ptr=d.ptr.value)],  # width not 1, 3 or 9
           [_insn(gc, LOAD, 0)],                         # null field
           [_insn(gc, 42, 0)]]                           # unknown op
    for prog in bad:
        with pytest.raises(gc.GridcalcError):
            _run(gc, ctx, g, prog, 0, 1)
