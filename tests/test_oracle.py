"""Pins the CPU oracle (oracle/gridcalc_oracle.c) against facts that do not come from its
own code: exact stencil rationals and moment conditions, the exact discrete symbols of
every operator on grid Fourier modes, the reference tests' known answers, hand-derived
known answers for the pairwise summation tree, and the RK4/Euler amplification
polynomials.  CPU only."""
from fractions import Fraction as Fr

import numpy as np
import pytest

# ---- stencil tables ---------------------------------------------------------------------

EXACT = {
    (2, 2): [Fr(1), Fr(-2), Fr(1)],
    (2, 4): [Fr(-1, 12), Fr(4, 3), Fr(-5, 2), Fr(4, 3), Fr(-1, 12)],
    (1, 2): [Fr(-1, 2), Fr(0), Fr(1, 2)],
    (1, 4): [Fr(1, 12), Fr(-2, 3), Fr(0), Fr(2, 3), Fr(-1, 12)],
    (3, 2): [Fr(-1, 2), Fr(1), Fr(0), Fr(-1), Fr(1, 2)],
    (3, 4): [Fr(1, 8), Fr(-1), Fr(13, 8), Fr(0), Fr(-13, 8), Fr(1), Fr(-1, 8)],
    (4, 2): [Fr(1), Fr(-4), Fr(6), Fr(-4), Fr(1)],
    (4, 4): [Fr(-1, 6), Fr(2), Fr(-13, 2), Fr(28, 3), Fr(-13, 2), Fr(2), Fr(-1, 6)],
}


@pytest.mark.parametrize("kind,order", list(EXACT))
def test_stencil_tables_exact(orc, kind, order):
    r, offs, w = orc.stencil_table(kind, order)
    ex = EXACT[(kind, order)]
    assert offs == list(range(-r, r + 1))
    assert w == [float(x) for x in ex]  # the reference's EXPECT_DOUBLE_EQ audits
    # moment conditions of a derivative of order `kind` with accuracy `order`:
    # sum w m^p = kind! for p = kind, 0 for other p < kind + order
    fact = {1: 1, 2: 2, 3: 6, 4: 24}[kind]
    for p in range(kind + order):
        mom = sum(wi * Fr(m) ** p for wi, m in zip(ex, offs))
        assert mom == (fact if p == kind else 0), (p, mom)


# ---- discrete symbols -------------------------------------------------------------------

def _sym1(order, th):   # first derivative, times h
    return np.sin(th) if order == 2 else 4.0 / 3.0 * np.sin(th) - 1.0 / 6.0 * np.sin(2 * th)


def _sym2(order, th):   # second derivative, times h^2
    if order == 2:
        return 2.0 * np.cos(th) - 2.0
    return -1.0 / 6.0 * np.cos(2 * th) + 8.0 / 3.0 * np.cos(th) - 5.0 / 2.0


def _sym4(order, th):   # fourth derivative, times h^4
    if order == 2:
        return 2.0 * np.cos(2 * th) - 8.0 * np.cos(th) + 6.0
    return -1.0 / 3.0 * np.cos(3 * th) + 4.0 * np.cos(2 * th) - 13.0 * np.cos(th) + 28.0 / 3.0


def _mode(shape, h, k, phase=0.3):
    nz, ny, nx = shape
    x = np.arange(nx) * h[0]
    y = np.arange(ny) * h[1]
    z = np.arange(nz) * h[2]
    L = (nx * h[0], ny * h[1], nz * h[2])
    arg = (2 * np.pi * k[0] / L[0] * x[None, None, :] + 2 * np.pi * k[1] / L[1] * y[None, :, None]
           + 2 * np.pi * k[2] / L[2] * z[:, None, None] + phase)
    th = [2 * np.pi * k[a] / (shape[2 - a]) for a in range(3)]  # k_a * h_a
    return arg, th


CASES = [((12, 10, 14), (0.3, 0.2, 0.5), (2, 1, 3)),
         ((9, 16, 1), (1.0, 0.5, 1.0), (1, 3, 0)),
         ((20, 7, 11), (0.05, 0.1, 0.2), (3, 2, 4))]


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("case", CASES)
def test_laplacian_symbol(orc, order, case):
    (nx, ny, nz), h, k = case
    g = orc.Grid.make(nx, ny, nz, h)
    arg, th = _mode(g.shape, h, k)
    f = np.sin(arg)
    lam = sum(_sym2(order, th[a]) / h[a] ** 2 for a in range(3))
    out = orc.laplacian(g, f, order)
    assert np.abs(out - lam * f).max() < 1e-12 * max(1.0, abs(lam))


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("case", CASES)
def test_gradient_symbol(orc, order, case):
    (nx, ny, nz), h, k = case
    g = orc.Grid.make(nx, ny, nz, h)
    arg, th = _mode(g.shape, h, k)
    out = orc.gradient(g, np.sin(arg), order)
    for a in range(3):
        want = _sym1(order, th[a]) / h[a] * np.cos(arg)
        assert np.abs(out[..., a] - want).max() < 1e-12 / h[a]


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("case", CASES)
def test_divergence_symbol(orc, order, case):
    (nx, ny, nz), h, k = case
    g = orc.Grid.make(nx, ny, nz, h)
    arg, th = _mode(g.shape, h, k)
    phases = (0.1, 0.7, 1.9)
    v = np.stack([np.sin(arg + p) for p in phases], axis=-1)
    want = sum(_sym1(order, th[a]) / h[a] * np.cos(arg + phases[a]) for a in range(3))
    assert np.abs(orc.divergence(g, v, order) - want).max() < 1e-11 / min(h)


@pytest.mark.parametrize("order", [2, 4])
def test_vector_gradient_and_tensor_divergence_symbol(orc, order):
    (nx, ny, nz), h, k = CASES[0]
    g = orc.Grid.make(nx, ny, nz, h)
    arg, th = _mode(g.shape, h, k)
    phases = (0.1, 0.7, 1.9)
    v = np.stack([np.sin(arg + p) for p in phases], axis=-1)
    M = orc.gradient_vec(g, v, order)
    for comp in range(3):
        for ax in range(3):  # column-major: M(comp, ax) at comp + 3*ax
            want = _sym1(order, th[ax]) / h[ax] * np.cos(arg + phases[comp])
            assert np.abs(M[..., comp + 3 * ax] - want).max() < 1e-11 / min(h)
    # tensor field with m(r, c) = sin(arg + q_rc): div gives v_r = sum_c d_c m(r, c)
    q = np.arange(9).reshape(3, 3) * 0.37
    m = np.stack([np.sin(arg + q[i % 3, i // 3]) for i in range(9)], axis=-1)
    dv = orc.divergence_mat(g, m, order)
    for r in range(3):
        want = sum(_sym1(order, th[c]) / h[c] * np.cos(arg + q[r, c]) for c in range(3))
        assert np.abs(dv[..., r] - want).max() < 1e-10 / min(h)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("case", CASES)
def test_biharmonic_symbol(orc, order, case):
    (nx, ny, nz), h, k = case
    if order == 4 and min(n for n in (nx, ny, nz) if n > 1) < 7:
        pytest.skip("extent too small for radius 3")
    g = orc.Grid.make(nx, ny, nz, h)
    arg, th = _mode(g.shape, h, k)
    f = np.sin(arg)
    d2 = [_sym2(order, th[a]) / h[a] ** 2 for a in range(3)]
    lam = sum(_sym4(order, th[a]) / h[a] ** 4 for a in range(3))
    lam += 2 * (d2[0] * d2[1] + d2[0] * d2[2] + d2[1] * d2[2])
    out = orc.biharmonic(g, f, order)
    assert np.abs(out - lam * f).max() < 1e-11 * max(1.0, abs(lam))


def test_degenerate_axis_terms_vanish(orc):
    """2-D (Nz = 1): the z terms of every operator are exactly zero."""
    g3 = orc.Grid.make(16, 12, 1, (0.5, 0.25, 0.7))
    f = np.sin(np.arange(16 * 12)).reshape(1, 12, 16)
    assert np.all(orc.gradient(g3, f)[..., 2] == 0.0)
    g1 = orc.Grid.make(16, 12, 1, (0.5, 0.25, 123.0))  # hz must not matter
    assert np.array_equal(orc.laplacian(g3, f), orc.laplacian(g1, f))


def test_extent_rule(orc):
    for n, ok in [(1, True), (2, False), (3, True), (4, True)]:
        g = orc.Grid.make(n, 8, 8)
        if ok:
            orc.laplacian(g, np.zeros(g.shape), 2)
        else:
            with pytest.raises(orc.OracleError):
                orc.laplacian(g, np.zeros(g.shape), 2)
    with pytest.raises(orc.OracleError):
        g = orc.Grid.make(8, 8, 6)
        orc.biharmonic(g, np.zeros(g.shape), 4)  # radius 3 needs N >= 7


# ---- finite volume ----------------------------------------------------------------------

def test_fvm_known_answer(orc):
    # fvm_cell_laplacian_test.cpp:90-120
    g = orc.Grid.make(4, 1, 1)
    out = orc.cell_laplacian(g, [0.0, 1.0, 1.0, 0.0], [1.0, 4.0, 4.0, 1.0])
    assert out.ravel().tolist() == [1.6, -1.6, -1.6, 1.6]


def test_fvm_constant_d_matches_laplacian(orc, gc):
    # fvm_cell_laplacian_test.cpp:30-70: D = c reduces to c * laplacian
    g = orc.Grid.make(10, 12, 8, (0.3, 0.2, 0.4))
    psi = gc.inputs.symmetric(1, g.size).reshape(g.shape)
    c = 2.5
    out = orc.cell_laplacian(g, psi, np.full(g.shape, c))
    assert np.abs(out - c * orc.laplacian(g, psi)).max() < 1e-12 * np.abs(out).max()


def test_fvm_conservative(orc, gc):
    # fluxes telescope: the cell Laplacian sums to zero
    g = orc.Grid.make(9, 8, 7)
    psi = gc.inputs.symmetric(2, g.size).reshape(g.shape)
    D = gc.inputs.lognormal(3, g.size).reshape(g.shape)
    out = orc.cell_laplacian(g, psi, D)
    assert abs(out.sum()) < 1e-12 * np.abs(out).sum()


def test_fvm_nonpositive(orc):
    g = orc.Grid.make(4, 4, 4)
    D = np.ones(g.shape)
    D.ravel()[9] = -0.5
    with pytest.raises(orc.OracleError, match="flat index 9"):
        orc.cell_laplacian(g, np.zeros(g.shape), D)


# ---- reductions -------------------------------------------------------------------------

TWO53 = 2.0 ** 53


def test_pairwise_leaf_of_64(orc):
    # n = 64 is one leaf summed left to right: every +1 after 2^53 rounds away
    x = [TWO53] + [1.0] * 63
    assert orc.pairwise_sum(x) == TWO53
    assert orc.neumaier_sum(x) == TWO53 + 63


def test_pairwise_65_splits(orc):
    # n = 65 is split: the right half's 33 ones survive as 32 (ties to even)
    x = [TWO53] + [1.0] * 64
    assert orc.pairwise_sum(x) == TWO53 + 32


def test_pairwise_split_is_floor_half(orc):
    # split at n/2 = 32: left = 32 ones, right = 2^53 + 32 ones -> 2^53; total 2^53 + 32.
    # A split at 33 would give 2^53 + 64.
    x = [1.0] * 32 + [TWO53] + [1.0] * 32
    assert orc.pairwise_sum(x) == TWO53 + 32


def test_pairwise_deeper_tree(orc):
    # n = 256: four leaves of 64.  Leaf 1 = 2^53 (the ones above 2^53 round away),
    # leaf 2 = 64, leaf 3 = -2^53 + 63 (exact: spacing below 2^53 is 1), leaf 4 = 64;
    # tree (2^53 + 64) + (-2^53 + 127) = 191.  Leaves of 128 would give 127.
    x = [TWO53] + [1.0] * 63 + [1.0] * 64 + [-TWO53] + [1.0] * 63 + [1.0] * 64
    assert orc.pairwise_sum(x) == 191.0


def test_neumaier_known_answers(orc):
    assert orc.neumaier_sum([1e100, 1.0, -1e100]) == 1.0
    assert orc.neumaier_sum([1.0, 1e100, 1.0, -1e100]) == 2.0
    assert orc.pairwise_sum([1e100, 1.0, -1e100]) == 0.0


def test_integrate_known_answers(orc):
    # integrate_test.cpp:31-48
    g = orc.Grid.make(8, 16, 4, (0.25, 0.125, 0.5))
    ones = np.ones(g.shape)
    vol = (8 * 0.25) * (16 * 0.125) * (4 * 0.5)
    assert orc.integrate(g, ones, orc.PAIRWISE) == vol
    assert orc.integrate(g, ones, orc.KAHAN) == vol
    # integrate_test.cpp:81-101: pairwise vs Kahan within 1e-13 relative
    n = 64
    h = 2 * np.pi / n
    x = np.arange(n) * h
    f = np.broadcast_to(1.0 + np.sin(x)[None, None, :] * np.cos(x)[None, :, None], (n, n, n))
    gp = orc.Grid.make(n, n, n, h)
    a, b = orc.integrate(gp, f, orc.PAIRWISE), orc.integrate(gp, f, orc.KAHAN)
    assert abs(a - b) / max(abs(a), abs(b)) < 1e-13
    assert abs(a - (2 * np.pi) ** 3) < 1e-9


def test_free_energy_closed_form(orc):
    """F for psi = A sin(k x) on [0, L)^3: the O2 discrete gradient is A sin(th)/h cos(kx),
    so F = V*(-A^2/4 + 3A^4/32 + kappa/4 * A^2 sin(th)^2 / h^2) (grid sums of sin^2, sin^4,
    cos^2 over whole periods)."""
    n, A, kappa = 16, 0.7, 0.9
    h = 2 * np.pi / n
    g = orc.Grid.make(n, n, n, h)
    x = np.arange(n) * h
    psi = np.broadcast_to(A * np.sin(2 * x)[None, None, :], (n, n, n)).copy()
    V = (n * h) ** 3
    th = 2 * h
    want = V * (-A * A / 4 + 3 * A ** 4 / 32 + kappa / 4 * A * A * np.sin(th) ** 2 / h ** 2)
    assert abs(orc.free_energy(g, psi, kappa) - want) < 1e-12 * abs(want) * 10
    wantE = V * A * A * np.sin(th) ** 2 / h ** 2 / 2
    assert abs(orc.gradient_energy(g, psi) - wantE) < 1e-11 * wantE


# ---- time integration ---------------------------------------------------------------------

def _g_rk4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


@pytest.mark.parametrize("method", [0, 1])
def test_diffusion_eigenmode_amplification(orc, method):
    """psi0 = sin x sin y sin z is an eigenvector of the O2 Laplacian with eigenvalue
    lam = 3*(2cos h - 2)/h^2, so n steps multiply it by g(D dt lam)^n exactly
    (g = 1 + z for Euler, the RK4 polynomial for RK4) up to rounding."""
    N = 12
    h = 2 * np.pi / N
    g = orc.Grid.make(N, N, N, h)
    x = np.arange(N) * h
    psi = np.sin(x)[None, None, :] * np.sin(x)[None, :, None] * np.sin(x)[:, None, None]
    lam = 3 * (2 * np.cos(h) - 2) / h ** 2
    D, dt, n = 0.1, 0.3 * h * h, 40
    out = orc.diffuse_const(g, psi, D, dt, n, method)
    z = D * dt * lam
    amp = (_g_rk4(z) if method == 1 else 1 + z) ** n
    assert np.abs(out - amp * psi).max() < 1e-13


def test_rk4_time_order(orc):
    """integrator_test.cpp:154-205 analogue on a grid eigenmode: halving dt cuts the
    time-discretisation error by ~16 (order 4) against the exact exp(D lam T)."""
    N = 8
    h = 2 * np.pi / N
    g = orc.Grid.make(N, N, N, h)
    x = np.arange(N) * h
    psi = np.sin(x)[None, None, :] * np.sin(x)[None, :, None] * np.sin(x)[:, None, None]
    lam = 3 * (2 * np.cos(h) - 2) / h ** 2
    D, T = 1.0, 1.0
    errs = []
    for n in (4, 8, 16):
        dt = T / n
        out = orc.diffuse_const(g, psi, D, dt, n, 1) if D * dt * 3 / h ** 2 <= 0.6963 else None
        if out is None:
            continue
        errs.append(np.abs(out - np.exp(D * lam * T) * psi).max())
    assert len(errs) >= 2
    for a, b in zip(errs, errs[1:]):
        assert 12 < a / b < 20


def test_cfl_rule(orc):
    g = orc.Grid.make(8, 8, 8, 0.5)
    # D*dt*(3/h^2) = 12*D*dt
    assert not orc.cfl_violated(g, 1.0, 0.5 / 12, 0.5)
    assert orc.cfl_violated(g, 1.0, 0.5 / 12 * 1.001, 0.5)
    with pytest.raises(orc.OracleError):
        orc.diffuse_const(g, np.zeros(g.shape), 1.0, 0.1, 1, 0)


def test_ch_rhs_invariants(orc, gc):
    # cahn_hilliard_test.cpp:47-78: zero for a constant field; conservative
    g = orc.Grid.make(8, 8, 8)
    assert np.abs(orc.ch_rhs(g, np.full(g.shape, 0.42))).max() < 1e-12
    g = orc.Grid.make(16, 16, 16)
    psi = gc.inputs.ch_initial(16, 16, 16, 0.3, seed=123)
    r = orc.ch_rhs(g, psi)
    assert abs(r.sum()) < 1e-10 * np.abs(r).max() * r.size


def test_ch_rhs_linear_symbol(orc):
    """For small amplitude the CH right-hand side of A sin(kx) is, to O(A^3),
    M*lam*(-1 - kappa*lam)*A sin(kx) with lam the O2 Laplacian symbol."""
    n, A, M, kappa = 16, 1e-6, 1.3, 0.7
    g = orc.Grid.make(n, 1, 1)
    x = np.arange(n)
    psi = A * np.sin(2 * np.pi * 3 * x / n).reshape(1, 1, n)
    lam = 2 * np.cos(2 * np.pi * 3 / n) - 2
    want = M * lam * (-1 - kappa * lam) * psi
    assert np.abs(orc.ch_rhs(g, psi, M, kappa) - want).max() < 1e-15


def test_ch_free_energy_decreases(orc, gc):
    g = orc.Grid.make(24, 24, 1)
    psi = gc.inputs.ch_initial(24, 24, 1, 0.3, seed=5)
    F = [orc.free_energy(g, psi, 1.0)]
    for _ in range(4):
        psi = orc.ch_integrate(g, psi, 1.0, 1.0, 0.02, 60, orc.RK4)
        F.append(orc.free_energy(g, psi, 1.0))
    assert all(b <= a for a, b in zip(F, F[1:])), F


# ---- closed forms added for the CH right-hand side and the face-D scan --------------------

def _plane_wave(n3, h3, k3, amp, phase=0.3):
    """amp*sin(k.x + phase) on an (nx, ny, nz) grid with integer mode numbers k3; returns the
    (nz, ny, nx) field and theta = k.x + phase."""
    (nx, ny, nz), (hx, hy, hz) = n3, h3
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    th = 2 * np.pi * (k3[0] * i / nx + k3[1] * j / ny + k3[2] * k / nz) + phase
    return amp * np.sin(th), th


def _lap2_symbol(n3, h3, k3, mult=1):
    """Eigenvalue of the order-2 discrete Laplacian on sin(mult*(k.x) + phase)."""
    return sum((2 * np.cos(2 * np.pi * mult * kk / nn) - 2) / (hh * hh)
               for nn, hh, kk in zip(n3, h3, k3))


@pytest.mark.parametrize("n3,h3,k3,amp", [
    ((16, 1, 1), (1.0, 1.0, 1.0), (3, 0, 0), 0.6),
    ((12, 10, 9), (0.8, 1.1, 1.3), (2, 1, 4), 0.9),
    ((20, 16, 1), (1.0, 1.0, 1.0), (1, 5, 0), 1.2),
])
def test_ch_rhs_two_mode_closed_form(orc, n3, h3, k3, amp):
    """psi = A sin(th) with th = k.x + phase.  sin^3 th = (3 sin th - sin 3th)/4, and a plane
    wave of mode m*k is an eigenvector of the O2 Laplacian with eigenvalue L(m*k), so
      mu  = (-A - kappa*L(k)*A + 3A^3/4) sin th - (A^3/4) sin 3th
      rhs = M*[L(k)*(-A - kappa*L(k)*A + 3A^3/4) sin th - L(3k)*(A^3/4) sin 3th]
    exactly (aliasing of 3k on the grid changes nothing: the symbol is evaluated at the grid's
    own mode).  Pins the cubic term's sign and factor, kappa, M and the per-axis spacing."""
    M, kappa = 1.3, 0.7
    g = orc.Grid.make(*n3, h3)
    psi, th = _plane_wave(n3, h3, k3, amp)
    L1, L3 = _lap2_symbol(n3, h3, k3, 1), _lap2_symbol(n3, h3, k3, 3)
    want = M * (L1 * (-amp - kappa * L1 * amp + 0.75 * amp ** 3) * np.sin(th)
                - L3 * 0.25 * amp ** 3 * np.sin(3 * th))
    got = orc.ch_rhs(g, psi, M, kappa)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    # the same field with the cubic term's sign flipped differs by far more than rounding
    wrong = M * (L1 * (-amp - kappa * L1 * amp - 0.75 * amp ** 3) * np.sin(th)
                 + L3 * 0.25 * amp ** 3 * np.sin(3 * th))
    assert np.abs(got - wrong).max() > 1e-3 * np.abs(want).max()


def test_max_face_d_known_answers(orc):
    """Largest harmonic-mean face D (solve/Diffusion.hpp:187-207), hand-computed:
    - x strip D = [1, 4, 4, 1]: x faces 1.6, 4, 1.6, 1; the degenerate y and z axes wrap to
      the cell itself, so their faces are hmean(D, D) = D: maximum 4;
    - 2x2x2 checkerboard of 1 and 9: every face joins 1 and 9, 2*9/10 = 1.8 (an arithmetic
      face mean would give 5, max(D) 9);
    - D = 1 except two cells of 8 adjacent along one axis a: the face between them is 8, all
      others at most hmean(8, 1) = 16/9; scanning every axis but a gives 16/9."""
    g = orc.Grid.make(4, 1, 1)
    assert orc.max_face_d(g, [1.0, 4.0, 4.0, 1.0]) == 4.0
    g = orc.Grid.make(2, 2, 2)
    i, j, k = np.meshgrid(np.arange(2), np.arange(2), np.arange(2), indexing="ij")
    D = np.where((i + j + k) % 2 == 0, 1.0, 9.0).transpose(2, 1, 0)
    assert orc.max_face_d(g, D) == 2.0 * 1.0 * 9.0 / 10.0
    g = orc.Grid.make(5, 6, 7)
    for axis in range(3):
        D = np.ones(g.shape)  # (nz, ny, nx)
        a = [2, 3, 4]  # (i, j, k) of the first cell
        b = list(a)
        b[axis] += 1
        D[a[2], a[1], a[0]] = 8.0
        D[b[2], b[1], b[0]] = 8.0
        assert orc.max_face_d(g, D) == 8.0, axis
    # a periodic face: the last and first cells along y
    D = np.ones(g.shape)
    D[0, 0, 0] = D[0, g.ny - 1, 0] = 3.0
    assert orc.max_face_d(g, D) == 3.0
    # no pair of large neighbours: the largest face joins 8 and 1
    D = np.ones(g.shape)
    D[3, 3, 3] = 8.0
    assert orc.max_face_d(g, D) == 2.0 * 8.0 / 9.0


def test_hetero_diffusion_matches_constant_d(orc):
    """diffusion_test.cpp:216-237: with D uniform the heterogeneous path follows the
    constant-D trajectory within 1e-12 (sin x cos y sin z, N = 16, 50 Euler steps)."""
    N = 16
    h = 2 * np.pi / N
    g = orc.Grid.make(N, N, N, h)
    x = np.arange(N) * h
    psi = np.sin(x)[None, None, :] * np.cos(x)[None, :, None] * np.sin(x)[:, None, None]
    D, dt, n = 0.1, 0.4 * h * h / (6 * 0.1), 50
    a = orc.diffuse_const(g, psi, D, dt, n, orc.EULER)
    b = orc.diffuse_field(g, psi, np.full(g.shape, D), dt, n, orc.EULER)
    assert np.abs(a - b).max() < 1e-12
    a = orc.diffuse_const(g, psi, D, dt, n, orc.RK4)
    b = orc.diffuse_field(g, psi, np.full(g.shape, D), dt, n, orc.RK4)
    assert np.abs(a - b).max() < 1e-12


def test_hetero_cfl_uses_max_face_d(orc):
    """diffusion_test.cpp:239-254: D = 0.5 + 0.5 cos x (bumped to 1e-6 where it vanishes),
    N = 32: the largest face D is hmean of the two cells beside x = 0, a hair under 1, so
    dt = h^2/3 (CFL 1 > 0.5) is rejected and dt at 0.499 of the bound by that face passes."""
    N = 32
    h = 2 * np.pi / N
    g = orc.Grid.make(N, N, N, h)
    x = np.arange(N) * h
    Dx = 0.5 + 0.5 * np.cos(x)
    Dx[Dx <= 0.0] = 1e-6
    D = np.broadcast_to(Dx[None, None, :], g.shape).copy()
    face = max(2 * Dx[i] * Dx[(i + 1) % N] / (Dx[i] + Dx[(i + 1) % N]) for i in range(N))
    # y and z faces join two cells of equal D: hmean(D, D), which rounds like any face mean
    face = max(face, max(2 * d * d / (d + d) for d in Dx))
    assert orc.max_face_d(g, D) == face
    psi = np.ones(g.shape)
    with pytest.raises(orc.OracleError) as e:
        orc.diffuse_field(g, psi, D, h * h / 3.0, 1, orc.EULER)
    assert e.value.code == orc.CFL
    ok_dt = 0.499 * h * h / (3 * face)
    orc.diffuse_field(g, psi, D, ok_dt, 1, orc.EULER)


def test_ch_energy_rise_from_white_noise(orc, gc):
    """From white noise c0 = 0.1*U(-1,1) (kappa = 0.5, dt = 0.005, 3-D) the reference's free
    energy F (central-difference gradient) rises at first while the semi-discrete energy
    E = h^3 * sum(-c^2/2 + c^4/4 - kappa/2 * c * lap(c)), with the compact O2 Laplacian the
    dynamics use, falls at every step checked: dE/dt = mu . lap(mu) <= 0 for
    dc/dt = lap(mu), mu = -c + c^3 - kappa*lap(c), the Laplacian being symmetric negative
    semidefinite.  A central difference sees no gradient in a near-Nyquist mode, so the
    fast decay of those modes removes energy -c^2/2 < 0 from F and F increases."""
    n = 24
    g = orc.Grid.make(n, n, n)
    c = gc.inputs.ch_initial(n, n, n, 0.1, seed=3)
    kappa = 0.5

    def E(c):
        return float(np.sum(-0.5 * c * c + 0.25 * c ** 4 - 0.5 * kappa * c * orc.laplacian(g, c)))
    Es, Fs = [E(c)], [orc.free_energy(g, c, kappa)]
    for _ in range(10):
        c = orc.ch_integrate(g, c, 1.0, kappa, 0.005, 25, orc.RK4)
        Es.append(E(c))
        Fs.append(orc.free_energy(g, c, kappa))
    assert all(b < a for a, b in zip(Es, Es[1:])), Es
    assert Fs[-1] > Fs[0], Fs


# ---- func::evaluate densities: exact discrete closed forms ---------------------------------

def _sss(N):
    h = 2 * np.pi / N
    x = np.arange(N) * h
    s, c = np.sin(x), np.cos(x)
    return h, x, s, c


@pytest.mark.parametrize("N", [16, 12])
def test_evaluate_scalar_densities_closed_form(orc, N):
    """psi = sin x sin y sin z on [0, 2pi)^3.  The central difference gives
    g = (sin h / h) (cos x sin y sin z, ...), lap psi = L psi, biharm psi = B psi with the O2
    symbols L = 3 s2, B = 3 s4 + 6 s2^2 (s2 = (2 cos h - 2)/h^2, s4 = (2 cos 2h - 8 cos h + 6)/h^4),
    and whole-period grid sums give sum psi^2 = (N/2)^3, sum psi^4 = (3N/8)^3,
    sum |g|^2 = 3 (sin h/h)^2 (N/2)^3; every density's integral follows exactly."""
    h, x, s, c = _sss(N)
    g = orc.Grid.make(N, N, N, h)
    psi = s[None, None, :] * s[None, :, None] * s[:, None, None]
    V = h ** 3
    S2, S4 = (N / 2) ** 3, (3 * N / 8) ** 3
    G2 = 3 * (np.sin(h) / h) ** 2 * S2
    s2 = (2 * np.cos(h) - 2) / h ** 2
    s4 = (2 * np.cos(2 * h) - 8 * np.cos(h) + 6) / h ** 4
    L, B = 3 * s2, 3 * s4 + 6 * s2 * s2
    kappa, W, q0, eps = 0.7, 1.3, 1.1, 0.1
    want = {
        ("value", 0.0, 0.0): 0.25 * N ** 3 * V,  # on psi + 0.25
        ("gradient_energy", 0.0, 0.0): G2 * V,
        ("cahn_hilliard", kappa, 0.0): (-0.5 * S2 + 0.25 * S4 + 0.5 * kappa * G2) * V,
        ("ginzburg_landau", W, 0.0): (0.5 * G2 + W * (S4 - 2 * S2 + N ** 3)) * V,
        ("pfc", q0, eps): (0.5 * (q0 ** 4 - eps) * S2 + q0 ** 2 * L * S2 + 0.5 * B * S2
                           + 0.25 * S4) * V,
    }
    for (kind, a, b), w in want.items():
        f = psi + 0.25 if kind == "value" else psi
        for method in (orc.PAIRWISE, orc.KAHAN):
            got = orc.evaluate(g, kind, f, a, b, method)
            assert abs(got - w) <= 1e-12 * max(abs(w), G2 * V), (kind, got, w)
    # a plausible slip (PFC without the biharmonic's mixed terms) is far outside that bound
    wrong = (0.5 * (q0 ** 4 - eps) * S2 + q0 ** 2 * L * S2 + 0.5 * 3 * s4 * S2 + 0.25 * S4) * V
    assert abs(orc.evaluate(g, "pfc", psi, q0, eps) - wrong) > 1e-3 * abs(wrong)


def test_evaluate_linear_elastic_closed_form(orc):
    """u = (sin x + sin y, sin z, sin x): G = grad u has G00 = s cos x, G01 = s cos y,
    G12 = s cos z, G20 = s cos x (s = sin h / h); e = sym(G) gives tr e = s cos x and
    e:e = s^2 (1.5 cos^2 x + 0.5 cos^2 y + 0.5 cos^2 z), so the integral of
    (lambda tr^2 + 2 mu e:e)/2 is h^3 (lambda s^2 N^3/2 + 2 mu 1.25 s^2 N^3)/2 exactly.  Using
    G instead of its symmetric part would give 2 mu (1.5 + 1 + 1) s^2 N^3 / 2 instead."""
    N = 16
    h, x, sx, cx = _sss(N)
    g = orc.Grid.make(N, N, N, h)
    X = np.broadcast_to(x[None, None, :], (N, N, N))
    Y = np.broadcast_to(x[None, :, None], (N, N, N))
    Z = np.broadcast_to(x[:, None, None], (N, N, N))
    u = np.stack([np.sin(X) + np.sin(Y), np.sin(Z), np.sin(X)], axis=-1)
    lam, mu = 1.7, 0.6
    s = np.sin(h) / h
    want = h ** 3 * 0.5 * (lam * s * s * N ** 3 / 2 + 2 * mu * 1.25 * s * s * N ** 3)
    got = orc.evaluate(g, "linear_elastic", u, lam, mu)
    assert abs(got - want) <= 1e-12 * abs(want)


def test_evaluate_reference_hand_computed(orc):
    """The reference tests' continuum answers at N = 64 (functional_test.cpp:27-56,
    pfc_functional_test.cpp:52-101) within their own tolerances."""
    N = 64
    h, x, s, c = _sss(N)
    g = orc.Grid.make(N, N, N, h)
    psi = s[None, None, :] * s[None, :, None] * s[:, None, None]
    pi3 = np.pi ** 3
    gl = orc.evaluate(g, "ginzburg_landau", psi, 1.0)
    assert abs(gl - 507 / 64 * pi3) < 1e-2 * 507 / 64 * pi3
    pfc = orc.evaluate(g, "pfc", psi, 1.0, 0.1)
    assert abs(pfc - 2.05546875 * pi3) < 2e-2 * 2.05546875 * pi3
    assert abs(orc.evaluate(g, "gradient_energy", psi) / 2 - 1.5 * pi3) < 1e-2 * 1.5 * pi3


# ---- mixed partial derivatives ---------------------------------------------------------------

# every multi-index with 1 <= nx + ny + nz <= 4 (the 31 named partials plus the three first
# derivatives)
MULTI = [(a, b, c) for a in range(5) for b in range(5) for c in range(5) if 1 <= a + b + c <= 4]


def _axis_symbol(n, order, t):
    """sum_m w_m e^{i m t} of the n-th derivative stencil (identity for n = 0)."""
    if n == 0:
        return 1.0 + 0j
    w = EXACT[(n, order)]
    r = (len(w) - 1) // 2
    return sum(float(wm) * np.exp(1j * (m - r) * t) for m, wm in enumerate(w))


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("multi", MULTI, ids=lambda m: "".join(map(str, m)))
def test_mixed_partial_symbol(orc, order, multi):
    """On sin(th), th = k.x + phase, the tensor-product stencil acts as the product of the
    per-axis complex symbols S_a = sum_m w_m e^{i m k_a h_a}: the result is
    Im(prod_a S_a / h_a^n_a * e^{i th}) exactly (to rounding)."""
    (nx, ny, nz), h, k = (14, 11, 9), (0.3, 0.2, 0.5), (2, 1, 3)
    g = orc.Grid.make(nx, ny, nz, h)
    arg, th = _mode(g.shape, h, k)
    S = 1.0 + 0j
    for a in range(3):
        S *= _axis_symbol(multi[a], order, th[a]) / h[a] ** multi[a]
    want = np.imag(S * np.exp(1j * arg))
    got = orc.mixed_partial(g, np.sin(arg), *multi, order)
    assert np.abs(got - want).max() < 1e-10 * max(1.0, abs(S)), (multi, order)


def test_mixed_partial_known_values_and_extents(orc):
    """Closed forms of the named partials the reference tests sweep: on
    psi = sin x sin y sin z, d2/dxdy at O2 gives (sin h / h)^2 cos x cos y sin z; the
    third-derivative stencil's symbol at O2 is i (sin 2t - 2 sin t) / h^3; and the per-axis
    extent rule uses that axis's radius (an axis of order 0 accepts any N)."""
    N = 16
    h = 2 * np.pi / N
    g = orc.Grid.make(N, N, N, h)
    x = np.arange(N) * h
    s, c = np.sin(x), np.cos(x)
    psi = s[None, None, :] * s[None, :, None] * s[:, None, None]
    want = (np.sin(h) / h) ** 2 * c[None, None, :] * c[None, :, None] * s[:, None, None]
    assert np.abs(orc.mixed_partial(g, psi, 1, 1, 0, 2) - want).max() < 1e-12
    want3 = (np.sin(2 * h) - 2 * np.sin(h)) / h ** 3 * c[None, None, :] * s[None, :, None] * s[:, None, None]
    assert np.abs(orc.mixed_partial(g, psi, 3, 0, 0, 2) - want3).max() < 1e-9
    g2 = orc.Grid.make(5, 2, 16)  # y of extent 2 cannot take any radius >= 1
    orc.mixed_partial(g2, np.zeros(g2.shape), 2, 0, 2, 2)
    with pytest.raises(orc.OracleError):
        orc.mixed_partial(g2, np.zeros(g2.shape), 1, 1, 0, 2)
    with pytest.raises(orc.OracleError):  # third derivative O4 has radius 3: N = 5 too small
        orc.mixed_partial(g2, np.zeros(g2.shape), 3, 0, 0, 4)


# ---- tensor contraction restatements ----------------------------------------------------------

def test_tensor_restatements_pinned(orc):
    """The per-site trace, symmetric part, matrix product and double contraction of the oracle
    (column-major storage) against numpy's own trace, transpose, matmul and einsum."""
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((2, 50, 9))
    A = a.reshape(50, 3, 3).transpose(0, 2, 1)  # column-major storage -> A[r, c]
    B = b.reshape(50, 3, 3).transpose(0, 2, 1)
    assert np.allclose(orc.mat_trace(a), np.trace(A, axis1=1, axis2=2), rtol=0, atol=1e-14)
    assert np.allclose(orc.mat_sym(a).reshape(50, 3, 3).transpose(0, 2, 1),
                       0.5 * (A + A.transpose(0, 2, 1)), rtol=0, atol=1e-15)
    assert np.allclose(orc.mat_mul(a, b).reshape(50, 3, 3).transpose(0, 2, 1), A @ B, rtol=0,
                       atol=1e-14)
    assert np.allclose(orc.mat_ddot(a, b), np.einsum("nij,nij->n", A, B), rtol=0, atol=1e-14)
    # a transposed operand in the product would not pass
    assert not np.allclose(orc.mat_mul(a, b).reshape(50, 3, 3).transpose(0, 2, 1),
                           A.transpose(0, 2, 1) @ B, atol=1e-6)
