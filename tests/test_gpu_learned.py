"""Learned time stepping with the plain-RK4 fallback (gc_learned_integrate), on the GPU.

The acceptance rule: a macro step's proposal is kept unless the free energy rises or the mean
drifts by more than mass_tol; otherwise the macro step is recomputed by the fused RK4, so a
run in which every proposal is rejected is bit-identical to plain RK4, and a network that
proposes no change leaves the state untouched.  The trained-network tests show what the
network gains over the method without it (c(t + 16 dt) ~ c(t)) on held-out trajectories."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = W = 256
DT = 0.01


def _ic(gc, seed, dtype=np.float64):
    return (0.05 * gc.inputs.symmetric(seed, H * W, 1.0)).reshape(1, H, W).astype(dtype)


def _zero_net(gc):
    return np.zeros(gc.N_CNN_PARAMS, np.float32)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_all_rejected_is_plain_rk4(gc, ctx, dtype):
    """Random weights with a large output scale raise the free energy at every macro step:
    every step falls back, and the result equals plain RK4 bit for bit."""
    g = gc.Grid.make(W, H, 1)
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.5, 0.5, gc.N_CNN_PARAMS).astype(np.float32)
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(p, 1.0)
    c0 = _ic(gc, 11, dtype)
    a = gc.DeviceArray.from_numpy(ctx, c0)
    b = gc.DeviceArray.from_numpy(ctx, c0)
    nf = net.integrate(g, a, 1.0, 1.0, DT, 16, 5)
    gc.ch_integrate(ctx, g, b, 1.0, 1.0, DT, 80, gc.RK4)
    assert nf == 5
    assert np.array_equal(a.numpy(ctx), b.numpy(ctx))
    net.close()


def test_zero_network_keeps_state(gc, ctx):
    """A network with all-zero parameters proposes c + 0: accepted (F unchanged, no drift)
    and the state is unchanged."""
    g = gc.Grid.make(W, H, 1)
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(_zero_net(gc), 1e-2)
    c0 = _ic(gc, 12)
    a = gc.DeviceArray.from_numpy(ctx, c0)
    nf = net.integrate(g, a, 1.0, 1.0, DT, 16, 4)
    assert nf == 0
    assert np.array_equal(a.numpy(ctx), c0)
    net.close()


def test_bias_only_network_is_mass_free(gc, ctx):
    """A constant network output is removed by the mean projection: no change either."""
    g = gc.Grid.make(W, H, 1)
    p = _zero_net(gc)
    p[-1] = 3.0  # b6
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(p, 1e-2)
    c0 = _ic(gc, 13)
    a = gc.DeviceArray.from_numpy(ctx, c0)
    assert net.integrate(g, a, 1.0, 1.0, DT, 16, 3) == 0
    assert np.abs(a.numpy(ctx) - c0).max() < 1e-9
    net.close()


def test_mass_tolerance_forces_fallback(gc, ctx):
    """With mass_tol < 0 no proposal can pass: every macro step is plain RK4."""
    g = gc.Grid.make(W, H, 1)
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(_zero_net(gc), 1e-2)
    c0 = _ic(gc, 14)
    a = gc.DeviceArray.from_numpy(ctx, c0)
    b = gc.DeviceArray.from_numpy(ctx, c0)
    assert net.integrate(g, a, 1.0, 1.0, DT, 16, 3, mass_tol=-1.0) == 3
    gc.ch_integrate(ctx, g, b, 1.0, 1.0, DT, 48, gc.RK4)
    assert np.array_equal(a.numpy(ctx), b.numpy(ctx))
    net.close()


@pytest.mark.parametrize("n_macro,dtype", [(48, np.float64), (45, np.float64), (48, np.float32)])
def test_mixed_decisions_match_stepwise_rule(gc, ctx, n_macro, dtype):
    """A random network with a tiny output scale proposes c + O(1e-6) changes that sometimes
    raise the free energy: accepted and rejected macro steps interleave across several
    16-step proposal chunks (the next chunk proposed while the host decides, discarded after
    a rejection; 45 leaves a shorter last chunk), so a rejection inside a chunk discards the
    proposals after it.  The result must equal a host loop that decides after every macro
    step (free energy and mass from the library's reductions), for an FP64 and an FP32
    state."""
    g = gc.Grid.make(W, H, 1)
    rng = np.random.default_rng(5)
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(rng.uniform(-0.3, 0.3, gc.N_CNN_PARAMS).astype(np.float32), 1e-6)
    c0 = _ic(gc, 15, dtype)
    tol = 1e-6
    a = gc.DeviceArray.from_numpy(ctx, c0)
    nf = net.integrate(g, a, 1.0, 1.0, DT, 16, n_macro, tol)

    psi = gc.DeviceArray.from_numpy(ctx, c0)
    prop = gc.DeviceArray.from_numpy(ctx, c0)
    F0 = gc.ch_free_energy(ctx, g, psi, 1.0, gc.KAHAN)
    m_init = gc.integrate(ctx, g, psi, gc.KAHAN)
    want_nf = 0
    for _ in range(n_macro):
        net.predict(psi, prop, 1)
        F1 = gc.ch_free_energy(ctx, g, prop, 1.0, gc.KAHAN)
        m1 = gc.integrate(ctx, g, prop, gc.KAHAN)
        if np.isfinite(F1) and F1 <= F0 and abs(m1 - m_init) / (H * W) <= tol:
            psi.copy_from(ctx, prop)
            F0 = F1
        else:
            want_nf += 1
            gc.ch_integrate(ctx, g, psi, 1.0, 1.0, DT, 16, gc.RK4)
            F0 = gc.ch_free_energy(ctx, g, psi, 1.0, gc.KAHAN)
    print(f"fallbacks {nf} of {n_macro}")
    assert 1 <= want_nf <= n_macro - 1
    assert nf == want_nf
    assert np.array_equal(a.numpy(ctx), psi.numpy(ctx))
    net.close()


def test_rejects_wrong_grid(gc, ctx):
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(_zero_net(gc), 1e-2)
    a = gc.DeviceArray((1, 128, 256), np.float64)
    with pytest.raises(gc.GridcalcError, match="corrector was built"):
        net.integrate(gc.Grid.make(256, 128, 1), a, 1.0, 1.0, DT, 16, 1)
    net.close()


@pytest.fixture(scope="module")
def trained(gc):
    from gridcalc_amd import corrector as cr
    p, meta = cr.load_params(H, W)
    if p is None:
        pytest.skip("no trained corrector parameters in the package data")
    return cr, p, meta


def test_trained_network_beats_identity_one_step(gc, ctx, trained):
    """Held-out trajectories (seeds never used in training): the trained network's one
    macro-step error is well below that of the method without it, and below the untrained
    network's."""
    cr, p, meta = trained
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(p, meta["scale"])
    e_net, e_id = cr.one_step_errors(net, n_traj=2, log=lambda *a: None)
    net.set_params(cr.init_params(cr.CONFIG["seed"]), meta["scale"])
    e_init, _ = cr.one_step_errors(net, n_traj=2, log=lambda *a: None)
    net.close()
    print(f"one macro step rel L2: trained {e_net:.3e}, identity {e_id:.3e}, untrained {e_init:.3e}")
    assert e_net < 0.02 * e_id
    assert e_net < 0.02 * e_init


def test_trained_learned_integration(gc, ctx, trained):
    """T = 20 from a held-out initial state with the single network: mass is conserved to the
    tolerance, and the learned path stays within 5e-3 (relative L2) of the FP64 RK4 solution."""
    cr, p, meta = trained
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(p, meta["scale"])
    r = cr.integration_error(net, n_steps=2000)
    net.close()
    print("learned integration to T=20:", r)
    assert r["mass_drift"] <= 1e-6
    assert r["rel_l2"] < 5e-3
    assert r["fallbacks"] < r["n_macro"]


def test_phased_meets_accuracy_target(gc, ctx):
    """The learned-corrector target on four held-out initial states (seeds never used in
    training): the phase networks reach T = 20 (125 macro steps of 16 RK4 steps) within a
    relative L2 error of 1e-3 of the FP64 RK4 solution on every state, with no fallback and
    the mass conserved.  The speed ratio against the fastest plain GPU RK4 (the one-member
    ensemble kernel replayed in HIP graphs, tile shape tuned) is measured beside it (median of
    three runs each) and printed; it is asserted only loosely (>= 3x) so that the test does not
    depend on the box's clocks; the bench records the measured ratio."""
    from gridcalc_amd import corrector as cr
    nets = cr.phase_nets(ctx, H, W)
    if nets is None:
        pytest.skip("no phase parameters in the package data")
    g = gc.Grid.make(W, H, 1)
    seeds = [991001 + i for i in range(4)]

    def timed(fn):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        return time.perf_counter() - t0

    ens = gc.Ensemble(ctx, [dict(nx=W, ny=H, kappa=1.0, M=1.0)], DT, gc.F64)
    ens.tune()
    errs, t_plain, t_learned = [], [], []
    for k, seed in enumerate(seeds):
        c0 = cr.initial_state(H, W, seed).astype(np.float64)
        ens.set_state(0, c0)
        if k == 0:
            ens.step(32)
            ens.set_state(0, c0)
        t_plain.append(timed(lambda: ens.step(2000)))
        r = ens.get_state(0).reshape(c0.shape)
        d = gc.DeviceArray.from_numpy(ctx, c0)
        if k == 0:
            cr.integrate_phased(ctx, g, d, nets, 1.0, 1.0, DT, 16, 4)
            d.upload(ctx, c0)
        nf = [0]
        t_learned.append(timed(lambda: nf.__setitem__(
            0, cr.integrate_phased(ctx, g, d, nets, 1.0, 1.0, DT, 16, 125))))
        a = d.numpy(ctx)
        errs.append(float(np.linalg.norm(a - r) / np.linalg.norm(r)))
        assert nf[0] == 0
        assert abs(a.mean() - c0.mean()) <= 1e-6
        d.free()
    for k in range(2):  # two more timed runs of each path on the first state
        c0 = cr.initial_state(H, W, seeds[0]).astype(np.float64)
        ens.set_state(0, c0)
        t_plain.append(timed(lambda: ens.step(2000)))
        d = gc.DeviceArray.from_numpy(ctx, c0)
        t_learned.append(timed(lambda: cr.integrate_phased(ctx, g, d, nets, 1.0, 1.0, DT, 16, 125)))
        d.free()
    ens.close()
    for n, _, _ in nets:
        n.close()
    ratio = float(np.median(t_plain) / np.median(t_learned))
    print(f"T=20 phased rel L2 per state {['%.3e' % e for e in errs]}; plain graph RK4 "
          f"{np.median(t_plain) * 1e3:.2f} ms, learned {np.median(t_learned) * 1e3:.2f} ms, "
          f"ratio {ratio:.2f}")
    assert max(errs) <= 1e-3
    assert ratio >= 3.0


def test_phased_all_warm_is_plain_rk4(gc, ctx):
    """integrate_phased with every macro step inside the warm-up is plain RK4, bit for bit."""
    from gridcalc_amd import corrector as cr
    g = gc.Grid.make(W, H, 1)
    c0 = _ic(gc, 21)
    a = gc.DeviceArray.from_numpy(ctx, c0)
    b = gc.DeviceArray.from_numpy(ctx, c0)
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(_zero_net(gc), 1.0)
    assert cr.integrate_phased(ctx, g, a, [(net, 0, 100)], 1.0, 1.0, DT, 16, 3, warm=3) == 0
    gc.ch_integrate(ctx, g, b, 1.0, 1.0, DT, 48, gc.RK4)
    assert np.array_equal(a.numpy(ctx), b.numpy(ctx))
    net.close()


def test_phased_schedule(gc, ctx):
    """Networks that propose no change on macro steps [2, 4) and [5, 7) of 8, one warm step:
    the state equals plain RK4 over the other four macro steps (0, 1, 4, 7), so every step
    outside the phases is plain RK4 and every step inside is the phase's network."""
    from gridcalc_amd import corrector as cr
    g = gc.Grid.make(W, H, 1)
    c0 = _ic(gc, 22)
    a = gc.DeviceArray.from_numpy(ctx, c0)
    b = gc.DeviceArray.from_numpy(ctx, c0)
    nets = []
    for lo, hi in ((2, 4), (5, 7)):
        n = gc.Corrector(ctx, H, W, 1)
        n.set_params(_zero_net(gc), 1.0)
        nets.append((n, lo, hi))
    assert cr.integrate_phased(ctx, g, a, nets, 1.0, 1.0, DT, 16, 8, warm=1) == 0
    gc.ch_integrate(ctx, g, b, 1.0, 1.0, DT, 4 * 16, gc.RK4)
    assert np.array_equal(a.numpy(ctx), b.numpy(ctx))
    for n, _, _ in nets:
        n.close()


def test_phased_trained_networks_beat_single(gc, ctx, trained):
    """T = 20 from a held-out initial state: the phase-specialised networks (first macro step
    plain) are more accurate than the single trained network with the same plain first step,
    and conserve mass."""
    cr, p, meta = trained
    nets = cr.phase_nets(ctx, H, W)
    if nets is None:
        pytest.skip("no phase parameters in the package data")
    g = gc.Grid.make(W, H, 1)
    c0 = cr.initial_state(H, W, 991007).astype(np.float64)
    ref = gc.DeviceArray.from_numpy(ctx, c0)
    ph = gc.DeviceArray.from_numpy(ctx, c0)
    one = gc.DeviceArray.from_numpy(ctx, c0)
    gc.ch_integrate(ctx, g, ref, 1.0, 1.0, DT, 2000, gc.RK4)
    nf = cr.integrate_phased(ctx, g, ph, nets, 1.0, 1.0, DT, 16, 125)
    net = gc.Corrector(ctx, H, W, 1)
    net.set_params(p, meta["scale"])
    gc.ch_integrate(ctx, g, one, 1.0, 1.0, DT, 16, gc.RK4)
    net.integrate(g, one, 1.0, 1.0, DT, 16, 124)
    r, a, b = ref.numpy(ctx), ph.numpy(ctx), one.numpy(ctx)
    e_ph = np.linalg.norm(a - r) / np.linalg.norm(r)
    e_one = np.linalg.norm(b - r) / np.linalg.norm(r)
    print(f"T=20 rel L2: phased {e_ph:.3e}, single network {e_one:.3e}, fallbacks {nf}")
    assert e_ph < 0.9 * e_one
    assert abs(a.mean() - c0.mean()) <= 1e-6
    net.close()
    for n, _, _ in nets:
        n.close()


def test_phases_mixed_decisions_match_stepwise_rule(gc, ctx):
    """gc_learned_integrate_phases with two random networks (tiny output scale) on macro
    steps [0, 20) and [20, 45): accepts and rejects interleave and the proposal pipeline
    crosses the segment boundary; the result equals a host loop that decides after every
    macro step with the segment's network (mass drift from the initial state)."""
    from gridcalc_amd import corrector as cr
    g = gc.Grid.make(W, H, 1)
    nets = []
    for seed, lo, hi in ((5, 0, 20), (6, 20, 45)):
        n = gc.Corrector(ctx, H, W, 1)
        n.set_params(np.random.default_rng(seed).uniform(-0.3, 0.3, gc.N_CNN_PARAMS)
                     .astype(np.float32), 1e-6)
        nets.append((n, lo, hi))
    c0 = _ic(gc, 16)
    tol, n_macro = 1e-6, 45
    a = gc.DeviceArray.from_numpy(ctx, c0)
    nf = cr.integrate_phased(ctx, g, a, nets, 1.0, 1.0, DT, 16, n_macro, tol, warm=0)

    psi = gc.DeviceArray.from_numpy(ctx, c0)
    prop = gc.DeviceArray.from_numpy(ctx, c0)
    F0 = gc.ch_free_energy(ctx, g, psi, 1.0, gc.KAHAN)
    m_init = gc.integrate(ctx, g, psi, gc.KAHAN)
    want_nf = 0
    for m in range(n_macro):
        net = nets[0][0] if m < 20 else nets[1][0]
        net.predict(psi, prop, 1)
        F1 = gc.ch_free_energy(ctx, g, prop, 1.0, gc.KAHAN)
        m1 = gc.integrate(ctx, g, prop, gc.KAHAN)
        if np.isfinite(F1) and F1 <= F0 and abs(m1 - m_init) / (H * W) <= tol:
            psi.copy_from(ctx, prop)
            F0 = F1
        else:
            want_nf += 1
            gc.ch_integrate(ctx, g, psi, 1.0, 1.0, DT, 16, gc.RK4)
            F0 = gc.ch_free_energy(ctx, g, psi, 1.0, gc.KAHAN)
    print(f"fallbacks {nf} of {n_macro}")
    assert 1 <= want_nf <= n_macro - 1
    assert nf == want_nf
    assert np.array_equal(a.numpy(ctx), psi.numpy(ctx))
    for n, _, _ in nets:
        n.close()


def test_phases_rejects_bad_arguments(gc, ctx):
    """gc_learned_integrate_phases refuses networks built for another grid or no networks."""
    import ctypes as C
    g = gc.Grid.make(W, H, 1)
    small = gc.Corrector(ctx, 32, 64, 1)
    a = gc.DeviceArray.from_numpy(ctx, _ic(gc, 17))
    P = gc.CHParams(1.0, 1.0)
    handles = (C.c_void_p * 1)(small._h.value)
    lo, hi = (C.c_int * 1)(0), (C.c_int * 1)(4)
    st = gc.lib().gc_learned_integrate_phases(handles, lo, hi, 1, C.byref(g), a.gc_dtype,
                                              C.byref(P), a.ptr, DT, 16, 4, 1e-6, None)
    assert st != 0 and b"built for" in gc.lib().gc_last_error()
    st = gc.lib().gc_learned_integrate_phases(handles, lo, hi, 0, C.byref(g), a.gc_dtype,
                                              C.byref(P), a.ptr, DT, 16, 4, 1e-6, None)
    assert st != 0
    small.close()
