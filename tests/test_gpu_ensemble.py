"""GPU parity of the batched 2-D Cahn-Hilliard ensemble (one temporally blocked kernel per
RK4 step, hipGraph replay) against the CPU oracle run member by member."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEMBERS = [
    dict(nx=64, ny=64, kappa=1.0),
    dict(nx=96, ny=80, kappa=0.7, M=1.3, hx=1.1, hy=0.9),
    dict(nx=128, ny=48, kappa=1.6),
    dict(nx=37, ny=29, kappa=0.55),   # ragged: smaller than every tile shape
    dict(nx=8, ny=16, kappa=2.0),     # smaller than the 8-site halo
]


def _ics(gc, members, seed=5):
    return [gc.inputs.ch_initial(m["nx"], m["ny"], 1, 0.05, seed=seed + i,
                                 mean=0.1 * (i % 3 - 1)) for i, m in enumerate(members)]


def _oracle(orc, m, c0, dt, steps):
    og = orc.Grid.make(m["nx"], m["ny"], 1, (m.get("hx", 1.0), m.get("hy", 1.0), 1.0))
    return orc.ch_integrate(og, c0, m.get("M", 1.0), m["kappa"], dt, steps, orc.RK4)


@pytest.mark.parametrize("cfg", range(4))
def test_ensemble_matches_oracle_f64(gc, orc, ctx, cfg):
    dt, steps = 0.01, 30
    ens = gc.Ensemble(ctx, MEMBERS, dt, gc.F64)
    ens.set_config(cfg)
    ics = _ics(gc, MEMBERS)
    for i, c in enumerate(ics):
        ens.set_state(i, c)
    ens.step(steps, use_graph=False)
    for i, m in enumerate(MEMBERS):
        want = _oracle(orc, m, ics[i], dt, steps)
        got = ens.get_state(i)
        assert np.array_equal(got, want), (cfg, i, np.abs(got - want).max())
    ens.close()


def test_graph_replays_match_eager(gc, ctx):
    """Graph replays (20-step graphs) interleaved with eager steps and host syncs give the
    same states as purely eager stepping, every time."""
    dt = 0.01
    ics = _ics(gc, MEMBERS, seed=9)
    a = gc.Ensemble(ctx, MEMBERS, dt, gc.F64)
    b = gc.Ensemble(ctx, MEMBERS, dt, gc.F64)
    for i, c in enumerate(ics):
        a.set_state(i, c)
        b.set_state(i, c)
    for rep, n in enumerate([40, 3, 45, 20, 1, 21]):
        a.step(n, use_graph=True)
        b.step(n, use_graph=False)
        ctx.sync()
        for i in range(len(MEMBERS)):
            assert np.array_equal(a.get_state(i), b.get_state(i)), (rep, i)
    a.close()
    b.close()


def test_tune_preserves_state_and_diagnostics(gc, orc, ctx):
    dt, steps = 0.01, 10
    ens = gc.Ensemble(ctx, MEMBERS, dt, gc.F64)
    ics = _ics(gc, MEMBERS, seed=3)
    for i, c in enumerate(ics):
        ens.set_state(i, c)
    best, ms = ens.tune()
    assert 0 <= best < 4 and all(t > 0 for t in ms)
    for i, c in enumerate(ics):
        assert np.array_equal(ens.get_state(i), c.reshape(1, MEMBERS[i]["ny"], MEMBERS[i]["nx"]))
    ens.step(steps)
    mass, F = ens.diagnostics(gc.PAIRWISE)
    for i, m in enumerate(MEMBERS):
        want = _oracle(orc, m, ics[i], dt, steps)
        og = orc.Grid.make(m["nx"], m["ny"], 1, (m.get("hx", 1.0), m.get("hy", 1.0), 1.0))
        assert mass[i] == orc.integrate(og, want)
        assert F[i] == orc.free_energy(og, want, m["kappa"])
    ens.close()


def test_ensemble_f32_close_to_f64(gc, orc, ctx):
    dt, steps = 0.005, 200
    ens = gc.Ensemble(ctx, MEMBERS[:3], dt, gc.F32)
    ics = _ics(gc, MEMBERS[:3], seed=1)
    for i, c in enumerate(ics):
        ens.set_state(i, c.astype(np.float32))
    ens.step(steps)
    for i, m in enumerate(MEMBERS[:3]):
        want = _oracle(orc, m, ics[i].astype(np.float32).astype(np.float64), dt, steps)
        got = ens.get_state(i).astype(np.float64)
        assert np.abs(got - want).max() < 1e-4 * max(1.0, np.abs(want).max())
    ens.close()
