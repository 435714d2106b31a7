"""Training data, training and evaluation of the learned Cahn-Hilliard corrector.

Configuration (the learned-corrector workload): 2-D Cahn-Hilliard on a 256x256 periodic
grid, dx = 1, M = 1, kappa = 1, dt = 0.01; the network maps c(t) to c(t + 16 dt).
Training trajectories start from U(-0.05, 0.05) noise and are advanced by the FP64 GPU RK4
(the batched ensemble kernel, bit-identical per member to solve::integrate with the
Cahn-Hilliard right-hand side).  Snapshots every 16 steps are kept on the device in FP32;
a training pair is (snapshot m, snapshot m+1) of one trajectory.

    python -m gridcalc_amd.corrector  (or corrector.main()) trains and writes
    data/corrector_<H>x<W>.npy (FP32 parameters) and .json (scale, history, config).
    corrector.main(["phases=1"]) then fine-tunes one copy of those parameters per phase of a
    run (PHASES) and writes data/corrector_<H>x<W>_<phase>.npy / .json.

Phase-specialised inference (integrate_phased): the first WARM macro steps by plain RK4,
then each phase's network on its own range of macro steps.  The network is small; one set
of weights fitted to every stage of the dynamics (fast early coarsening, slow late
coarsening) leaves more one-step error than one set per stage.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import (F32, F64, RK4, CHParams, Context, Corrector, DeviceArray, Ensemble, Grid,
               N_CNN_PARAMS, _check, ch_integrate, inputs, lib)

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# skip_first: macro steps at the start of every trajectory left out of training (the learned
# integrator takes them with plain RK4: the first step from the noise state is the one the
# network predicts worst); sample_pow: training pairs are drawn with probability proportional
# to rms(increment of their macro index)^-sample_pow (0: uniform), so that the slow late
# increments, whose small errors accumulate over most of a run, are not drowned by the large
# early ones in the squared loss.
CONFIG = dict(H=256, W=256, kappa=1.0, M=1.0, dt=0.01, substeps=16, n_traj=2000, n_steps=2000,
              batch_traj=100, minibatch=16, iters_per_batch=9000, lr_max=2e-3, lr_min=2e-5,
              seed=12345, skip_first=0, sample_pow=0.0)


# Phases of a run, as half-open ranges of macro-step indices (macro step m advances the state
# from t = 16 m dt): (name, first, end, fine-tuning stages (batches, iterations per batch,
# learning rate max -> min, seed, pushforward period)).  WARM macro steps at the start are
# plain RK4.  With pushforward period k > 0, every k-th iteration from the second on trains
# on (the network's own one-step prediction from snapshot m, snapshot m + 2) instead of
# (snapshot m, snapshot m + 1): the network learns to advance the states its own errors lead
# to, which cuts the error that accumulates over the slow late phase (the pushforward trick of
# Brandstetter, Worrall and Welling, "Message passing neural PDE solvers", ICLR 2022).
# A seventh stage entry K > 1 trains instead on unrolled chains: from snapshot m, K
# applications of the residual update, the loss on every one of snapshots m + 1 .. m + K
# (weight 1/K each) and its gradient carried back through all K applications
# (gc_cnn_loss_grad_chain), so the weights are fitted to the error after several steps of
# their own, not only to the one-step error.
# The shipped parameters were made by these stages; each phase's json file records, per stage,
# the forward arithmetic of its 32 -> 32 layers (the first stages ran when those layers were
# evaluated in BF16; the later ones, with chains, re-fitted the weights to the FP16 forward
# pass the kernels use now).
WARM = 1
PHASES = (("early", 1, 40, ((12, 2000, 3e-4, 3e-6, 66666, 0), (8, 2000, 1e-4, 1e-6, 31001, 0, 1),
                            (8, 2000, 3e-5, 3e-7, 32001, 0, 3), (8, 2000, 2e-5, 2e-7, 33001, 0, 4))),
          ("late", 40, 125, ((12, 2000, 3e-4, 3e-6, 55555, 0), (12, 2000, 1e-4, 1e-6, 77777, 0),
                             (20, 2000, 1e-4, 1e-6, 133331, 2), (16, 2000, 5e-5, 5e-7, 155551, 2),
                             (8, 2000, 5e-5, 5e-7, 31002, 0, 1), (8, 2000, 3e-5, 3e-7, 31003, 0, 3),
                             (8, 2000, 2e-5, 2e-7, 32002, 0, 4),
                             (10, 2000, 1.5e-5, 1.5e-7, 33002, 0, 4))))


def params_path(H, W, phase=None):
    tag = f"_{phase}" if phase else ""
    return os.path.join(DATA_DIR, f"corrector_{H}x{W}{tag}.npy")


def load_params(H=256, W=256, phase=None):
    """Trained parameters and metadata (of one phase's fine-tuned copy if phase is given), or
    (None, None) if not trained."""
    p = params_path(H, W, phase)
    if not os.path.exists(p):
        return None, None
    meta = json.load(open(p[:-4] + ".json"))
    return np.load(p), meta


def init_params(seed=0):
    """He-style uniform initialisation (last layer scaled down)."""
    rng = np.random.default_rng(seed)
    parts = [rng.uniform(-1, 1, 288) * math.sqrt(6.0 / 9.0), np.zeros(32)]
    for _ in range(4):
        parts += [rng.uniform(-1, 1, 9216) * math.sqrt(6.0 / 288.0), np.zeros(32)]
    parts += [rng.uniform(-1, 1, 288) * math.sqrt(6.0 / 288.0) * 0.1, np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == N_CNN_PARAMS
    return p


def initial_state(H, W, seed):
    """U(-0.05, 0.05) noise."""
    return inputs.symmetric(seed, H * W, 0.05).reshape(1, H, W)


def state_ptr(ens, i):
    p = C.c_void_p()
    _check(lib().gc_ensemble_state_ptr(ens._h, i, C.byref(p)))
    return p


def generate(ctx, n_traj, n_steps, seed, cfg=CONFIG, log=print):
    """Runs n_traj FP64 trajectories for n_steps; returns a device FP32 array of shape
    (n_traj, n_macro + 1, H, W) with the snapshots every `substeps` steps."""
    H, W, sub = cfg["H"], cfg["W"], cfg["substeps"]
    n_macro = n_steps // sub
    member = dict(nx=W, ny=H, kappa=cfg["kappa"], M=cfg["M"])
    data = DeviceArray((n_traj, n_macro + 1, H, W), np.float32)
    ens = Ensemble(ctx, [member] * n_traj, cfg["dt"], F64)
    for i in range(n_traj):
        ens.set_state(i, initial_state(H, W, seed + i))
    t0 = time.perf_counter()
    hw = H * W
    for m in range(n_macro + 1):
        for i in range(n_traj):
            dst = C.c_void_p(data.ptr.value + ((i * (n_macro + 1) + m) * hw) * 4)
            _check(lib().gc_cast(ctx.handle, F64, F32, hw, state_ptr(ens, i), dst))
        if m < n_macro:
            ens.step(sub)
    ctx.sync()
    ens.close()
    log(f"  generated {n_traj} trajectories x {n_steps} steps in {time.perf_counter() - t0:.1f} s")
    return data, n_macro


def snapshot(ctx, data, n_macro, i, m):
    """Host copy of snapshot m of trajectory i."""
    H, W = data.shape[2], data.shape[3]
    out = np.empty((H, W), np.float32)
    src = C.c_void_p(data.ptr.value + ((i * (n_macro + 1) + m) * H * W) * 4)
    _check(lib().gc_memcpy_dtoh(ctx.handle, out.ctypes.data_as(C.c_void_p), src, out.nbytes))
    return out


def increment_rms_by_macro(ctx, data, n_macro, n_traj=4):
    """RMS of (snapshot m+1 - snapshot m) for every macro index m, over the first n_traj
    trajectories."""
    rms = np.zeros(n_macro)
    for m in range(n_macro):
        sq = 0.0
        for i in range(n_traj):
            d = snapshot(ctx, data, n_macro, i, m + 1).astype(np.float64) - snapshot(ctx, data, n_macro, i, m)
            sq += float((d * d).mean())
        rms[m] = math.sqrt(sq / n_traj)
    return rms


def increment_scale(ctx, data, n_macro, n_sample=64, seed=0):
    """RMS of (snapshot m+1 - snapshot m) over a sample of pairs."""
    rng = np.random.default_rng(seed)
    n_traj = data.shape[0]
    sq, cnt = 0.0, 0
    for _ in range(n_sample):
        i, m = int(rng.integers(n_traj)), int(rng.integers(n_macro))
        d = snapshot(ctx, data, n_macro, i, m + 1).astype(np.float64) - snapshot(ctx, data, n_macro, i, m)
        sq += float((d * d).sum())
        cnt += d.size
    return math.sqrt(sq / cnt)


def train(cfg=CONFIG, log=print, out_dir=DATA_DIR):
    cfg = dict(cfg)
    H, W = cfg["H"], cfg["W"]
    n_batches = cfg["n_traj"] // cfg["batch_traj"]
    total_iters = n_batches * cfg["iters_per_batch"]
    rng = np.random.default_rng(cfg["seed"])
    history = []
    t_start = time.perf_counter()
    with Context() as ctx:
        net = Corrector(ctx, H, W, cfg["minibatch"])
        scale = None
        it = 0
        for bi in range(n_batches):
            data, n_macro = generate(ctx, cfg["batch_traj"], cfg["n_steps"],
                                     cfg["seed"] + 1000 * bi, cfg, log)
            if scale is None:
                scale = increment_scale(ctx, data, n_macro)
                net.set_params(init_params(cfg["seed"]), scale)
                log(f"  increment scale {scale:.4e}")
                ms = np.arange(cfg["skip_first"], n_macro)
                w = increment_rms_by_macro(ctx, data, n_macro)[ms] ** -cfg["sample_pow"]
                probs = w / w.sum()
                log(f"  macro indices {ms[0]}..{ms[-1]}, sampling weight max/min {w.max() / w.min():.1f}")
            per_traj = n_macro + 1
            target = C.c_void_p(data.ptr.value + H * W * 4)  # snapshot m+1 of pair index m

            class _Shifted:  # device array view starting one snapshot later
                ptr = target
                gc_dtype = F32

            losses = []
            for _ in range(cfg["iters_per_batch"]):
                tr = rng.integers(0, cfg["batch_traj"], cfg["minibatch"])
                m = rng.choice(ms, cfg["minibatch"], p=probs)
                idx = (tr * per_traj + m).astype(np.int32)
                frac = it / max(1, total_iters - 1)
                lr = cfg["lr_min"] + 0.5 * (cfg["lr_max"] - cfg["lr_min"]) * (1 + math.cos(math.pi * frac))
                loss = net.loss_grad(data, _Shifted, idx=idx)
                net.adam(lr)
                losses.append(loss)
                it += 1
            history.append(float(np.mean(losses)))
            log(f"  batch {bi + 1}/{n_batches}: mean loss {history[-1]:.4e} "
                f"(last {losses[-1]:.4e}), {time.perf_counter() - t_start:.0f} s")
            data.free()
        params = net.params()
        net.close()
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"corrector_{H}x{W}.npy")
    np.save(path, params)
    meta = dict(scale=scale, history=history, config=cfg, seconds=time.perf_counter() - t_start)
    json.dump(meta, open(path[:-4] + ".json", "w"), indent=1)
    log(f"  wrote {path}")
    return params, meta


def finetune(params, scale, m_lo, m_hi, stages, cfg=CONFIG, log=print):
    """Continues training `params` on pairs whose first snapshot is macro step m_lo..m_hi-1
    (uniformly), through `stages` = ((batches, iterations per batch, lr max, lr min, seed,
    pushforward period[, chain length K]), ...), each a cosine schedule over fresh
    trajectories (see PHASES for the pushforward and chain iterations).  Returns (params,
    loss history)."""
    H, W, mb = cfg["H"], cfg["W"], cfg["minibatch"]
    hw = H * W
    history = []
    t_start = time.perf_counter()
    p = np.asarray(params, np.float32)
    with Context() as ctx:
        net = Corrector(ctx, H, W, mb)
        xs, ys, ts = (DeviceArray((mb, H, W), np.float32) for _ in range(3))

        def gather(dst, data, rows):  # snapshots `rows` of data into the images of dst
            for b, r in enumerate(rows):
                _check(lib().gc_memcpy_dtod(ctx.handle, C.c_void_p(dst.ptr.value + b * hw * 4),
                                            C.c_void_p(data.ptr.value + int(r) * hw * 4), hw * 4))

        for n_batches, iters, lr_max, lr_min, seed, pf, *chain in stages:
            K = chain[0] if chain else 1
            net.set_params(p, scale)  # fresh Adam state per stage
            rng = np.random.default_rng(seed)
            total = n_batches * iters
            it = 0
            for bi in range(n_batches):
                data, n_macro = generate(ctx, cfg["batch_traj"], cfg["n_steps"], seed + 1000 * bi,
                                         cfg, lambda *a: None)
                per_traj = n_macro + 1

                class _Shifted:  # device array view starting one snapshot later
                    ptr = C.c_void_p(data.ptr.value + H * W * 4)
                    gc_dtype = F32

                losses = []
                for _ in range(iters):
                    frac = it / max(1, total - 1)
                    lr = lr_min + 0.5 * (lr_max - lr_min) * (1 + math.cos(math.pi * frac))
                    tr = rng.integers(0, cfg["batch_traj"], mb)
                    if K > 1:  # unrolled chain: snapshot m through K applications
                        m = rng.integers(m_lo, min(m_hi, n_macro - K + 1), mb)
                        losses.append(net.chain_loss_grad(data, tr * per_traj + m, K))
                    elif pf and it % pf:  # pushforward: net(snapshot m) -> snapshot m + 2
                        m = rng.integers(m_lo, min(m_hi, n_macro - 1), mb)
                        gather(xs, data, tr * per_traj + m)
                        gather(ts, data, tr * per_traj + m + 2)
                        net.predict(xs, ys, mb)
                        losses.append(net.loss_grad(ys, ts, B=mb))
                    else:
                        m = rng.integers(m_lo, min(m_hi, n_macro), mb)
                        idx = (tr * per_traj + m).astype(np.int32)
                        losses.append(net.loss_grad(data, _Shifted, idx=idx))
                    net.adam(lr)
                    it += 1
                data.free()
                history.append(float(np.mean(losses)))
                log(f"  [{m_lo}, {m_hi}) batch {bi + 1}/{n_batches}: mean loss {history[-1]:.4e}, "
                    f"{time.perf_counter() - t_start:.0f} s")
            p = net.params()
        net.close()
        for a in (xs, ys, ts):
            a.free()
    return p, history


def train_phases(cfg=CONFIG, log=print, out_dir=DATA_DIR, only=None):
    """Fine-tunes the trained parameters once per phase (PHASES, or those named in `only`)
    and writes them.  A phase whose shipped parameters were made by a prefix of its declared
    stages continues from them with the remaining stages (each stage starts from the previous
    stage's parameters with a fresh optimiser state and its own seed, so the result is that of
    running every stage)."""
    H, W = cfg["H"], cfg["W"]
    base, meta = load_params(H, W)
    if base is None:
        raise RuntimeError("train the corrector first")
    for name, m_lo, m_hi, stages in PHASES:
        if only is not None and name not in only:
            continue
        t0 = time.perf_counter()
        start, done, hist0, sec0, arith0 = base, 0, [], 0.0, []
        prev, pmeta = load_params(H, W, name)
        if prev is not None and pmeta.get("config") == cfg:
            k = len(pmeta["stages"])
            if 0 < k <= len(stages) and pmeta["stages"] == [list(st) for st in stages[:k]]:
                if k == len(stages):
                    log(f"  {name}: every stage done")
                    continue
                start, done, hist0 = prev, k, list(pmeta["history"])
                sec0 = float(pmeta.get("seconds", 0.0))
                arith0 = list(pmeta.get("forward_arithmetic", ["bf16"] * k))
                log(f"  {name}: continuing after stage {k} of {len(stages)}")
        p, history = finetune(start, meta["scale"], m_lo, m_hi, stages[done:], cfg, log)
        history = hist0 + history
        path = os.path.join(out_dir, f"corrector_{H}x{W}_{name}.npy")
        np.save(path, p)
        json.dump(dict(scale=meta["scale"], history=history, macro_steps=[m_lo, m_hi],
                       stages=[list(st) for st in stages], config=cfg,
                       seconds=sec0 + time.perf_counter() - t0,
                       forward_arithmetic=arith0 + ["fp16"] * (len(stages) - done)),
                  open(path[:-4] + ".json", "w"), indent=1)
        log(f"  wrote {path}")


def phase_nets(ctx, H=256, W=256):
    """One Corrector per phase with its fine-tuned parameters: [(net, first, end), ...], or
    None if the phase parameters are not there."""
    nets = []
    for name, m_lo, m_hi, _ in PHASES:
        p, meta = load_params(H, W, name)
        if p is None:
            for n, _, _ in nets:
                n.close()
            return None
        net = Corrector(ctx, H, W, 1)
        net.set_params(p, meta["scale"])
        nets.append((net, m_lo, m_hi))
    return nets


def integrate_phased(ctx, g, psi, nets, M, kappa, dt, substeps, n_macro, mass_tol=1e-6,
                     warm=WARM):
    """n_macro macro steps of substeps RK4 steps each: the first `warm` by plain RK4, then
    every macro step m by the network of the phase whose range holds m (plain RK4 where no
    phase does), each with the mass / free-energy fallback, in one gc_learned_integrate_phases
    call (the proposal pipeline runs on across phase boundaries; mass drift is measured from
    the initial state).  Returns the fallback count."""
    segs = [(net, max(lo, warm), hi) for net, lo, hi in nets if hi > max(lo, warm)]
    if not segs:
        ch_integrate(ctx, g, psi, M, kappa, dt, n_macro * substeps, RK4)
        return 0
    k = len(segs)
    handles = (C.c_void_p * k)(*[net._h.value for net, _, _ in segs])
    first = (C.c_int * k)(*[lo for _, lo, _ in segs])
    end = (C.c_int * k)(*[hi for _, _, hi in segs])
    P = CHParams(M, kappa)
    nf = C.c_int()
    _check(lib().gc_learned_integrate_phases(handles, first, end, k, C.byref(g), psi.gc_dtype,
                                             C.byref(P), psi.ptr, dt, substeps, n_macro, mass_tol,
                                             C.byref(nf)))
    return nf.value


def one_step_errors(net, cfg=CONFIG, n_traj=8, seed=777001, log=print):
    """Relative L2 error of one macro step (substeps RK4 steps) on held-out trajectories:
    returns (network, identity) where identity is the error of c(t+16 dt) ~ c(t), i.e. of the
    method without the network.  Errors are over all macro steps of n_traj trajectories."""
    ctx = net.ctx
    data, n_macro = generate(ctx, n_traj, cfg["n_steps"], seed, cfg, log)
    H, W = cfg["H"], cfg["W"]
    hw = H * W
    cur = DeviceArray((1, H, W), np.float32)
    nxt = DeviceArray((1, H, W), np.float32)
    e_net = e_id = norm = 0.0
    for i in range(n_traj):
        for m in range(n_macro):
            src = C.c_void_p(data.ptr.value + ((i * (n_macro + 1) + m) * hw) * 4)
            _check(lib().gc_memcpy_dtod(ctx.handle, cur.ptr, src, hw * 4))
            net.predict(cur, nxt, 1)
            a = snapshot(ctx, data, n_macro, i, m).astype(np.float64)
            b = snapshot(ctx, data, n_macro, i, m + 1).astype(np.float64)
            p = nxt.numpy(ctx).astype(np.float64).reshape(H, W)
            e_net += float(((p - b) ** 2).sum())
            e_id += float(((a - b) ** 2).sum())
            norm += float((b * b).sum())
    cur.free()
    nxt.free()
    data.free()
    return math.sqrt(e_net / norm), math.sqrt(e_id / norm)


def integration_error(net, cfg=CONFIG, seed=991001, n_steps=None, mass_tol=1e-6, warm=None):
    """Runs the learned integrator and the plain FP64 RK4 from the same U(-0.05, 0.05) state
    over n_steps steps (default cfg n_steps, i.e. T = n_steps * dt); returns a dict with the
    relative L2 and max errors of the learned result against the RK4 result, the fallback
    count and both wall times (device synchronised).  The first `warm` macro steps (default
    cfg skip_first) are plain RK4 in the learned run too."""
    ctx = net.ctx
    warm = cfg.get("skip_first", 0) if warm is None else warm
    H, W, dt, sub = cfg["H"], cfg["W"], cfg["dt"], cfg["substeps"]
    n_steps = cfg["n_steps"] if n_steps is None else n_steps
    assert n_steps % sub == 0
    g = Grid.make(W, H, 1)
    c0 = initial_state(H, W, seed).astype(np.float64)
    ref = DeviceArray.from_numpy(ctx, c0)
    lrn = DeviceArray.from_numpy(ctx, c0)
    ctx.sync()
    t0 = time.perf_counter()
    ch_integrate(ctx, g, ref, cfg["M"], cfg["kappa"], dt, n_steps, RK4)
    ctx.sync()
    t_rk4 = time.perf_counter() - t0
    t0 = time.perf_counter()
    if warm:
        ch_integrate(ctx, g, lrn, cfg["M"], cfg["kappa"], dt, warm * sub, RK4)
    nf = net.integrate(g, lrn, cfg["M"], cfg["kappa"], dt, sub, n_steps // sub - warm, mass_tol)
    ctx.sync()
    t_lrn = time.perf_counter() - t0
    r, l = ref.numpy(ctx), lrn.numpy(ctx)
    ref.free()
    lrn.free()
    return dict(rel_l2=float(np.linalg.norm(l - r) / np.linalg.norm(r)),
                max_abs=float(np.abs(l - r).max()), fallbacks=nf, n_macro=n_steps // sub,
                seconds_rk4=t_rk4, seconds_learned=t_lrn,
                mass_drift=float(abs(l.mean() - c0.mean())))


def main(argv=None):
    """key=value arguments override CONFIG; out=<dir> changes the output directory;
    phases=1 fine-tunes the per-phase copies of the trained parameters instead of training
    (only=a,b: just those phases)."""
    argv = sys.argv[1:] if argv is None else argv
    cfg = dict(CONFIG)
    out_dir = DATA_DIR
    phases = False
    only = None
    for a in argv:
        k, v = a.split("=", 1)
        if k == "out":
            out_dir = v
        elif k == "phases":
            phases = v == "1"
        elif k == "only":
            only = v.split(",")
        else:
            cfg[k] = type(CONFIG[k])(v)
    if phases:
        train_phases(cfg, out_dir=out_dir, only=only)
    else:
        train(cfg, out_dir=out_dir)


if __name__ == "__main__":
    main()
