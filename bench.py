#!/usr/bin/env python3
"""Benchmark: fused RK4 Cahn-Hilliard site-updates/s on one MI355X.

Default workload (the RK4 configuration of the headline metric): 3-D Cahn-Hilliard on a
256^3 periodic grid, dx=1, M=1, kappa=0.5, dt=0.005, c0 = 0.1*U(-1,1); stencils and RK4
stages in FP64 (the reference's precision) for the headline line, and the FP32 form
(FP32 stencils and stages) in extra.fp32; total mass and free energy are reduced in FP64
(compensated) every 500 steps and at the end, as the reference driver prints diagnostics
per snapshot chunk.  One "step" is one RK4 step of the whole grid: two two-stage kernels.
Inputs are resident in HBM before timing; every kernel of the timed region has run once
before it.

Also reported: the FP32 result's L2/Linf error against the FP64 GPU run of the same steps,
the roofline position of the fused step (kernel time from HIP events over 200 steps), and
the single-core CPU oracle (a C restatement of the reference path, FP64, gcc -O3) timed on
a bounded prefix of the same workload, with the host CPU model and the oracle's flags.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu] [--no-fp32]
  python bench.py --workload sweep [--n 512]      # operator GB/s (3-D operator sweep)
  python bench.py --workload ensemble             # 64 uneven 2-D members, graph-replayed
  python bench.py --workload learned [--n-ic 4]   # learned corrector vs plain RK4 to T=20
  python bench.py --workload tune [--n 256]       # autotuner: ms/step of every tile shape
  python bench.py --workload diffuse [--n 256]    # constant- and heterogeneous-D diffusion RK4
  python bench.py --workload reduce [--n 256]     # integrate() and the fused free energy
  python bench.py --workload rhs                  # RK4 right-hand sides: host lambda, expr, fused
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as entry  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM words per site per fused RK4 step: a low-storage RK4 reads and writes
# 16 words per site (4 stages x (stencil input + stage state)).  The bit-exact stage
# pipeline here moves 17 (it keeps k1 and k2+k3 apart as the reference combines them).
WORDS_PER_SITE_STEP = 16


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=None)
    p.add_argument("--nz", type=int, default=None, help="tune workload: Nz (default n; 1: 2-D)")
    p.add_argument("--workload", default="ch3d", choices=["ch3d", "sweep", "copy", "ensemble",
                                                          "learned", "tune", "diffuse", "reduce",
                                                          "rhs"])
    p.add_argument("--n-ic", type=int, default=4)
    p.add_argument("--no-learned", action="store_true")
    p.add_argument("--diag-every", type=int, default=500)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--no-tune", action="store_true")
    p.add_argument("--ch-config", type=int, default=None,
                   help="force this tile candidate (gc_ch_set_config) in both precisions")
    p.add_argument("--cpu-steps", type=int, default=4)
    return p.parse_args()


def profiled_traffic(name, key="traffic_bytes_per_step"):
    """Bytes per fused step measured by rocprofv3 FETCH_SIZE/WRITE_SIZE (scaled by the
    device-copy calibration) for this workload, as recorded in profiles/<name>."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + " ="):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def run_timed(gc, ctx, g, d, steps, diag_every, kappa, dt):
    diags = []
    ctx.sync()
    ctx.timer_start()
    done = 0
    while done < steps:
        chunk = min(diag_every, steps - done)
        gc.ch_integrate(ctx, g, d, 1.0, kappa, dt, chunk, gc.RK4)
        done += chunk
        mass = gc.integrate(ctx, g, d, gc.KAHAN)
        F = gc.ch_free_energy(ctx, g, d, kappa, gc.KAHAN)
        diags.append((done, mass, F))
    ms = ctx.timer_stop()
    return ms, diags


def warm(gc, ctx, g, src, steps, kappa, dt):
    """Untimed first launches of every kernel the timed region runs (the stage pipeline and
    both compensated reductions) on a device copy of `src`; the last reduction returns just
    before the timed region starts, so the GPU does not idle in between."""
    w = gc.DeviceArray(src.shape, src.dtype)
    w.copy_from(ctx, src)
    gc.ch_integrate(ctx, g, w, 1.0, kappa, dt, max(steps, 2), gc.RK4)
    gc.integrate(ctx, g, w, gc.KAHAN)
    gc.ch_free_energy(ctx, g, w, kappa, gc.KAHAN)
    w.free()


def step_ms(gc, ctx, g, src, steps, kappa, dt):
    """ms per fused RK4 step: HIP events on the context's stream around `steps` (even) steps
    and nothing else (2 two-stage or 4 one-stage kernel launches per step), on a device copy
    of `src` (no host upload, so the GPU does not idle before the window)."""
    d = gc.DeviceArray(src.shape, src.dtype)
    d.copy_from(ctx, src)
    gc.ch_integrate(ctx, g, d, 1.0, kappa, dt, 2, gc.RK4)
    ctx.sync()
    ctx.timer_start()
    gc.ch_integrate(ctx, g, d, 1.0, kappa, dt, steps, gc.RK4)
    ms = ctx.timer_stop()
    d.free()
    return ms / steps


def cpu_record(orc, extra_sample):
    """The oracle's build and this host's CPU, recorded beside every CPU timing."""
    info = orc.host_info()
    return {"cores": 1, "kind": "port", "cpu_model": info["cpu_model"],
            "host_cores_available": _HOST_CORES if _HOST_CORES is not None else info["nproc"],
            "host_logical_cpus": os.cpu_count(), "build": info["oracle_build"],
            "sample": extra_sample}


KERNEL_TIMING_STEPS = 200


def ch3d(a):
    gc = entry.load_package()
    n = a.n or 256
    kappa, dt = 0.5, 0.005
    g = gc.Grid.make(n, n, n, 1.0)
    c0 = gc.inputs.ch_initial(n, n, n, 0.1, seed=3)
    sites = n ** 3
    extra = {}
    legs = {}
    with gc.Context() as ctx:
        # inputs first: the GPU idles during a host upload, and its clocks take tens of
        # milliseconds of load to settle again, so uploads go before the tuner and warm-up
        legs_wanted = [("f64", gc.F64, np.float64, 8)]
        if not a.no_fp32:
            legs_wanted.append(("f32", gc.F32, np.float32, 4))
        arrays = {name: gc.DeviceArray.from_numpy(ctx, c0.astype(np_dtype))
                  for name, _, np_dtype, _ in legs_wanted}
        tiles, kinds = {}, {}
        for dtype, name in ((gc.F64, "f64"), (gc.F32, "f32")):
            if a.ch_config is not None:
                gc.ch_set_config(g, dtype, a.ch_config)
            elif not a.no_tune:
                gc.ch_tune(ctx, g, dtype)
            cfg = gc.ch_get_config(g, dtype)
            tile = list(gc.ch_config_shape(g, cfg))
            kinds[name] = gc.ch_config_kind(g, cfg)
            if name == "f64" and kinds[name] in (1, 2):
                # the shape is the FP32 one; wider two-stage tiles run 64 x 16 in FP64
                tile = [min(tile[0], 64), min(tile[1], 16), tile[2]]
            tiles[name] = tile
        extra["tile_ox_oy_zc"] = tiles
        extra["two_stage_kernels"] = {k: ["none", "plane rings", "register queues"][v]
                                      for k, v in kinds.items()}
        for name, dtype, np_dtype, w in legs_wanted:
            d = arrays[name]
            # the kernel timing goes first: right after the tuner's burst of launches the
            # GPU's clocks sag for some tens of milliseconds (timed steps ran up to 12 % slower
            # there than in a steady run), and its 200 steps carry the GPU past that
            sms = step_ms(gc, ctx, g, d, KERNEL_TIMING_STEPS, kappa, dt)
            warm(gc, ctx, g, d, a.warmup, kappa, dt)
            ms, diags = run_timed(gc, ctx, g, d, a.steps, a.diag_every, kappa, dt)
            bytes_step = WORDS_PER_SITE_STEP * w * sites
            legs[name] = {"ms": ms, "diags": diags, "step_kernels_ms": sms,
                          "bytes_step": bytes_step,
                          "achieved": bytes_step / (sms / 1e3) / 1e9}
            arrays[name] = d
        if "f32" in legs:
            r64 = arrays["f64"].numpy(ctx)
            r32 = arrays["f32"].numpy(ctx).astype(np.float64)
            L = legs["f32"]
            extra["fp32"] = {
                "site_updates_per_s": sites * a.steps / (L["ms"] / 1e3),
                "ms_per_step": L["ms"] / a.steps,
                "step_kernels_ms": L["step_kernels_ms"],
                "hbm_GBs_alg": L["achieved"], "hbm_frac": L["achieved"] / HBM_PEAK_GBS,
                "kernel": _kernel_name(kinds["f32"], "float", tiles["f32"]),
                "traffic": profiled_traffic(PROFILE_CH3D, "traffic_bytes_per_step_f32"),
                "diagnostics": [{"step": s, "mass": m, "F": f} for s, m, f in L["diags"]],
                "vs_fp64": {
                    "l2_rel": float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64)),
                    "linf_rel": float(np.abs(r32 - r64).max() / np.abs(r64).max()),
                    "mass_fp32": L["diags"][-1][1], "mass_fp64": legs["f64"]["diags"][-1][1],
                    "F_fp32": L["diags"][-1][2], "F_fp64": legs["f64"]["diags"][-1][2]}}
        for d in arrays.values():
            d.free()
    H = legs["f64"]
    traffic64 = profiled_traffic(PROFILE_CH3D, "traffic_bytes_per_step_f64")
    extra["diagnostics_fp64"] = [{"step": s, "mass": m, "F": f} for s, m, f in H["diags"]]

    if not a.no_learned:
        extra["learned"] = learned(a)

    cpu = None
    if not a.no_cpu:
        orc = entry.load_oracle()
        orc.lib()
        _pin_one_core()
        og = orc.Grid.make(n, n, n, 1.0)
        t0 = time.perf_counter()
        orc.ch_integrate(og, c0, 1.0, kappa, dt, a.cpu_steps, orc.RK4)
        el = time.perf_counter() - t0
        cpu = {"value": sites * a.cpu_steps / el, "unit": "site-updates/s",
               **cpu_record(orc, f"CPU oracle (C restatement of the reference path, FP64, "
                                 f"1 thread pinned to one core): RK4 on the same {n}^3 input; the "
                                 f"rate of the first {a.cpu_steps} steps ({el:.1f} s), not a full "
                                 f"{a.steps}-step run")}

    return {
        "metric": "RK4 site-updates/s",
        "value": sites * a.steps / (H["ms"] / 1e3),
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": H["ms"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Cahn-Hilliard fused RK4, {n}^3, dx=1, M=1, kappa=0.5, "
                               f"dt=0.005, c0=0.1*U(-1,1) seed 3; FP64 stages (the FP32 form in "
                               f"extra.fp32), FP64 compensated mass/F every {a.diag_every} steps "
                               f"and at the end",
                   "grid": [n, n, n], "dt": dt, "kappa": kappa, "M": 1.0},
        "roofline": {"bound": "hbm", "achieved": H["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": H["achieved"] / HBM_PEAK_GBS,
                     "traffic": traffic64,
                     "kernel": _kernel_name(kinds["f64"], "double", tiles["f64"]),
                     "step_kernels_ms": H["step_kernels_ms"],
                     "kernel_timing_steps": KERNEL_TIMING_STEPS,
                     "alg_bytes_per_step": H["bytes_step"],
                     # the kernels move 9 of the 16 algorithmic words per site and step: the
                     # HBM rate they run at, from the profiled traffic and this run's kernel time
                     "traffic_GBs": (traffic64 / (H["step_kernels_ms"] / 1e3) / 1e9
                                     if traffic64 else None)},
        "cpu_baseline": cpu,
        "extra": extra,
    }


PROFILE_CH3D = "ch3d_rk4_pairq3d_256.md"


def _kernel_name(kind, T, tile):
    if kind == 0:
        return f"k_ch_stage3d<{T},...> x4 (one RK4 step = 4 stage launches)"
    kname = "k_ch_pairq3d" if kind == 2 else "k_ch_pair3d"
    return (f"{kname}<{T},{tile[0]},{tile[1]},{tile[2]},...> x2 (one RK4 step = S1+S2 and "
            "S3+S4 launches)")


def sweep(a):
    """3-D operator sweep: GB/s per operator from algorithmic bytes, FP64 and FP32, and the
    FP32 result's relative L2 and max error against the FP64 result of the same operator
    (each precision on its own-precision copy of the inputs; the FP64 results are the
    reference's bits, which the parity tests pin)."""
    gc = entry.load_package()
    n = a.n or 512
    h = 1.0 / n
    g = gc.Grid.make(n, n, n, h)
    u = gc.inputs.sweep_field(n, h, seed=7)
    rows = []
    words = {"laplacian": 2, "biharmonic": 2, "gradient": 4, "divergence": 4, "cell_laplacian": 3,
             "gradient_vec": 12, "divergence_mat": 12}
    ops = [("laplacian", 2), ("laplacian", 4), ("gradient", 2), ("gradient", 4),
           ("divergence", 2), ("divergence", 4), ("biharmonic", 2), ("biharmonic", 4),
           ("cell_laplacian", 2), ("gradient_vec", 2), ("gradient_vec", 4),
           ("divergence_mat", 2), ("divergence_mat", 4)]
    # three independent components for the vector field; D ~ lognormal(0, 1)
    vv = np.empty((n, n, n, 3), np.float64)
    for c in range(3):
        vv[..., c] = gc.inputs.sweep_field(n, h, seed=11 + c)
    D = gc.inputs.lognormal(5, n ** 3).reshape(n, n, n)
    with gc.Context() as ctx:
        dev = {}
        for dtype in (np.float64, np.float32):
            d_u = gc.DeviceArray.from_numpy(ctx, u.astype(dtype))
            d_v = gc.DeviceArray.from_numpy(ctx, vv.astype(dtype))
            d_D = gc.DeviceArray.from_numpy(ctx, D.astype(dtype))
            outs = {"laplacian": gc.DeviceArray((n, n, n), dtype),
                    "gradient": gc.DeviceArray((n, n, n, 3), dtype),
                    "mat": gc.DeviceArray((n, n, n, 9), dtype)}
            # the tensor input of divergence_mat: the order-2 gradient of the vector field (in
            # FP32 from the FP32 vector field, so each precision runs on its own inputs)
            gc.gradient_vec(ctx, g, d_v, 2, out=outs["mat"])
            d_m = gc.DeviceArray((n, n, n, 9), dtype)
            d_m.copy_from(ctx, outs["mat"])
            dev[dtype] = (d_u, d_v, d_D, d_m, outs)
        for name, order in ops:
            out64 = None
            for dtype, w in ((np.float64, 8), (np.float32, 4)):
                d_u, d_v, d_D, d_m, outs = dev[dtype]
                key = {"gradient_vec": "mat", "divergence_mat": "gradient",
                       "gradient": "gradient"}.get(name, "laplacian")

                def run():
                    if name == "gradient_vec":
                        gc.gradient_vec(ctx, g, d_v, order, out=outs["mat"])
                    elif name == "divergence_mat":
                        gc.divergence_mat(ctx, g, d_m, order, out=outs["gradient"])
                    elif name == "laplacian":
                        gc.laplacian(ctx, g, d_u, order, out=outs["laplacian"])
                    elif name == "biharmonic":
                        gc.biharmonic(ctx, g, d_u, order, out=outs["laplacian"])
                    elif name == "gradient":
                        gc.gradient(ctx, g, d_u, order, out=outs["gradient"])
                    elif name == "divergence":
                        gc.divergence(ctx, g, d_v, order, out=outs["laplacian"])
                    else:
                        gc.cell_laplacian(ctx, g, d_u, d_D, out=outs["laplacian"])
                run()
                ctx.sync()
                reps = 10
                ctx.timer_start()
                for _ in range(reps):
                    run()
                ms = ctx.timer_stop() / reps
                gbs = words[name] * w * n ** 3 / (ms / 1e3) / 1e9
                rows.append({"op": name, "order": order, "dtype": "f64" if w == 8 else "f32",
                             "ms": ms, "GBs_alg": gbs, "hbm_frac": gbs / HBM_PEAK_GBS})
                res = outs[key].numpy(ctx)
                if dtype is np.float64:
                    out64 = res
                else:  # FP32 against FP64 (each on its own-precision input), in chunks
                    sq = ref = 0.0
                    mx = mref = 0.0
                    a64, a32 = out64.reshape(-1), res.reshape(-1)
                    for lo in range(0, a64.size, 1 << 26):
                        r64 = a64[lo:lo + (1 << 26)]
                        dlt = a32[lo:lo + (1 << 26)].astype(np.float64) - r64
                        sq += float(np.dot(dlt, dlt))
                        ref += float(np.dot(r64, r64))
                        mx = max(mx, float(np.abs(dlt).max()))
                        mref = max(mref, float(np.abs(r64).max()))
                    rows[-1]["vs_f64"] = {"l2_rel": math.sqrt(sq / ref), "linf_rel": mx / mref}
                    out64 = None
                del res
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        for d_u, d_v, d_D, d_m, outs in dev.values():
            for x in (d_u, d_v, d_D, d_m, *outs.values()):
                x.free()
    cpu = None
    if not a.no_cpu:  # the oracle (FP64, 1 thread) once per operator on the same input
        orc = entry.load_oracle()
        orc.lib()
        _pin_one_core()
        og = orc.Grid.make(n, n, n, h)
        cpu = {}
        mat = orc.gradient_vec(og, vv, 2)
        for name, order in ops:
            t0 = time.perf_counter()
            if name == "gradient_vec":
                orc.gradient_vec(og, vv, order)
            elif name == "divergence_mat":
                orc.divergence_mat(og, mat, order)
            elif name == "cell_laplacian":
                orc.cell_laplacian(og, u, D)
            elif name == "divergence":
                orc.divergence(og, vv, order)
            else:
                getattr(orc, name)(og, u, order)
            el = time.perf_counter() - t0
            gpu = [r for r in rows if r["op"] == name and r["order"] == order and r["dtype"] == "f64"][0]
            gpu["cpu_ms"] = el * 1e3
            gpu["speedup_vs_cpu"] = el * 1e3 / gpu["ms"]
            cpu[f"{name}{order}"] = el * 1e3
        cpu = {"ms_per_op": cpu,
               **cpu_record(orc, f"CPU oracle (C restatement, FP64, 1 thread), each operator "
                                 f"once on the same {n}^3 input")}
    return {"metric": "operator GB/s (algorithmic bytes)", "n": n, "rows": rows, "cpu_baseline": cpu}


_HOST_CORES = None  # cores this process could use before it pinned itself to one


def _pin_one_core():
    global _HOST_CORES
    try:
        avail = sorted(os.sched_getaffinity(0))
        if _HOST_CORES is None:
            _HOST_CORES = len(avail)
        os.sched_setaffinity(0, {avail[0]})
    except (AttributeError, OSError):
        pass


def ensemble_members(gc, n=64, seed=4):
    """64 members, N in {64,128,256,512,1024} (square), kappa ~ U(0.5, 2), mean ~ U(-0.3, 0.3),
    h = 1, M = 1, drawn from one seed."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 256, 512, 1024], size=n)
    kappas = rng.uniform(0.5, 2.0, size=n)
    means = rng.uniform(-0.3, 0.3, size=n)
    members = [dict(nx=int(s), ny=int(s), kappa=float(k)) for s, k in zip(sizes, kappas)]
    ics = [gc.inputs.ch_initial(m["nx"], m["ny"], 1, 0.05, seed=1000 + i, mean=float(means[i]))
           for i, m in enumerate(members)]
    return members, ics


def ensemble(a):
    """Batched ensemble of 64 uneven 2-D Cahn-Hilliard instances, FP32, dt = 0.005: one
    temporally blocked kernel per RK4 step replayed from a hipGraph, tiles autotuned."""
    gc = entry.load_package()
    members, ics = ensemble_members(gc)
    dt = 0.005
    steps = a.steps
    with gc.Context() as ctx:
        ens = gc.Ensemble(ctx, members, dt, gc.F32)
        for i, c in enumerate(ics):
            ens.set_state(i, c.astype(np.float32))
        mass0, F0 = ens.diagnostics(gc.KAHAN)
        best, tune_ms = ens.tune()
        tile = ens.config_shape(best)
        ens.step(a.warmup)
        ctx.sync()
        ctx.timer_start()
        ens.step(steps, use_graph=True)
        ms = ctx.timer_stop()
        mass, F = ens.diagnostics(gc.KAHAN)
        sites = ens.sites()
        ens.close()
    value = sites * steps / (ms / 1e3)
    cpu = None
    if not a.no_cpu:  # the oracle's RK4 (FP64, 1 thread), 2 steps of every member
        orc = entry.load_oracle()
        orc.lib()
        _pin_one_core()
        t0 = time.perf_counter()
        for m, c in zip(members, ics):
            og = orc.Grid.make(m["nx"], m["ny"], 1, 1.0)
            orc.ch_integrate(og, c, 1.0, m["kappa"], dt, 2, orc.RK4)
        el = time.perf_counter() - t0
        cpu = {"value": sites * 2 / el, "unit": "site-updates/s",
               **cpu_record(orc, f"CPU oracle RK4 (FP64, 1 thread): the rate of the first 2 "
                                 f"steps of all 64 members ({el:.1f} s)")}
    return {"metric": "ensemble RK4 site-updates/s", "value": value, "unit": "site-updates/s",
            "cpu_baseline": cpu,
            "dtype": "f32", "steps": steps, "warmup": a.warmup, "ms_per_step": ms / steps,
            "sites": sites, "members": len(members), "tile_ox_oy": tile,
            "tune_ms_per_step": tune_ms,
            "state_GBs": 2 * 4 * sites / (ms / steps / 1e3) / 1e9,
            # the C4 algorithmic figure: 16 words (64 B in FP32) per site and RK4 step
            "roofline": {"bound": "hbm", "achieved": WORDS_PER_SITE_STEP * 4 * sites / (ms / steps / 1e3) / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": WORDS_PER_SITE_STEP * 4 * sites / (ms / steps / 1e3) / 1e9 / HBM_PEAK_GBS,
                         "note": "the kernel moves 2 words per site per step (temporal blocking)"},
            "max_rel_mass_drift": max(abs(m1 - m0) / max(abs(m0), 1e-300) for m0, m1 in zip(mass0, mass)),
            "F_increases": sum(1 for f0, f1 in zip(F0, F) if f1 > f0)}


def learned(a):
    """Learned corrector against plain RK4 on 2-D Cahn-Hilliard, 256^2, dx=1, M=1, kappa=1,
    dt=0.01, from U(-0.05, 0.05) to T = 20 (2000 RK4 steps = 125 macro steps of 16).  The
    learned path advances the FP64 state with the trained network and falls back to plain
    RK4 for a macro step whose proposal drifts the mass by more than 1e-6 or raises the free
    energy.  Error: relative L2 against the plain FP64 GPU RK4 result (bit-identical to the
    CPU oracle's RK4; the parity tests pin that).  The plain path's tile shape is autotuned
    first.  Times are wall-clock, device-synchronised, after one untimed run of each path;
    n_ic initial states, median time, max error.  Variants: the first macro step by plain RK4
    (warm1), and the phase-specialised networks (corrector.integrate_phased: first macro step
    plain, then one fine-tuned copy of the network per range of macro steps)."""
    gc = entry.load_package()
    from gridcalc_amd import corrector as cr  # noqa: E402
    params, meta = cr.load_params(256, 256)
    if params is None:
        return {"metric": "learned vs plain time-to-T", "error": "no trained parameters"}
    cfg = dict(cr.CONFIG)
    H, W, dt, sub = cfg["H"], cfg["W"], cfg["dt"], cfg["substeps"]
    n_steps = 2000
    g = gc.Grid.make(W, H, 1)
    seeds = [991001 + i for i in range(a.n_ic)]
    res = {"plain_fp64": [], "plain_fp32": [], "plain_fp64_graph": [], "learned_fp64": [],
           "learned_fp64_warm1": []}
    errs, errs32, errs_w, errs_p, fallbacks, fallbacks_p = [], [], [], [], [], []
    with gc.Context() as ctx:
        tiles = {}
        for dtype, name in ((gc.F64, "f64"), (gc.F32, "f32")):  # the plain path at its best
            best, _ = gc.ch_tune(ctx, g, dtype)
            tiles[name] = gc.ch_config_shape(g, best)[:2]
        net = gc.Corrector(ctx, H, W, 1)
        net.set_params(params, meta["scale"])
        # phase-specialised copies of the network (first macro step plain, then one fine-tuned
        # network per range of macro steps), if shipped
        phased = cr.phase_nets(ctx, H, W)
        if phased is not None:
            res["learned_fp64_phased"] = []

        def timed(fn):
            ctx.sync()
            t0 = time.perf_counter()
            r = fn()
            ctx.sync()
            return time.perf_counter() - t0, r

        ens = gc.Ensemble(ctx, [dict(nx=W, ny=H, kappa=1.0, M=1.0)], dt, gc.F64)
        ens.tune()
        for k, seed in enumerate(seeds):
            c0 = cr.initial_state(H, W, seed).astype(np.float64)
            ref = gc.DeviceArray.from_numpy(ctx, c0)
            f32 = gc.DeviceArray.from_numpy(ctx, c0.astype(np.float32))
            lrn = gc.DeviceArray.from_numpy(ctx, c0)
            if k == 0:  # untimed first run of each path
                for arr in (ref, f32):
                    gc.ch_integrate(ctx, g, arr, 1.0, 1.0, dt, 32, gc.RK4)
                net.integrate(g, lrn, 1.0, 1.0, dt, sub, 2)
                for pn, _, _ in phased or []:
                    pn.integrate(g, lrn, 1.0, 1.0, dt, sub, 2)
                ref.upload(ctx, c0)
                f32.upload(ctx, c0.astype(np.float32))
                lrn.upload(ctx, c0)
            t, _ = timed(lambda: gc.ch_integrate(ctx, g, ref, 1.0, 1.0, dt, n_steps, gc.RK4))
            res["plain_fp64"].append(t)
            t, _ = timed(lambda: gc.ch_integrate(ctx, g, f32, 1.0, 1.0, dt, n_steps, gc.RK4))
            res["plain_fp32"].append(t)
            # the fastest plain path for one small grid: the ensemble kernel with one member
            # (all four stages in one kernel per step, 20 steps per hipGraph replay), tuned
            # once before the loop like the other plain paths
            ens.set_state(0, c0)
            ens.step(32)
            ens.set_state(0, c0)
            t, _ = timed(lambda: ens.step(n_steps))
            res["plain_fp64_graph"].append(t)
            ge = ens.get_state(0).reshape(c0.shape)
            t, nf = timed(lambda: net.integrate(g, lrn, 1.0, 1.0, dt, sub, n_steps // sub, 1e-6))
            res["learned_fp64"].append(t)
            fallbacks.append(nf)
            # variant: the first macro step (the white-noise initial state, far from the smooth
            # states the network is trained to advance) by plain RK4, the rest learned
            wrm = gc.DeviceArray.from_numpy(ctx, c0)

            def warm():
                gc.ch_integrate(ctx, g, wrm, 1.0, 1.0, dt, sub, gc.RK4)
                return net.integrate(g, wrm, 1.0, 1.0, dt, sub, n_steps // sub - 1, 1e-6)
            t, _ = timed(warm)
            res["learned_fp64_warm1"].append(t)
            if phased is not None:
                phs = gc.DeviceArray.from_numpy(ctx, c0)
                t, nfp = timed(lambda: cr.integrate_phased(ctx, g, phs, phased, 1.0, 1.0, dt, sub,
                                                           n_steps // sub))
                res["learned_fp64_phased"].append(t)
                fallbacks_p.append(nfp)
            r, l, q = ref.numpy(ctx), lrn.numpy(ctx), f32.numpy(ctx).astype(np.float64)
            assert np.array_equal(ge, r), "one-member ensemble differs from plain RK4"
            w = wrm.numpy(ctx)
            if phased is not None:
                ph = phs.numpy(ctx)
                errs_p.append(float(np.linalg.norm(ph - r) / np.linalg.norm(r)))
                phs.free()
            errs.append(float(np.linalg.norm(l - r) / np.linalg.norm(r)))
            errs32.append(float(np.linalg.norm(q - r) / np.linalg.norm(r)))
            errs_w.append(float(np.linalg.norm(w - r) / np.linalg.norm(r)))
            for arr in (ref, f32, lrn, wrm):
                arr.free()
        ens.close()
        net.close()
        for pn, _, _ in phased or []:
            pn.close()
    med = {k: float(np.median(v)) for k, v in res.items()}
    cpu = None
    if not a.no_cpu:  # the oracle's RK4 (FP64, 1 thread) on a prefix, scaled to T
        orc = entry.load_oracle()
        orc.lib()
        _pin_one_core()
        og = orc.Grid.make(W, H, 1, 1.0)
        c0 = cr.initial_state(H, W, seeds[0]).astype(np.float64)
        t0 = time.perf_counter()
        orc.ch_integrate(og, c0, 1.0, 1.0, dt, 160, orc.RK4)
        el = time.perf_counter() - t0
        cpu = {"value": el * n_steps / 160, "unit": "s to T=20",
               **cpu_record(orc, f"CPU oracle RK4 (FP64, 1 thread): the first 160 of 2000 steps "
                                 f"({el:.2f} s), extrapolated to T=20")}
    best_plain = min(med["plain_fp64"], med["plain_fp64_graph"])
    variants = {"learned_fp64": max(errs), "learned_fp64_warm1": max(errs_w)}
    if errs_p:
        variants["learned_fp64_phased"] = max(errs_p)
    phased_out = {}
    if errs_p:
        phased_out = {"rel_l2_learned_phased_vs_fp64": max(errs_p),
                      "speedup_phased_vs_plain_fp64": med["plain_fp64"] / med["learned_fp64_phased"],
                      "speedup_phased_vs_plain_fp64_graph":
                          med["plain_fp64_graph"] / med["learned_fp64_phased"],
                      "fallbacks_phased": fallbacks_p,
                      "phases": [[lo, hi] for _, lo, hi, _ in cr.PHASES], "phased_warm": cr.WARM}
    return {"metric": "learned vs plain time-to-T", "T": n_steps * dt, "grid": [H, W],
            "cpu_baseline": cpu,
            "macro_steps": n_steps // sub, "n_ic": len(seeds),
            "seconds": med,
            "speedup_vs_plain_fp64": med["plain_fp64"] / med["learned_fp64"],
            "speedup_vs_plain_fp32": med["plain_fp32"] / med["learned_fp64"],
            "speedup_vs_plain_fp64_graph": med["plain_fp64_graph"] / med["learned_fp64"],
            "rel_l2_learned_vs_fp64": max(errs), "rel_l2_plain_fp32_vs_fp64": max(errs32),
            "rel_l2_learned_warm1_vs_fp64": max(errs_w),
            "speedup_warm1_vs_plain_fp64": med["plain_fp64"] / med["learned_fp64_warm1"],
            **phased_out,
            "fallbacks": fallbacks, "target_rel_l2": 1e-3, "plain_tile_ox_oy": tiles,
            # any variant within the error bound at >= 5x the fastest plain path
            "target_met": any(e <= 1e-3 and best_plain / med[k] >= 5 for k, e in variants.items()),
            "training": {"seconds": meta.get("seconds"), "final_loss": meta["history"][-1],
                         "pairs": "snapshot m -> m+1 of 2000 FP64 GPU RK4 trajectories"}}


def diffuse(a):
    """Heterogeneous-D diffusion (fvm::cellLaplacian right-hand side) and constant-D diffusion,
    3-D 256^3, h = 2 pi / 256, RK4: ms per step and the roofline on 16 W per site per step (the
    stage pipeline) + 4 W for reading D once per stage (heterogeneous D).  D ~ 0.1 + 0.05
    cos(x); psi ~ U(-1, 1); dt at 0.4 of the CFL bound.  Both FP64 and FP32."""
    gc = entry.load_package()
    n = a.n or 256
    h = 2 * np.pi / n
    g = gc.Grid.make(n, n, n, h)
    psi = gc.inputs.symmetric(61, n ** 3).reshape(n, n, n)
    x = np.arange(n) * h
    Dh = np.broadcast_to(0.1 + 0.05 * np.cos(x)[None, None, :], (n, n, n)).copy()
    dt = 0.4 * 0.6963 * h * h / (3 * 0.15)
    steps = min(a.steps, 200)
    rows = []
    with gc.Context() as ctx:
        for dtype, w in ((np.float64, 8), (np.float32, 4)):
            d_p = gc.DeviceArray.from_numpy(ctx, psi.astype(dtype))
            d_D = gc.DeviceArray.from_numpy(ctx, Dh.astype(dtype))
            for kind in ("field", "const"):
                def run(k):
                    if kind == "field":
                        gc.diffuse_field(ctx, g, d_p, d_D, dt, k, gc.RK4)
                    else:
                        gc.diffuse_const(ctx, g, d_p, 0.15, dt, k, gc.RK4)
                run(2)
                ctx.sync()
                ctx.timer_start()
                run(steps)
                ms = ctx.timer_stop() / steps
                words = 20 if kind == "field" else 16
                gbs = words * w * n ** 3 / (ms / 1e3) / 1e9
                rows.append({"D": kind, "dtype": "f64" if w == 8 else "f32", "ms_per_step": ms,
                             "site_updates_per_s": n ** 3 / (ms / 1e3), "GBs_alg": gbs,
                             "hbm_frac": gbs / HBM_PEAK_GBS})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
            d_p.free()
            d_D.free()
    cpu = None
    if not a.no_cpu:
        orc = entry.load_oracle()
        orc.lib()
        _pin_one_core()
        og = orc.Grid.make(n, n, n, h)
        t0 = time.perf_counter()
        orc.diffuse_field(og, psi, Dh, dt, 2, orc.RK4)
        el = time.perf_counter() - t0
        cpu = {"value": n ** 3 * 2 / el, "unit": "site-updates/s",
               **cpu_record(orc, f"CPU oracle heterogeneous-D RK4 (FP64, 1 thread): the rate of "
                                 f"2 steps of the same {n}^3 input ({el:.1f} s)")}
    return {"metric": "diffusion RK4 ms per step", "grid": [n, n, n], "steps": steps, "rows": rows,
            "cpu_baseline": cpu}


def tune(a):
    """The tile autotuner's measurements for an n x n x nz grid (default 256^3; --nz 1 for
    2-D): ms per RK4 step of every candidate shape and kernel family (kind 0: one kernel per
    stage, 1 and 2: two-stage kernels, 3: one kernel per step), FP32 and FP64."""
    gc = entry.load_package()
    n = a.n or 256
    nz = a.nz if a.nz is not None else n
    g = gc.Grid.make(n, n, nz, 1.0)
    rows = []
    with gc.Context() as ctx:
        for dtype, name, w in ((gc.F32, "f32", 4), (gc.F64, "f64", 8)):
            gc.ch_tune(ctx, g, dtype)  # first launches of every candidate
            best, ms = gc.ch_tune(ctx, g, dtype)
            for i, m in enumerate(ms):
                rows.append({"dtype": name, "id": i, "tile_ox_oy_zc": gc.ch_config_shape(g, i),
                             "kind": gc.ch_config_kind(g, i), "ms_per_step": m,
                             "hbm_frac_16W": WORDS_PER_SITE_STEP * w * n * n * nz / (m / 1e3)
                             / 1e9 / HBM_PEAK_GBS, "best": i == best})
    return {"metric": "RK4 ms per step by tile shape", "grid": [n, n, nz], "rows": rows}


def reduce(a):
    """The reduction path on the 3-D workload's grid (default 256^3): integrate() pairwise and
    Kahan, and the fused Cahn-Hilliard free energy (gradient + density + reduction) pairwise
    and Kahan, FP64 and FP32 inputs; ms per call (HIP events around 20 calls, each ending
    with its result on the host) and GB/s on the algorithmic 1 word per site."""
    gc = entry.load_package()
    n = a.n or 256
    g = gc.Grid.make(n, n, n, 1.0)
    c = gc.inputs.ch_initial(n, n, n, 0.1, seed=3)
    rows = []
    with gc.Context() as ctx:
        for np_dtype, w, name in ((np.float64, 8, "f64"), (np.float32, 4, "f32")):
            d = gc.DeviceArray.from_numpy(ctx, c.astype(np_dtype))
            calls = {"integrate_pairwise": lambda: gc.integrate(ctx, g, d, gc.PAIRWISE),
                     "integrate_kahan": lambda: gc.integrate(ctx, g, d, gc.KAHAN),
                     "free_energy_pairwise": lambda: gc.ch_free_energy(ctx, g, d, 0.5, gc.PAIRWISE),
                     "free_energy_kahan": lambda: gc.ch_free_energy(ctx, g, d, 0.5, gc.KAHAN)}
            for op, fn in calls.items():
                fn()
                reps = 20
                ctx.timer_start()
                for _ in range(reps):
                    fn()
                ms = ctx.timer_stop() / reps
                gbs = w * n ** 3 / (ms / 1e3) / 1e9
                rows.append({"op": op, "dtype": name, "ms_per_call": ms, "GBs_alg": gbs,
                             "hbm_frac": gbs / HBM_PEAK_GBS})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
            d.free()
    return {"metric": "reduction ms per call", "grid": [n, n, n], "rows": rows}


def rhs(a):
    """Right-hand-side paths of solve::integrate(..., RK4) through the C++ header API
    (grid-calculus_amd/cpp/examples/rhs_paths.cpp, built in-tree): the caller's host-loop lambda
    of the Cahn-Hilliard example, the same lambda written with tensor::expr expressions, and the
    fused CahnHilliardRhs, on 128^2 and 64^3 (ms per RK4 step, field transfers per step,
    identical bits); plus the wall clock of the unchanged reference example program
    (cpp/bin/ref_examples/cahn_hilliard, 64^3, 100 RK4 steps, one diagnostic line and one
    ASCII VTK snapshot at the end) when it was built."""
    import subprocess
    import tempfile
    cpp = os.path.join(ROOT, "grid-calculus_amd", "cpp", "bin")
    r = subprocess.run([os.path.join(cpp, "rhs_paths")], capture_output=True, text=True,
                       timeout=600, check=True)
    rows = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    out = {"metric": "RK4 ms per step by right-hand side", "rows": rows}
    ex = os.path.join(cpp, "ref_examples", "cahn_hilliard")
    if os.path.exists(ex):
        with tempfile.TemporaryDirectory() as tmp:
            r = subprocess.run([ex, "--n-x", "64", "--n-steps", "100", "--snapshot-every", "100",
                                "--out-dir", tmp], capture_output=True, text=True, timeout=600,
                               check=True)
        wall = [float(x.split("=")[1].rstrip("s")) for x in r.stdout.split()
                if x.startswith("wall_clock=")]
        out["reference_example_64cubed_100_steps_s"] = wall[0] if wall else None
    return out


def copy(a):
    """Plain device-to-device copy of a known size (256 MiB): the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    gc = entry.load_package()
    nbytes = 256 << 20
    with gc.Context() as ctx:
        src = gc.DeviceArray((nbytes // 8,), np.float64)
        dst = gc.DeviceArray((nbytes // 8,), np.float64)
        dst.copy_from(ctx, src)
        ctx.sync()
        reps = 10
        ctx.timer_start()
        for _ in range(reps):
            dst.copy_from(ctx, src)
        ms = ctx.timer_stop() / reps
    return {"metric": "device copy GB/s", "bytes_per_copy": nbytes, "ms": ms,
            "GBs": 2 * nbytes / (ms / 1e3) / 1e9}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the stencil and "
              "RK4 path would shard by z-slabs with a halo exchange per stage; not built)",
              file=sys.stderr)
        sys.exit(2)
    out = {"ch3d": ch3d, "sweep": sweep, "copy": copy, "ensemble": ensemble,
           "learned": learned, "tune": tune, "diffuse": diffuse, "reduce": reduce,
           "rhs": rhs}[a.workload](a)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
