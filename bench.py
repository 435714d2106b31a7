#!/usr/bin/env python3
"""Benchmark: fused RK4 Cahn-Hilliard site-updates/s on one MI355X.

Workload (the RK4 configuration of the headline metric): 3-D Cahn-Hilliard on a 256^3
periodic grid, dx=1, M=1, kappa=0.5, dt=0.005, c0 = 0.1*U(-1,1); stencils and RK4 stages
in FP32, total mass and free energy reduced in FP64 (compensated) every 500 steps, as the
reference driver prints diagnostics per snapshot chunk.  One "step" is one RK4 step of the
whole grid (4 fused stage kernels).  Inputs are resident in HBM before timing starts.

Also reported: the FP32 result's L2/Linf error against an FP64 GPU run of the same steps,
the FP64 pipeline's own throughput, the roofline position of the stage kernel, and the
single-core CPU oracle (a faithful C restatement of the reference path, FP64) timed on a
bounded prefix of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as entry  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORDS_PER_SITE_STEP = 16  # algorithmic HBM words per site per fused RK4 step (4 per stage)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--diag-every", type=int, default=500)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=0, help="CPU oracle sample steps (0=auto)")
    return p.parse_args()


def run_gpu(gc, ctx, g, c0, dtype, steps, warmup, diag_every, kappa, dt):
    d = gc.DeviceArray.from_numpy(ctx, c0.astype(dtype))
    if warmup:
        w = gc.DeviceArray.from_numpy(ctx, c0.astype(dtype))
        gc.ch_integrate(ctx, g, w, 1.0, kappa, dt, warmup, gc.RK4)
        ctx.sync()
        w.free()
    diags = []
    ctx.sync()
    t0 = time.perf_counter()
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
    wall = time.perf_counter() - t0
    return d, ms, wall, diags


def stage_kernel_ms(gc, ctx, g, c0, dtype, steps, kappa, dt):
    """Average duration of the fused stage kernel: HIP events on the context stream around
    `steps` RK4 steps with nothing else in between (4 launches per step)."""
    d = gc.DeviceArray.from_numpy(ctx, c0.astype(dtype))
    gc.ch_integrate(ctx, g, d, 1.0, kappa, dt, 2, gc.RK4)
    ctx.sync()
    ctx.timer_start()
    gc.ch_integrate(ctx, g, d, 1.0, kappa, dt, steps, gc.RK4)
    ms = ctx.timer_stop()
    d.free()
    return ms / (4 * steps)


def cpu_baseline(orc, n, c0, kappa, dt, steps):
    og = orc.Grid.make(n, n, n, 1.0)
    t0 = time.perf_counter()
    orc.ch_integrate(og, c0, 1.0, kappa, dt, steps, orc.RK4)
    el = time.perf_counter() - t0
    return n ** 3 * steps / el, el


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the stencil and "
              "RK4 path would shard by z-slabs with halo exchange, not built here)", file=sys.stderr)
        sys.exit(2)
    gc = entry.load_package()
    n, kappa, dt = a.n, 0.5, 0.005
    g = gc.Grid.make(n, n, n, 1.0)
    c0 = gc.inputs.ch_initial(n, n, n, 0.1, seed=3)
    sites = n ** 3

    with gc.Context() as ctx:
        d32, ms, wall, diags = run_gpu(gc, ctx, g, c0, np.float32, a.steps, a.warmup,
                                       a.diag_every, kappa, dt)
        ms_per_step = ms / a.steps
        value = sites * a.steps / (ms / 1e3)
        k_ms = stage_kernel_ms(gc, ctx, g, c0, np.float32, min(a.steps, 200), kappa, dt)
        bytes_launch = WORDS_PER_SITE_STEP / 4 * 4 * sites
        achieved = bytes_launch / (k_ms / 1e3) / 1e9
        extra = {}
        if not a.no_fp64:
            d64, ms64, _, diags64 = run_gpu(gc, ctx, g, c0, np.float64, a.steps, 0,
                                            a.diag_every, kappa, dt)
            r32 = d32.numpy(ctx).astype(np.float64)
            r64 = d64.numpy(ctx)
            l2 = float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))
            linf = float(np.abs(r32 - r64).max() / np.abs(r64).max())
            k64 = stage_kernel_ms(gc, ctx, g, c0, np.float64, min(a.steps, 100), kappa, dt)
            extra["fp32_vs_fp64"] = {"l2_rel": l2, "linf_rel": linf}
            extra["fp64"] = {
                "site_updates_per_s": sites * a.steps / (ms64 / 1e3),
                "ms_per_step": ms64 / a.steps,
                "stage_kernel_ms": k64,
                "hbm_frac": (WORDS_PER_SITE_STEP / 4 * 8 * sites) / (k64 / 1e3) / 1e9 / HBM_PEAK_GBS,
                "mass_final": diags64[-1][1], "F_final": diags64[-1][2],
            }
        extra["diagnostics_fp32"] = [{"step": s, "mass": m, "F": f} for s, m, f in diags]

    cpu = None
    if not a.no_cpu:
        orc = entry.load_oracle()
        orc.lib()
        steps_cpu = a.cpu_steps or 4
        cores = 1
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except Exception:
            pass
        v, el = cpu_baseline(orc, n, c0, kappa, dt, steps_cpu)
        cpu = {"value": v, "unit": "site-updates/s", "cores": cores, "kind": "port",
               "sample": f"oracle (C restatement of the reference path, FP64, -O2) RK4 on the "
                         f"same {n}^3 input, {steps_cpu} steps ({el:.1f} s), 1 thread"}

    out = {
        "metric": "RK4 site-updates/s",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3-D Cahn-Hilliard fused RK4, {n}^3, dx=1, M=1, kappa=0.5, "
                               "dt=0.005, c0=0.1*U(-1,1); FP32 stages, FP64 compensated mass/F "
                               f"every {a.diag_every} steps",
                   "grid": [n, n, n], "dt": dt, "kappa": kappa},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_stage<float, CHRhs<float>>",
                     "kernel_avg_ms": k_ms, "bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
