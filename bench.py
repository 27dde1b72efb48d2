"""Benchmark of the dynamical-core step on one MI355X.

Workload (default): the 1024x1024x96 mountain-wave configuration (Witch-of-
Agnesi ridge, N = 0.01 1/s, U = 10 m/s, dt = 6 s, 6 acoustic substeps per RK3
stage).  The headline value is the FP32-mixed step (FP64 state and RK3
accumulation, FP32 tendencies/substeps, FP64 compensated reductions) in
cell-updates/s; the plain FP64 step is timed beside it on the same inputs.

One "step" = one WS-RK3 step of the whole domain (3 stages, 3 x 6 substeps,
sponge).  Inputs are resident in HBM before timing starts.  Timing: W untimed
warm-up steps, then K steps bracketed by device synchronisation, measured
with HIP events on the domain's stream; per-kernel times from events around
each stage kernel group and each substep group.

Roofline: algorithmic bytes per cell-update (compulsory prognostic traffic)
  FP64   8 * (48 + 30 * n_s) = 1824 B     (n_s = 6)
  mixed  8 * 48 + 4 * 30 * n_s = 1104 B
against 8.0 TB/s.  The dominant kernel group (acoustic substeps) is reported
with its own algorithmic bytes: 10 field accesses per cell per substep.

cpu_baseline: the FP64 CPU oracle (a restatement of the reference's step)
timed on this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vibe-model_amd"))
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_cell_update(nsub: int, mixed: bool) -> int:
    return 8 * 48 + (4 if mixed else 8) * 30 * nsub


def run_gpu(case, precision, steps, warmup):
    import vibe_mi

    dom = vibe_mi.Domain(case, precision)
    if warmup:
        dom.step(case.dt, case.nsub, warmup)
    ms = dom.step_timed(case.dt, case.nsub, steps)
    kt = dom.kernel_times()
    e = dom.energy_budget()
    nbytes = dom.device_bytes()
    dom.close()
    return ms, kt, e, nbytes


def cpu_baseline(name, nsub, threads):
    """Times the CPU oracle on a column subset of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleDycore
    from vibe_mi import configs

    nxs = 96
    case = configs.CONFIGS[name](nx=nxs, ny=nxs)
    orc = OracleDycore(case, threads=threads)
    orc.step(case.dt, nsub, 1)  # warm
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.step(case.dt, nsub, 1)
        n += 1
        el = time.perf_counter() - t0
        if el > 8.0 or n >= 20:
            break
    cups = case.ncell * n / el
    orc.close()
    return cups, f"{nxs}x{nxs}x{case.nz} columns of the same configuration, {n} RK3 steps, FP64 oracle"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="mountain_wave_large")
    ap.add_argument("--nx", type=int, default=0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import vibe_mi
    from vibe_mi import configs

    kw = {}
    if args.nx:
        kw = {"nx": args.nx, "ny": args.nx}
    t0 = time.perf_counter()
    case = configs.CONFIGS[args.config](**kw)
    log(f"[bench] built {case.name} {case.nx}x{case.ny}x{case.nz} in {time.perf_counter() - t0:.1f}s")
    cells = case.ncell

    log("[bench] FP32-mixed timed run")
    ms_m, kt_m, e_m, mem_m = run_gpu(case, vibe_mi.MIXED, args.steps, args.warmup)
    cups_m = cells * args.steps / (ms_m * 1e-3)
    log(f"[bench] mixed: {ms_m / args.steps:.2f} ms/step, {cups_m:.3e} cell-updates/s")
    res64 = None
    if not args.no_fp64:
        log("[bench] FP64 timed run")
        ms_d, kt_d, e_d, mem_d = run_gpu(case, vibe_mi.FP64, args.steps, args.warmup)
        cups_d = cells * args.steps / (ms_d * 1e-3)
        log(f"[bench] fp64: {ms_d / args.steps:.2f} ms/step, {cups_d:.3e} cell-updates/s")
        res64 = {"value": cups_d, "ms_per_step": ms_d / args.steps,
                 "hbm_frac": cups_d * bytes_per_cell_update(case.nsub, False) / HBM_PEAK,
                 "ms_stage_per_launch_group": kt_d["ms_stage"] / max(kt_d["n_stage"], 1),
                 "ms_substep_per_launch_group": kt_d["ms_substep"] / max(kt_d["n_substep"], 1)}

    # dominant kernel group: acoustic substep (10 field accesses per cell)
    sub_ms = kt_m["ms_substep"] / max(kt_m["n_substep"], 1)
    sub_bytes = 10 * 4 * cells
    achieved = sub_bytes / (sub_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved * 1e9 / HBM_PEAK, "traffic": None,
                "kernel": "acoustic substep (k_sub_h + k_sub_v), FP32", "algorithmic_bytes_per_launch": sub_bytes,
                "avg_launch_ms": sub_ms}

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] CPU oracle baseline on {threads} threads")
        cpu_cups, sample = cpu_baseline(args.config, case.nsub, threads)
        cpu = {"value": cpu_cups, "unit": "cell-updates/s", "cores": threads, "kind": "port", "sample": sample}

    out = {
        "metric": "cell-updates/s per RK3 step + % HBM roofline, FP32-mixed/batched vs FP64",
        "value": cups_m,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 compute, f64 state",
        "data": "synthetic (seeded terrain/noise initial state)",
        "config": {"workload": f"{case.name} {case.nx}x{case.ny}x{case.nz}, dt={case.dt}s, "
                               f"{case.nsub} substeps/stage", "cells": cells, "parallelism": "single GPU"},
        "hbm_frac_step": cups_m * bytes_per_cell_update(case.nsub, True) / HBM_PEAK,
        "bytes_per_cell_update": bytes_per_cell_update(case.nsub, True),
        "fp64": res64,
        "speedup_mixed_vs_fp64": (cups_m / res64["value"]) if res64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device_bytes": mem_m,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
