"""Benchmark of the dynamical-core step on one MI355X.

Workload (default): the 1024x1024x96 mountain-wave configuration (Witch-of-
Agnesi ridge, N = 0.01 1/s, U = 10 m/s, dt = 6 s, 6 acoustic substeps per RK3
stage).  --config ensemble runs the 48-member uneven ensemble (nx, ny in
96..320, nz in 40/60/80, dx = 2 km, dt = 8 s) as one batch; --config
nest_hierarchy the ratio-3 nest hierarchy (one step = one parent step).  The headline
value is the FP32-compensated batched step (VIBE_MI_MIXED_COMP: FP64 state and
RK3 accumulation, FP32 tendencies/substeps with u carrying an 8-bit remainder
and the x part of the horizontal mass-flux divergence in FP64, FP64
compensated reductions; one persistent launch per kernel phase over the tile
work list of all members, each step replayed from a hipGraph) in
cell-updates/s: the FP32 mode that keeps every prognostic within 1e-4 of
FP64 after 1 h.  --headline mixed reports the plain FP32-mixed mode
(VIBE_MI_MIXED) there instead.  Beside it: the plain FP64 batched step on the
same inputs, the other FP32 mode, and every member stepped alone through the
single-domain path.

One "step" = one WS-RK3 step of every member (3 stages, 3 x 6 substeps,
sponge).  Inputs are resident in HBM before timing starts.  Timing: W untimed
warm-up steps, then K steps bracketed by device synchronisation, measured
with HIP events on the batch stream.  Per-group kernel times come from a
second, eager (no graph) run with HIP events around each stage kernel and
around each stage's substep launches.

Roofline: algorithmic bytes per cell-update (compulsory prognostic traffic)
  FP64          8 * (48 + 30 * n_s) = 1824 B     (n_s = 6)
  mixed         8 * 48 + 4 * 30 * n_s = 1104 B
  compensated   mixed + u's 1-byte remainder on 6 n_s substep accesses = 1140 B
against 8.0 TB/s.  The dominant kernel is the fused acoustic substep
kb_substep (3 n_s launches per step, about 70 % of the step).  Its algorithmic
bytes per launch: read u, v, w, rho', pi' and write them once, 4 B each in
the FP32 paths (40 B per cell; compensated 42 B with u's remainder), except
the step's last substep, which writes the FP64 state (20 + 40 = 60 B per
cell, compensated 62 B; the intermediate stage states stay FP32); averaged
over the 3 n_s launches of a step.  This implementation moves less
than that model: its FP32 substeps carry chi = rho' - pi'/kappa, which no
substep but a stage's last reads, in place of rho' (32 B per cell on the
first n_s - 1 substeps of a stage); "bytes_moved_per_cell_update" and the
roofline's "moved_bytes_per_launch" give what it moves, while "achieved"
and "hbm_frac_step" keep the model's bytes.  The launch duration is measured
with HIP events on the batch stream around each stage's substep group (only
kb_substep launches are in the group), summed over the step.
"traffic" (HBM bytes per launch) comes from the rocprofv3 FETCH_SIZE /
WRITE_SIZE per-dispatch summary committed for the same workload, each
scaled by what the counter reads for a plain 1 GiB elementwise copy
(--copy-calibration).

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


def bytes_per_cell_update(nsub: int, mixed: bool, compensated: bool = False) -> int:
    """Compulsory bytes per cell-update of the reference's step structure:
    FP64 stage traffic (48 values) plus the substeps' 30 nsub values per cell
    (u, v, w, rho', pi' read and written), 4 B in FP32, 8 B in FP64; the
    compensated mode adds u's 1-byte remainder to its 6 nsub u accesses."""
    return 8 * 48 + (4 if mixed else 8) * 30 * nsub + (6 * nsub if compensated else 0)


def bytes_moved_per_cell_update(nsub: int, compensated: bool = False) -> int:
    """What this implementation's FP32 step moves per cell-update under the
    same model: the substeps carry chi = rho' - pi'/kappa in place of rho' and
    leave it untouched, so the first nsub-1 substeps of a stage read and write
    u, v, w, pi' only (8 values); a stage's last substep also reads chi and
    writes rho' (10 values).  Compensated: plus u's 1-byte remainder read and
    written by every substep."""
    per_stage = 4 * (8 * (nsub - 1) + 10) + (2 * nsub if compensated else 0)
    return 8 * 48 + 3 * per_stage


def substep_moved_bytes_per_launch(cells: int, nsub: int, compensated: bool = False) -> float:
    """Bytes this implementation's kb_substep moves per launch (FP32 batch),
    averaged over the 3 nsub launches of a step: 32 B per cell for the first
    nsub-1 substeps of a stage (u, v, w, pi' read and written), 40 B for the
    last of stages 1 and 2 (chi read, rho' written) and 60 B for the step's
    last (the FP64 state written); the compensated mode adds u's 1-byte
    remainder read and written (2 B) to every launch."""
    n = 3 * nsub
    extra = 2.0 if compensated else 0.0
    return cells * ((32.0 + extra) * 3 * (nsub - 1) + (40.0 + extra) * 2 + 60.0 + extra) / n


def substep_bytes_per_launch(cells: int, nsub: int, compensated: bool = False) -> float:
    """Algorithmic bytes of one kb_substep launch (FP32 batch), averaged over
    the 3 nsub launches of a step: 40 B per cell, 60 B for the step's last
    (FP64 state written); compensated plus 2 B (u's remainder)."""
    n = 3 * nsub
    extra = 2.0 if compensated else 0.0
    return cells * ((40.0 + extra) * (n - 1) + 60.0 + extra) / n


def measured_traffic(cases, compensated: bool = False):
    """HBM bytes per kb_substep launch from the committed rocprofv3 summary of
    this workload and FP32 mode (FETCH_SIZE + WRITE_SIZE per dispatch,
    copy-calibrated)."""
    mode = "compensated_" if compensated else ""
    if len(cases) == 1:
        c = cases[0]
        name = f"kb_substep_{mode}{c.name}_{c.nx}x{c.ny}x{c.nz}_hbm_traffic.txt"
    else:
        name = f"kb_substep_{mode}ensemble{len(cases)}_hbm_traffic.txt"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    if "traffic_bytes_per_launch" not in vals:
        return None, None
    return vals["traffic_bytes_per_launch"], "profiles/" + name


def copy_calibration(nbytes: int = 1 << 30, reps: int = 3):
    """Plain copies of nbytes for scaling the FETCH_SIZE / WRITE_SIZE
    counters (run under rocprofv3 --pmc): the runtime's device-to-device copy
    (dispatched as a blit kernel) and torch's elementwise kernel (b = a * 1,
    ordinary vector loads and stores like the dycore kernels); the
    elementwise kernel sets the scale factors.  Also prints the bandwidth
    each copy reaches (bytes read + written over the HIP-event time of the
    last reps - 1 copies), the practical ceiling a streaming kernel meets."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(reps - 1):
            fn()
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / max(reps - 1, 1)
        return 2.0 * nbytes / (ms * 1e-3) / 1e9

    gbs_copy = timed(lambda: b.copy_(a))
    ok = bool(torch.equal(a, b))
    b.zero_()
    gbs_mul = timed(lambda: torch.mul(a, 1.0, out=b))
    ok = ok and bool(torch.equal(a, b))
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ok": ok, "copy_GBps": round(gbs_copy, 1),
                      "elementwise_GBps": round(gbs_mul, 1)}))


def run_batch(cases, precision, steps, warmup, graph=True):
    """Batched path (persistent tile-list kernels, hipGraph replay)."""
    import vibe_mi

    b = vibe_mi.Batch(cases, precision, use_graph=graph)
    nsub = cases[0].nsub
    if warmup:
        b.step(None, nsub, warmup)
    ms = b.step_timed(None, nsub, steps)
    out = {"ms": ms, "kt": b.kernel_times(), "tuning": b.tuning(), "device_bytes": b.device_bytes()}
    b.close()
    return out


def run_alone(cases, precision, steps, warmup):
    """Each member stepped alone through the single-domain path, one after
    the other; returns the summed timed milliseconds."""
    import vibe_mi

    total = 0.0
    for c in cases:
        d = vibe_mi.Domain(c, precision)
        if warmup:
            d.step(c.dt, c.nsub, warmup)
        total += d.step_timed(c.dt, c.nsub, steps)
        d.close()
    return total


def cpu_baseline(name, nsub, threads):
    """Times the CPU oracle on a column subset of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleDycore
    from vibe_mi import configs

    nxs = 96
    if name == "ensemble":
        case = configs.ensemble_member(nxs, nxs, 60, seed=20240612)
        what = f"one ensemble member at {nxs}x{nxs}x60"
    else:
        case = configs.CONFIGS[name](nx=nxs, ny=nxs)
        what = f"{nxs}x{nxs}x{case.nz} columns of the same configuration"
    def timed(nthreads, budget_s, max_steps):
        orc = OracleDycore(case, threads=nthreads)
        orc.step(case.dt, nsub, 1)  # warm
        t0 = time.perf_counter()
        n = 0
        while True:
            orc.step(case.dt, nsub, 1)
            n += 1
            el = time.perf_counter() - t0
            if el > budget_s or n >= max_steps:
                break
        orc.close()
        return case.ncell * n / el, n

    cups, n = timed(threads, 8.0, 20)
    cups1, n1 = timed(1, 8.0, 5)
    return cups, f"{what}, {n} RK3 steps, FP64 oracle ({n1} steps on 1 thread)", cups1


def cpu_literal(name, nsub, budget_s=10.0, max_steps=3):
    """Times the oracle's LITERAL restatement of the reference's step
    (runge_kutta.cpp:77-118 on equations.cpp's operators: full prognostic
    values, explicit forward-backward substeps, the reference's loop nest and
    operation order; single thread, as the reference runs) on the same
    column subset as cpu_baseline.  The literal equations are unbalanced and
    diverge after a few steps, which does not matter for a rate."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    from literal_cases import LiteralDesc, oracle_step
    from oracle_bind import _PD
    from vibe_mi import configs

    nxs = 96
    case = (configs.ensemble_member(nxs, nxs, 60, seed=20240612) if name == "ensemble"
            else configs.CONFIGS[name](nx=nxs, ny=nxs))
    nx, ny, nz = case.nx, case.ny, case.nz
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (case.zeta, case.zs, case.rho0, case.theta0,
                                                                  case.pi0)]
    d = LiteralDesc(nx, ny, nz, case.dx, case.dy, case.ztop, *[a.ctypes.data_as(_PD) for a in keep],
                    2, 1, 45.0, 0.0, 2, 1)  # the reference's defaults: central2, f-plane 45, kessler
    st = case.copy_state()
    sh = (nz, ny, nx)
    def faces(a, axis):
        a = a.reshape(sh)
        return np.ascontiguousarray(np.concatenate([a, np.take(a, [0], axis=axis)], axis=axis).ravel())
    s = [faces(st["u"], 2), faces(st["v"], 1), np.ascontiguousarray(st["w"]),
         np.ascontiguousarray(keep[2] + st["rho"]), np.ascontiguousarray(keep[3] + st["theta"]),
         np.ascontiguousarray(keep[4] + st["pi"])] + [np.zeros(nx * ny * nz) for _ in range(6)]
    t0 = time.perf_counter()
    n = 0
    while True:
        oracle_step(d, s, case.dt, nsub, np.ascontiguousarray(case.sponge))
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= max_steps:
            break
    return {"value": case.ncell * n / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{nx}x{ny}x{nz} columns of the same configuration, {n} literal RK3 steps "
                      f"(runge_kutta.cpp:77-118 restated on the pinned equations.cpp operators), 1 thread"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_nests(args, t0):
    """Nest hierarchy (parent + ratio-3 nests, time refinement): one "step" is
    one parent step, i.e. r**L batched sub-steps; cell-updates count every
    domain's own steps.  The Davies relaxations are launched between the
    batched sub-steps and are inside the wall time measured here."""
    import vibe_mi
    from vibe_mi import configs
    from vibe_mi.nesting import NestedRun

    cases, links = configs.nest_hierarchy()
    cells = sum(c.ncell for c in cases)
    log(f"[bench] built nest hierarchy: {len(cases)} domains, {cells} cells in {time.perf_counter() - t0:.1f}s")
    res = {}
    head = "compensated" if args.headline == "comp" else "mixed"
    other = "mixed" if head == "compensated" else "compensated"
    precs = {"compensated": vibe_mi.MIXED_COMP, "mixed": vibe_mi.MIXED, "fp64": vibe_mi.FP64}
    for name in (head, "fp64", other):
        prec = precs[name]
        if (name == "fp64" and args.no_fp64) or (name == other and args.no_other):
            continue
        run = NestedRun(cases, links, precision=prec)
        cu = run.cell_updates_per_parent_step()
        if args.warmup:
            run.step(args.warmup)
        run.batch.members[0].energy_budget()  # synchronise
        w0 = time.perf_counter()
        ms_launch = run.step(args.steps, timed=True)
        run.batch.members[0].energy_budget()
        wall = time.perf_counter() - w0
        res[name] = {"value": cu * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
                     "ms_batched_launches_per_step": ms_launch / args.steps, "tiles": run.batch.tuning()}
        log(f"[bench] {name}: {wall * 1e3 / args.steps:.2f} ms per parent step, "
            f"{cu * args.steps / wall:.3e} cell-updates/s")
        run.close()
    m = res[head]
    out = {
        "metric": "cell-updates/s per RK3 step + % HBM roofline, FP32-mixed/batched vs FP64",
        "value": m["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 compute, f64 state" if head == "mixed" else
                 "f32 compute (u with an 8-bit remainder, x mass-flux divergence f64), f64 state",
        "data": "synthetic (seeded random terrain, random-phase theta')",
        "config": {"workload": "nest hierarchy: 384x384x60 parent at 3 km, nests 271x271, 211x331, 151x151 at "
                               "1 km, 181x181 at 333 m; dt 18/6/2 s, 6 substeps/stage; one step = one parent step",
                   "cells": cells, "parallelism": "single GPU"},
        "hbm_frac_step": m["value"] * bytes_per_cell_update(6, True, head == "compensated") / HBM_PEAK,
        "fp64": res.get("fp64"),
        "speedup_vs_fp64": (m["value"] / res["fp64"]["value"]) if "fp64" in res else None,
        other: res.get(other),
        "roofline": None, "cpu_baseline": None,
    }
    out["config"]["mode"] = "FP32-compensated" if head == "compensated" else "FP32-mixed"
    if other in res and "fp64" in res:
        res[other]["speedup_vs_fp64"] = res[other]["value"] / res["fp64"]["value"]
    out["config"]["cell_updates_per_step"] = sum(c.ncell * 3 ** c.meta["level"] for c in cases)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="mountain_wave_large",
                    help="mountain_wave_large (default), mountain_wave, warm_bubble, ensemble, nest_hierarchy")
    ap.add_argument("--nx", type=int, default=0)
    ap.add_argument("--members", type=int, default=48, help="ensemble size")
    ap.add_argument("--headline", choices=("comp", "mixed"), default="comp",
                    help="FP32 mode of the headline value: comp (VIBE_MI_MIXED_COMP, the mode within 1e-4 of "
                         "FP64 after 1 h; default) or mixed (VIBE_MI_MIXED)")
    ap.add_argument("--no-other", action="store_true", help="skip the other FP32 mode")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-alone", action="store_true", help="skip the unbatched comparison")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import vibe_mi
    from vibe_mi import configs

    t0 = time.perf_counter()
    if args.config == "nest_hierarchy":
        return bench_nests(args, t0)
    if args.config == "ensemble":
        cases = configs.ensemble(n=args.members)
        wl = (f"ensemble of {len(cases)} members (nx, ny in 96..320, nz in 40/60/80), dx=2 km, "
              f"dt={cases[0].dt}s, {cases[0].nsub} substeps/stage")
    else:
        kw = {"nx": args.nx, "ny": args.nx} if args.nx else {}
        c = configs.CONFIGS[args.config](**kw)
        cases = [c]
        wl = f"{c.name} {c.nx}x{c.ny}x{c.nz}, dt={c.dt}s, {c.nsub} substeps/stage"
    nsub = cases[0].nsub
    cells = sum(c.ncell for c in cases)
    log(f"[bench] built {wl}: {cells} cells in {time.perf_counter() - t0:.1f}s")

    head_comp = args.headline == "comp"
    modes = {True: (vibe_mi.MIXED_COMP, "FP32-compensated"), False: (vibe_mi.MIXED, "FP32-mixed")}
    head_prec, head_name = modes[head_comp]
    other_prec, other_name = modes[not head_comp]

    def mode_summary(cups, ms, tuning, comp_mode):
        return {"value": cups, "ms_per_step": ms, "tiles": tuning,
                "bytes_per_cell_update": bytes_per_cell_update(nsub, True, comp_mode),
                "bytes_moved_per_cell_update": bytes_moved_per_cell_update(nsub, comp_mode),
                "hbm_frac_step": cups * bytes_per_cell_update(nsub, True, comp_mode) / HBM_PEAK}

    log(f"[bench] {head_name} batched (hipGraph) timed run")
    rm = run_batch(cases, head_prec, args.steps, args.warmup, graph=True)
    cups_m = cells * args.steps / (rm["ms"] * 1e-3)
    log(f"[bench] {head_name}: {rm['ms'] / args.steps:.2f} ms/step, {cups_m:.3e} cell-updates/s, tiles {rm['tuning']}")
    log(f"[bench] {head_name} batched eager run with per-group HIP events")
    re_ = run_batch(cases, head_prec, args.steps, 1, graph=False)
    kt_m = re_["kt"]
    res64 = None
    if not args.no_fp64:
        log("[bench] FP64 batched (hipGraph) timed run")
        rd = run_batch(cases, vibe_mi.FP64, args.steps, args.warmup, graph=True)
        cups_d = cells * args.steps / (rd["ms"] * 1e-3)
        log(f"[bench] fp64: {rd['ms'] / args.steps:.2f} ms/step, {cups_d:.3e} cell-updates/s")
        res64 = {"value": cups_d, "ms_per_step": rd["ms"] / args.steps,
                 "hbm_frac": cups_d * bytes_per_cell_update(nsub, False) / HBM_PEAK, "tiles": rd["tuning"]}
        log(f"[bench] {head_name}: {cups_m / cups_d:.3f}x FP64")
    other = None
    if not args.no_other:
        log(f"[bench] {other_name} batched (hipGraph) timed run")
        ro_ = run_batch(cases, other_prec, args.steps, args.warmup, graph=True)
        cups_o = cells * args.steps / (ro_["ms"] * 1e-3)
        other = mode_summary(cups_o, ro_["ms"] / args.steps, ro_["tuning"], not head_comp)
        other["speedup_vs_fp64"] = (cups_o / res64["value"]) if res64 else None
        log(f"[bench] {other_name}: {ro_['ms'] / args.steps:.2f} ms/step, {cups_o:.3e} cell-updates/s"
            + (f", {cups_o / res64['value']:.3f}x FP64" if res64 else ""))
    alone = None
    if not args.no_alone:
        log(f"[bench] each member alone (single-domain path), {head_name}")
        ms_a = run_alone(cases, head_prec, args.steps, args.warmup)
        alone = {"value": cells * args.steps / (ms_a * 1e-3), "ms_per_step": ms_a / args.steps,
                 "batched_speedup": ms_a / rm["ms"]}
        log(f"[bench] alone: {ms_a / args.steps:.2f} ms/step; batched speedup {ms_a / rm['ms']:.3f}x")

    # dominant kernel: the fused acoustic substep (kb_substep) of the headline mode
    sub_ms = kt_m["ms_substep"] / max(kt_m["n_substep"], 1)
    sub_bytes = sum(substep_bytes_per_launch(c.ncell, nsub, head_comp) for c in cases)
    achieved = sub_bytes / (sub_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(cases, head_comp)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved * 1e9 / HBM_PEAK, "traffic": traffic,
                "kernel": f"kb_substep (fused HEVI acoustic substep, {head_name}, batched persistent)",
                "algorithmic_bytes_per_launch": sub_bytes, "avg_launch_ms": sub_ms,
                "moved_bytes_per_launch": substep_moved_bytes_per_launch(cells, nsub, head_comp),
                "traffic_source": traffic_src}

    cpu = None
    if not args.no_cpu:
        # every core this process may run on, up to the host's CPU share per
        # GPU (OMP_NUM_THREADS, 16 on a one-GPU share); 1 thread beside it
        visible = len(os.sched_getaffinity(0))
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or visible
        threads = max(1, min(visible, share))
        log(f"[bench] CPU oracle baseline on {threads} threads ({visible} cores visible)")
        cpu_cups, sample, cpu_cups1 = cpu_baseline(args.config, nsub, threads)
        cpu = {"value": cpu_cups, "unit": "cell-updates/s", "cores": threads, "kind": "port", "sample": sample,
               "value_1_thread": cpu_cups1, "cores_visible": visible, "cpu_model": cpu_model()}
        log("[bench] literal CPU restatement of the reference's step, 1 thread")
        cpu["literal"] = cpu_literal(args.config, nsub)

    out = {
        "metric": "cell-updates/s per RK3 step + % HBM roofline, FP32-mixed/batched vs FP64",
        "value": cups_m,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rm["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 compute (u with an 8-bit remainder, x mass-flux divergence f64), f64 state" if head_comp
                  else "f32 compute, f64 state"),
        "data": "synthetic (seeded terrain/noise/bubble initial states)",
        "config": {"workload": wl, "cells": cells, "members": len(cases), "parallelism": "single GPU",
                   "mode": head_name},
        "hbm_frac_step": cups_m * bytes_per_cell_update(nsub, True, head_comp) / HBM_PEAK,
        "bytes_per_cell_update": bytes_per_cell_update(nsub, True, head_comp),
        "bytes_moved_per_cell_update": bytes_moved_per_cell_update(nsub, head_comp),
        "tiles": rm["tuning"],
        "fp64": res64,
        "speedup_vs_fp64": (cups_m / res64["value"]) if res64 else None,
        "unbatched": alone,
        # the headline mode: VIBE_MI_MIXED_COMP (FP32 compute with u carrying a
        # remainder and the x part of the horizontal mass-flux divergence in
        # FP64), within 1e-4 of FP64 after 1 h; the plain FP32-mixed mode
        # (VIBE_MI_MIXED, 1.24e-4 after 1 h) beside it, or the other way round
        # with --headline mixed
        "mixed" if head_comp else "compensated": other,
        "kernel_ms": {"stage_group_per_stage": kt_m["ms_stage"] / max(kt_m["n_stage"], 1),
                      "substep_per_launch": sub_ms},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device_bytes": rm["device_bytes"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
