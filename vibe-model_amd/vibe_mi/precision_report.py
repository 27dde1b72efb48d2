"""FP32-mixed versus FP64 error report on one GPU.

Runs one workload twice from the same seeded initial state, once with
VIBE_MI_FP64 (plain FP64, the GPU path that matches the CPU oracle to
round-off) and once with VIBE_MI_MIXED (FP64 state and RK3 accumulation,
FP32 tendencies and substeps, FP64 compensated reductions) or, with
--precision comp, VIBE_MI_MIXED_COMP (the same with compensated FP32 substep
state and FP64 mass-flux sums), for a stretch of
simulated time, and reports per prognostic:

  rel_l2   ||x_mixed - x_fp64||_2 / ||x_fp64||_2
  max_abs  max |x_mixed - x_fp64|
  rel_max  max_abs / max |x_fp64|

plus the mass and total-energy drift of each run relative to its initial
state (compensated FP64 sums, energy_budget role, src/dyn/diagnostics.cpp:255-277).

  python -m vibe_mi.precision_report --config mountain_wave_large --hours 1 \
      --out report.json

Prints one progress line per block of steps and the JSON report on stdout.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import FP64, MIXED, MIXED_COMP, Batch, Domain
from . import configs

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def _run(case, precision, nsteps, block, label, batched=True):
    """Steps one domain; batched (default) uses the persistent batched kernels
    with hipGraph replay, the path bench.py measures."""
    if batched:
        b = Batch([case], precision)
        dom = b.members[0]
        advance = lambda n: b.step([case.dt], case.nsub, n)
    else:
        b = None
        dom = Domain(case, precision)
        advance = lambda n: dom.step(case.dt, case.nsub, n)
    e0 = dom.energy_budget()
    t0 = time.perf_counter()
    done = 0
    while done < nsteps:
        n = min(block, nsteps - done)
        advance(n)
        done += n
        print(f"[precision] {label}: {done}/{nsteps} steps, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    e1 = dom.energy_budget()
    st = dom.get_state()
    (b or dom).close()
    return st, e0, e1


def compare(a: dict, b: dict) -> dict:
    """Per-field error of a (mixed) against b (FP64)."""
    out = {}
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        d = x - y
        ny = float(np.linalg.norm(y))
        mabs = float(np.max(np.abs(d))) if d.size else 0.0
        ymax = float(np.max(np.abs(y))) if y.size else 0.0
        out[f] = {"rel_l2": float(np.linalg.norm(d)) / ny if ny > 0 else float(np.linalg.norm(d)),
                  "max_abs": mabs, "rel_max": mabs / ymax if ymax > 0 else mabs,
                  "ref_l2": ny, "ref_max": ymax}
    return out


def drift(e0, e1) -> dict:
    """Relative change of the domain totals {mass, total energy}."""
    return {"mass": (e1[0] - e0[0]) / e0[0], "total_energy": (e1[5] - e0[5]) / e0[5]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="mountain_wave_large")
    ap.add_argument("--hours", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=0, help="overrides --hours")
    ap.add_argument("--nx", type=int, default=0)
    ap.add_argument("--block", type=int, default=50)
    ap.add_argument("--out", default="")
    ap.add_argument("--unbatched", action="store_true", help="single-domain kernels instead of the batched ones")
    ap.add_argument("--precision", choices=("mixed", "comp"), default="mixed",
                    help="comp: VIBE_MI_MIXED_COMP")
    args = ap.parse_args(argv)
    kw = {"nx": args.nx, "ny": args.nx} if args.nx else {}
    case = configs.CONFIGS[args.config](**kw)
    nsteps = args.steps or int(math.ceil(args.hours * 3600.0 / case.dt))
    print(f"[precision] {case.name} {case.nx}x{case.ny}x{case.nz}, dt={case.dt} s, {nsteps} steps "
          f"({nsteps * case.dt / 3600.0:.3f} h simulated)", file=sys.stderr, flush=True)
    ref, r0, r1 = _run(case, FP64, nsteps, args.block, "fp64", not args.unbatched)
    prec = MIXED_COMP if args.precision == "comp" else MIXED
    mix, m0, m1 = _run(case, prec, nsteps, args.block, args.precision, not args.unbatched)
    err = compare(mix, ref)
    rep = {
        "workload": f"{case.name} {case.nx}x{case.ny}x{case.nz}",
        "path": "single-domain kernels" if args.unbatched else "batched kernels (hipGraph)",
        "precision": args.precision,
        "dt": case.dt, "substeps": case.nsub, "steps": nsteps,
        "simulated_hours": nsteps * case.dt / 3600.0,
        "error_mixed_vs_fp64": err,
        "drift_fp64": drift(r0, r1),
        "drift_mixed": drift(m0, m1),
        "max_rel_l2": max(v["rel_l2"] for v in err.values()),
    }
    txt = json.dumps(rep)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(txt + "\n")
    print(txt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
