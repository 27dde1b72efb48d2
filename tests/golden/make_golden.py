"""Writes golden input/output vectors from the reference's own C kernels.

The reference's src/common/c_kernels.c is compiled (unmodified, from where it
lies under /root/reference) by oracle/ref.mk into oracle/_ref/; this script
calls it on seeded inputs and stores inputs and outputs as JSON fixtures.
Run:  make -C oracle -f ref.mk && python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_bind import ref_lib, _pd  # noqa: E402


def main():
    L = ref_lib()
    assert L is not None, "build oracle/_ref first: make -C oracle -f ref.mk"
    rng = np.random.default_rng(20241203)
    out = {}
    # WENO5: a fixed probe stencil plus seeded stencils of mixed scale
    cases = [[0.2, -0.4, 1.1, 0.9, 2.3], [3.5] * 5]
    for _ in range(60):
        cases.append(list(rng.normal(size=5) * float(rng.choice([1e-4, 1.0, 50.0]))))
    wc = []
    for q in cases:
        l, r = ctypes.c_double(), ctypes.c_double()
        L.vibe_weno5_reconstruct_f64(*q, ctypes.byref(l), ctypes.byref(r))
        wc.append({"q": [float(v) for v in q], "q_left": l.value, "q_right": r.value})
    json.dump({"source": "src/common/c_kernels.c:64-104", "cases": wc}, open(os.path.join(HERE, "weno5.json"), "w"),
              indent=1)
    # CFL scan and flux advection on one seeded 3-D case, with and without terrain
    nx, ny, nz = 7, 5, 6
    n = nx * ny * nz
    u, v, w = rng.normal(size=n) * 10, rng.normal(size=n) * 10, rng.normal(size=n)
    q, rho = rng.uniform(1, 2, n), rng.uniform(0.5, 1.2, n)
    dzeta = rng.uniform(0.05, 0.2, nz)
    zs = rng.uniform(0, 500, nx * ny)
    rec = {"source": "src/common/c_kernels.c:148-187,209-286", "nx": nx, "ny": ny, "nz": nz, "dx": 1000.0,
           "dy": 900.0, "z_top": 20000.0, "dt": 3.0, "u": u.tolist(), "v": v.tolist(), "w": w.tolist(),
           "q": q.tolist(), "rho": rho.tolist(), "dzeta": dzeta.tolist(), "zs": zs.tolist(), "results": []}
    for use_zs in (False, True):
        zz = zs if use_zs else None
        mw, md = ctypes.c_double(), ctypes.c_double()
        c = L.vibe_cfl_scan_f64(_pd(u), _pd(v), _pd(w), _pd(dzeta), _pd(zz), nx, ny, nz, 1000.0, 900.0, 20000.0,
                                3.0, ctypes.byref(mw), ctypes.byref(md))
        dq = np.zeros(n)
        L.vibe_advect_scalar_flux_f64(_pd(q), _pd(rho), _pd(u), _pd(v), _pd(w), _pd(dzeta), _pd(zz), nx, ny, nz,
                                      1000.0, 900.0, 20000.0, _pd(dq))
        rec["results"].append({"zs": use_zs, "cfl": c, "max_wind": mw.value, "min_dz": md.value,
                               "dqdt": dq.tolist()})
    json.dump(rec, open(os.path.join(HERE, "cfl_flux.json"), "w"))
    # bilinear
    halo, bx, by, bz = 2, 5, 4, 2
    nsx, nsy = bx + 2 * halo, by + 2 * halo
    f = rng.normal(size=nsx * nsy * (bz + 2 * halo))
    pts = [(2.0, 1.0, 1), (1.5, 1.0, 0), (-50.0, -50.0, 0), (3.7, 2.2, 5), (0.3, 9.0, -1), (4.0, 3.0, 1)]
    br = [{"xi": xi, "yj": yj, "k": k, "value": L.vibe_bilinear_f64(_pd(f), halo, nsx, nsy, bz, xi, yj, k)}
          for xi, yj, k in pts]
    json.dump({"source": "src/common/c_kernels.c:108-144", "halo": halo, "nsx": nsx, "nsy": nsy, "nz": bz,
               "f": f.tolist(), "points": br}, open(os.path.join(HERE, "bilinear.json"), "w"))
    # Neumaier (sequential order of the reference)
    arrs = [rng.normal(size=1000) * 10.0 ** rng.integers(-6, 6, size=1000), np.ones(10000)]
    arrs[1][0], arrs[1][-1] = 1e16, -1e16
    json.dump({"source": "src/common/c_kernels.c:190-206",
               "cases": [{"x": a.tolist(), "sum": L.vibe_neumaier_sum_f64(_pd(a), a.size)} for a in arrs]},
              open(os.path.join(HERE, "neumaier.json"), "w"))
    # Thomas: the reference's C implementation on its own unit-test system
    # (tests/unit/test_common.cpp:163-175), recorded to document its defect.
    nt = 24
    a, c, b, d = np.ones(nt - 1), np.ones(nt - 1), -4.0 * np.ones(nt), np.ones(nt)
    x, ws = np.zeros(nt), np.zeros(nt)
    rc = L.vibe_tridiagonal_solve_f64(_pd(a), _pd(b), _pd(c), _pd(d), nt, _pd(x), _pd(ws))
    json.dump({"source": "src/common/c_kernels.c:18-60", "n": nt, "rc": rc, "x_reference": x.tolist()},
              open(os.path.join(HERE, "thomas_reference_defect.json"), "w"), indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
