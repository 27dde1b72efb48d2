"""The C++ caller (vibe-model_amd/examples/run_warm_bubble.cpp, built by the
library's Makefile against include/vibe_mi/dycore.hpp) on the GPU, against
the FP64 oracle: the role of the reference's embed_model example driving
Integrator::step (examples/embed_model.cpp:61-131, integrator.hpp:62-82).

* Device loop (advance): 5 steps of the 24 x 20 x 16 warm bubble (dx = dz =
  250 m, dt = 2 s, 6 acoustic substeps) equal the oracle to 1e-12 relative
  L2 per prognostic.
* Reference-layout loop: every step converts the reference's 12-species
  halo-4 State to the perturbation state, calls Integrator::step through the
  base class and converts back (from_reference / to_reference).  Velocities
  match the oracle at 1e-12; rho, theta, pi are held as full values
  rho0 + rho' in that layout (as the reference holds them), so each step
  rounds the perturbation to the ulp of the full value: the bar is 1e-13 of
  the full value's magnitude.  IntegratorStats counts 5 steps and 3 x 6
  acoustic substeps per step, the boundary callback runs once per step.
* max_stable_dt printed by the caller equals min(0.8 / CFL(1), 2 / (cs_max
  max(1/dx + 1/dy + 1/dz))) of the final state computed here in numpy
  (runge_kutta.cpp:121-141, advection.cpp:325-340, equations.cpp:238-249),
  and the Python Domain gives the same number.
* StepContext::physics_on = false skips registered physics: a step with a
  held theta rate but physics disabled equals the step without physics bit
  for bit, and enabled again equals the oracle with the rate.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle_bind import OracleDycore

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "vibe-model_amd", "bin", "run_warm_bubble")
FIELDS = ("u", "v", "w", "rho", "theta", "pi")
NX, NY, NZ, STEPS = 24, 20, 16, 5
KRD, KCP, KCV = 287.05, 1004.64, 717.63  # include/vibe/common/constants.hpp:17-20


def say(*a):
    print("[cpp-caller]", *a, flush=True)
    sys.stdout.flush()


def _case():
    from vibe_mi import configs
    return configs.warm_bubble(nx=NX, ny=NY, nz=NZ, dx=250.0, dz=250.0)


@pytest.fixture(scope="module")
def oracle_final():
    case = _case()
    orc = OracleDycore(case)
    assert orc.step(2.0, 6, STEPS) == 0
    st = {k: v.copy() for k, v in orc.state.items()}
    orc.close()
    return case, st


def _run(tmp_path, *extra):
    if not os.access(EXE, os.X_OK):
        raise RuntimeError(f"{EXE} is not built: run make -C vibe-model_amd")
    dump = str(tmp_path / "final.bin")
    cmd = [EXE, "--nx", str(NX), "--ny", str(NY), "--nz", str(NZ), "--steps", str(STEPS), "--dump", dump, *extra]
    say("run", " ".join(cmd[1:]))
    env = {k: v for k, v in os.environ.items() if k != "RELAX"}  # RELAX turns on the example's physics
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    say(p.stdout.strip().replace("\n", " | "))
    assert p.returncode == 0, p.stderr
    raw = np.fromfile(dump, dtype=np.float64)
    c, nxy = NX * NY * NZ, NX * NY
    sizes = {"u": c, "v": c, "w": c + nxy, "rho": c, "theta": c, "pi": c}
    assert raw.size == sum(sizes.values())
    out, o = {}, 0
    for f in FIELDS:
        out[f] = raw[o:o + sizes[f]]
        o += sizes[f]
    return p.stdout, out


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _numpy_max_stable_dt(case, st):
    nx, ny, nz = case.nx, case.ny, case.nz
    sh = (nz, ny, nx)
    u, v, th, pi = (st[f].reshape(sh) for f in ("u", "v", "theta", "pi"))
    w = st["w"].reshape(nz + 1, ny, nx)
    dzeta = np.diff(case.zeta)
    dz = np.maximum(dzeta[:, None, None] * (case.ztop - case.zs.reshape(1, ny, nx)), 1.0)
    uu = 0.5 * (u + np.roll(u, -1, axis=2))
    vv = 0.5 * (v + np.roll(v, -1, axis=1))
    ww = 0.5 * (w[:-1] + w[1:])
    cfl = np.max(np.abs(uu) / case.dx + np.abs(vv) / case.dy + np.abs(ww) / dz)
    t = (case.theta0.reshape(sh) + th) * (case.pi0.reshape(sh) + pi)
    cs = np.max(np.sqrt((KCP / KCV) * KRD * np.maximum(t, 1.0)))
    rate = np.max(cs * (1.0 / case.dx + 1.0 / case.dy + 1.0 / dz))
    return min(0.8 / cfl if cfl > 0 else 1e9, 2.0 / rate if rate > 0 else 1e9)


def test_device_loop_matches_oracle(tmp_path, oracle_final):
    case, want = oracle_final
    out, got = _run(tmp_path)
    for f in FIELDS:
        e = _rel(got[f], want[f])
        assert e <= 1e-12, (f, e)
    m = re.findall(r"max stable dt ([0-9.eE+-]+) s", out)
    assert len(m) == 2
    ref = _numpy_max_stable_dt(case, got)
    assert abs(float(m[-1]) - ref) <= 1e-6 * ref + 1e-6, (m[-1], ref)  # printed with %.6f


def test_reference_layout_loop_matches_oracle(tmp_path, oracle_final):
    case, want = oracle_final
    out, got = _run(tmp_path, "--reference-layout")
    vscale = max(float(np.linalg.norm(want[f])) for f in ("u", "v", "w"))
    for f in ("u", "v", "w"):
        e = float(np.linalg.norm(got[f] - want[f])) / vscale
        assert e <= 1e-12, (f, e)
    for f, base in (("rho", case.rho0), ("theta", case.theta0), ("pi", case.pi0)):
        e = float(np.max(np.abs(got[f] - want[f]))) / float(np.max(np.abs(base + want[f])))
        assert e <= 1e-13, (f, e)
    m = re.search(r"reference layout: (\d+) steps, (\d+) acoustic substeps, (\d+) boundary calls", out)
    assert m is not None, out
    assert tuple(int(x) for x in m.groups()) == (STEPS, 3 * 6 * STEPS, STEPS)


def test_python_max_stable_dt_and_physics_switch(vmi, oracle_final):
    case, want = oracle_final
    d = vmi.Domain(case, vmi.FP64)
    d.set_state(want)
    ref = _numpy_max_stable_dt(case, want)
    got = d.max_stable_dt()
    assert abs(got - ref) <= 1e-14 * ref, (got, ref)

    rate = np.random.default_rng(3).normal(size=case.ncell) * 1e-3
    plain = vmi.Domain(case, vmi.FP64)
    plain.set_state(want)
    plain.step(2.0, 6, 1)
    d.set_physics_tendency(theta=rate)
    d.physics_enable(False)
    d.step(2.0, 6, 1)
    a, b = d.get_state(), plain.get_state()
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), ("physics off", f)
    d.physics_enable(True)
    d.step(2.0, 6, 1)
    orc = OracleDycore(case)
    orc.state = {k: v.copy() for k, v in b.items()}
    orc.set_physics(theta=rate)
    assert orc.step(2.0, 6, 1) == 0
    g = d.get_state()
    for f in FIELDS:
        e = _rel(g[f], orc.state[f])
        assert e <= 1e-12, ("physics on", f, e)
    orc.close()
    d.close()
    plain.close()
