"""The five benchmark workloads at their real shapes on the GPU (through the
C ABI), plus the substep kernel instantiations the 1024x1024x96 benchmark
runs, against the FP64 oracle where the oracle finishes in seconds and
through size-independent properties where it does not.

* bench kernels at nz = 96: the ridge of the mountain-wave workload and
  random terrain, 96 stretched levels (dz 50 -> 500 m), substep workgroups
  pinned to 4 waves (24 levels per wave) and 8 waves (12 levels per wave,
  the kernel the benchmark times).  FP64 single-domain against the oracle at
  1e-12 (2 steps); FP64 batched equal to single-domain bit for bit;
  FP32-mixed and FP32-compensated batched within 1e-4 of FP64.
* warm bubble 64x64x40, dx = dz = 250 m, 150 steps of 2 s (the whole run):
  FP64 single-domain and batched against the oracle at 1e-12 per prognostic,
  mass and energy totals at 1e-12; FP32-mixed within 1e-4 of FP64.
* mountain wave 512x512x80: one FP64 step against the oracle at 1e-12.
* mountain wave 1024x1024x96 (the precision workload): 3 steps with the
  sponge off conserve mass to 1e-13 relative in FP64 and in both FP32 modes
  (observed 5e-15; the totals are Neumaier FP64 sums of the FP64 state);
  every prognostic finite; FP32-mixed and FP32-compensated within 1e-4 of
  FP64 (observed 5e-6 after 3 steps).
* ensemble of 48 uneven members (1.25e8 cells): one FP64 batched step equals
  every member stepped alone bit for bit; FP32-mixed within 1e-4 of FP64.
* nest hierarchy at full size (parent 384x384x60 and four nests): one
  parent step of the batched driver equals the hand-written schedule of
  single-domain steps and nest relaxations bit for bit (FP64).
"""
import sys

import numpy as np
import pytest

from oracle_bind import OracleDycore

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def say(*a):
    print("[configs]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def _worst_vs(got, ref):
    """Relative L2 per prognostic, velocities against the largest velocity
    component (a near-zero v would otherwise be measured against itself)."""
    vscale = max(float(np.linalg.norm(ref[f])) for f in ("u", "v", "w"))
    out = {}
    for f in FIELDS:
        den = vscale if f in ("u", "v", "w") else float(np.linalg.norm(ref[f]))
        out[f] = float(np.linalg.norm(got[f] - ref[f])) / max(den, 1e-300)
    return out


def _assert_oracle(got, orc_state, bar, what):
    for f in FIELDS:
        e = rel_l2(got[f], orc_state[f])
        assert e <= bar, (what, f, e)


def _assert_equal(a, b, what):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), (what, f, float(np.max(np.abs(a[f] - b[f]))))


def _single(vmi, case, prec, nsteps, dt=None):
    d = vmi.Domain(case, prec)
    d.step(case.dt if dt is None else dt, case.nsub, nsteps)
    st = d.get_state()
    d.close()
    return st


def _batched(vmi, cases, prec, nsteps, graph=True):
    b = vmi.Batch(cases, prec, use_graph=graph)
    b.step(None, cases[0].nsub, nsteps)
    out = [m.get_state() for m in b.members]
    tn = b.tuning()
    b.close()
    return out, tn


# ------------------------------------------------------------- nz = 96 kernels

def _nz96_cases():
    from vibe_mi import configs
    return {
        "ridge": lambda: configs.mountain_wave(nx=128, ny=24, nz=96, dx=1000.0, noise_top=1000.0),
        "random": lambda: configs.random_terrain_case(nx=64, ny=8, nz=96, ztop=20000.0),
    }


@pytest.fixture(scope="module")
def nz96(vmi):
    out = {}
    for name, mk in _nz96_cases().items():
        case = mk()
        orc = OracleDycore(case)
        assert orc.step(case.dt, case.nsub, 2) == 0
        out[name] = (case, {k: v.copy() for k, v in orc.state.items()})
        orc.close()
    return out


@pytest.mark.parametrize("name", ["ridge", "random"])
def test_nz96_fp64_single_matches_oracle(vmi, nz96, name):
    case, ref = nz96[name]
    say("nz96 single FP64", name, case.nx, case.ny, case.nz)
    _assert_oracle(_single(vmi, case, vmi.FP64, 2), ref, 1e-12, name)


@pytest.mark.parametrize("waves", [4, 8])
@pytest.mark.parametrize("name", ["ridge", "random"])
def test_nz96_batched_kernels(vmi, nz96, name, waves, monkeypatch):
    """FP64 batched == single-domain bit for bit; FP32 modes within 1e-4 of
    FP64, with the substep workgroup pinned to `waves` waves."""
    case, ref = nz96[name]
    monkeypatch.setenv("VIBE_MI_SUBSTEP_WAVES", str(waves))
    single = _single(vmi, case, vmi.FP64, 2)
    say("nz96 batched FP64", name, "waves", waves)
    (b64,), tn = _batched(vmi, [case], vmi.FP64, 2)
    assert tn["substep_waves"] == waves
    _assert_equal(b64, single, (name, waves, "FP64"))
    for prec in ("MIXED", "MIXED_COMP"):
        say("nz96 batched", prec, name, "waves", waves)
        (bm,), tn = _batched(vmi, [case], getattr(vmi, prec), 2)
        assert tn["substep_waves"] == waves
        worst = _worst_vs(bm, single)
        print(name, waves, prec, worst)
        assert max(worst.values()) <= 1e-4, (name, waves, prec, worst)


# ------------------------------------------------------------ warm bubble

@pytest.fixture(scope="module")
def bubble(vmi):
    from vibe_mi import configs
    case = configs.warm_bubble()  # 64x64x40, 250 m, 150 steps of 2 s, 6 substeps
    assert (case.nx, case.ny, case.nz, case.steps) == (64, 64, 40, 150)
    say("warm bubble oracle 150 steps")
    orc = OracleDycore(case)
    assert orc.step(case.dt, case.nsub, case.steps) == 0
    ref = {k: v.copy() for k, v in orc.state.items()}
    e = orc.energy()
    orc.close()
    return case, ref, e


def test_warm_bubble_full_run_fp64(vmi, bubble):
    case, ref, e_orc = bubble
    say("warm bubble device FP64 single")
    d = vmi.Domain(case, vmi.FP64)
    d.step(case.dt, case.nsub, case.steps)
    _assert_oracle(d.get_state(), ref, 1e-12, "single")
    e = d.energy_budget()
    d.close()
    for q in range(6):
        assert abs(e[q] - e_orc[q]) <= 1e-12 * abs(e_orc[0] if q < 2 else e_orc[5]), (q, e[q], e_orc[q])
    assert np.max(ref["theta"]) > 1.0  # the bubble is still there and has risen
    say("warm bubble device FP64 batched")
    (bs,), _ = _batched(vmi, [case], vmi.FP64, case.steps)
    _assert_oracle(bs, ref, 1e-12, "batched")


def test_warm_bubble_full_run_mixed(vmi, bubble):
    case, ref, _ = bubble
    say("warm bubble device FP32-mixed batched")
    (bm,), _ = _batched(vmi, [case], vmi.MIXED, case.steps)
    worst = _worst_vs(bm, ref)
    print("mixed vs FP64 after 150 steps", worst)
    assert max(worst.values()) <= 1e-4, worst


# ----------------------------------------------------- mountain wave 512x512x80

def test_mountain_wave_512_fp64_matches_oracle(vmi):
    from vibe_mi import configs
    case = configs.mountain_wave()  # 512x512x80, 1 km, ridge, dt 6 s
    assert (case.nx, case.ny, case.nz) == (512, 512, 80)
    say("mountain wave 512 oracle 1 step")
    orc = OracleDycore(case, threads=16)
    assert orc.step(case.dt, case.nsub, 1) == 0
    ref = orc.state
    say("mountain wave 512 device")
    d = vmi.Domain(case, vmi.FP64)
    d.step(case.dt, case.nsub, 1)
    _assert_oracle(d.get_state(), ref, 1e-12, "512x512x80")
    d.close()
    orc.close()


# ---------------------------------------------------- precision workload 1024^2

def test_precision_workload_1024_conserves_mass(vmi):
    from vibe_mi import configs
    case = configs.mountain_wave_large(alpha_max=0.0)  # sponge off
    assert (case.nx, case.ny, case.nz) == (1024, 1024, 96)
    say("1024x1024x96 FP64 batched")
    b = vmi.Batch([case], vmi.FP64)
    m = b.members[0]
    e0 = m.energy_budget()
    b.step(None, case.nsub, 3)
    e1 = m.energy_budget()
    s64 = m.get_state()
    b.close()
    assert abs(e1[0] - e0[0]) <= 1e-13 * e0[0], (e0[0], e1[0])
    for f in FIELDS:
        assert np.all(np.isfinite(s64[f])), f
    for prec in ("MIXED", "MIXED_COMP"):
        say("1024x1024x96", prec, "batched")
        b = vmi.Batch([case], getattr(vmi, prec))
        m = b.members[0]
        b.step(None, case.nsub, 3)
        e2 = m.energy_budget()
        st = m.get_state()
        b.close()
        print(prec, "mass drift", (e2[0] - e0[0]) / e0[0])
        assert abs(e2[0] - e0[0]) <= 1e-13 * e0[0], (prec, e0[0], e2[0])
        worst = _worst_vs(st, s64)
        print(prec, "vs FP64 after 3 steps", worst)
        assert max(worst.values()) <= 1e-4, (prec, worst)
        del st


# ------------------------------------------------------------ 48-member ensemble

def test_ensemble_48_batched_equals_members_alone(vmi):
    from vibe_mi import configs
    cases = configs.ensemble()
    assert len(cases) == 48
    say("ensemble", sum(c.ncell for c in cases), "cells, FP64 batched")
    got, _ = _batched(vmi, cases, vmi.FP64, 1)
    for m, c in enumerate(cases):
        _assert_equal(got[m], _single(vmi, c, vmi.FP64, 1), ("member", m))
    say("ensemble FP32-mixed batched")
    gm, _ = _batched(vmi, cases, vmi.MIXED, 1)
    for m in range(len(cases)):
        worst = _worst_vs(gm[m], got[m])
        assert max(worst.values()) <= 1e-4, (m, worst)


# ------------------------------------------------------------ nest hierarchy

def test_nest_hierarchy_full_size_parent_step(vmi):
    from vibe_mi import configs
    from vibe_mi.nesting import NestedRun
    cases, links = configs.nest_hierarchy()
    assert [(c.nx, c.ny, c.nz) for c in cases] == [(384, 384, 60), (271, 271, 60), (211, 331, 60),
                                                   (151, 151, 60), (181, 181, 60)]
    say("nest hierarchy batched driver, 1 parent step")
    run = NestedRun(cases, links, precision=vmi.FP64)
    run.step(1)
    got = [m.get_state() for m in run.members]
    run.close()

    say("nest hierarchy hand schedule")
    doms = [vmi.Domain(c, vmi.FP64) for c in cases]
    descs = {c: (p, vmi.NestDesc(r, i0, j0, 5, 1.0)) for (c, p, r, i0, j0) in links}
    level = [0] * len(cases)
    for c in sorted(descs):
        level[c] = level[descs[c][0]] + 1
    for c in sorted(descs, key=lambda q: level[q]):
        p, d = descs[c]
        vmi.nest_interpolate(doms[c], doms[p], d)
    dt0, ns = cases[0].dt, cases[0].nsub
    L = max(level)
    for s in range(3 ** L):
        for q, dom in enumerate(doms):
            if s % 3 ** (L - level[q]) == 0:
                dom.step(dt0 / 3 ** level[q], ns, 1)
        for c in sorted(descs, key=lambda q: level[q]):
            if s % 3 ** (L - level[c]) == 0:
                p, d = descs[c]
                m = (s // 3 ** (L - level[c])) % 3
                vmi.nest_relax(doms[c], doms[p], d, (m + 1) / 3)
    for q, dom in enumerate(doms):
        _assert_equal(got[q], dom.get_state(), ("domain", q))
        dom.close()
