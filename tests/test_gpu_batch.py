"""Batched step (vibe_mi/batch.h) against each member stepped alone.

The batched kernels run the same tile bodies as the single-domain kernels on
the same operands, so in FP64 every member must match its single-domain run
bit for bit, with and without hipGraph replay, for any active subset and
per-member dt.  The FP32 batched kernels are compiled with multiply-add
contraction (the single-domain FP32 kernels are not), so in FP32-mixed and
FP32-compensated the members agree to FP32 rounding, checked at the
mixed-precision bar of 1e-4
relative L2 (observed up to 1.4e-5 for pi' after three steps), the velocity
components measured against the largest of them (v alone is a near-zero
field in these cases); graph replay matches the eager batched launches bit
for bit.  Members differ in size, depth and
geometry compaction (flat: 1-D reference arrays, ridge: (x, z), random
terrain: full 3-D)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def _cases(vmi):
    from vibe_mi import configs
    return [
        configs.ensemble_member(40, 24, 12, seed=11, dx=2000.0, dt=8.0),
        configs.mountain_wave(nx=24, ny=36, nz=16, noise=0.1),
        configs.random_terrain_case(nx=32, ny=20, nz=10),
    ]


def _single(vmi, case, prec, dts, nsub):
    d = vmi.Domain(case, prec)
    for dt in dts:
        d.step(dt, nsub, 1)
    st = d.get_state()
    d.close()
    return st


def _equal(a, b):
    return all(np.array_equal(a[f], b[f]) for f in FIELDS)


def _assert_match(prec_name, got, ref, what):
    """Bit-exact in FP64; FP32-mixed within the mixed-precision bar of 1e-4
    relative L2 (contracted FMAs round differently), velocities measured
    against the largest velocity component."""
    vscale = max(float(np.linalg.norm(ref[f])) for f in ("u", "v", "w"))
    for f in FIELDS:
        if prec_name == "FP64":
            assert np.array_equal(got[f], ref[f]), (what, f, float(np.max(np.abs(got[f] - ref[f]))))
        else:
            den = vscale if f in ("u", "v", "w") else float(np.linalg.norm(ref[f]))
            rel = float(np.linalg.norm(got[f] - ref[f])) / max(den, 1e-30)
            assert rel < 1e-4, (what, f, rel)


@pytest.mark.parametrize("prec_name", ["FP64", "MIXED", "MIXED_COMP"])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single(vmi, prec_name, graph):
    prec = getattr(vmi, prec_name)
    cases = _cases(vmi)
    dts = [c.dt for c in cases]
    b = vmi.Batch(cases, prec, use_graph=graph)
    for _ in range(3):
        b.step(dts, nsub=6, nsteps=1)
    got = [m.get_state() for m in b.members]
    b.close()
    for c, g in zip(cases, got):
        ref = _single(vmi, c, prec, [c.dt] * 3, 6)
        _assert_match(prec_name, g, ref, (c.name, graph))


@pytest.mark.parametrize("waves,rows", [(4, 2), (8, 4), (16, 8), (8, 8)])
def test_batch_tile_shapes_match_single(vmi, monkeypatch, waves, rows):
    """Every substep wave count and stage tile height the autotuner can pick
    gives the single-domain result bit for bit in FP64 (the shapes are pinned
    through the environment, which bypasses the tuner and its cache)."""
    monkeypatch.setenv("VIBE_MI_SUBSTEP_WAVES", str(waves))
    monkeypatch.setenv("VIBE_MI_STAGE_ROWS", str(rows))
    cases = _cases(vmi)
    b = vmi.Batch(cases, vmi.FP64, use_graph=False)
    b.step(None, nsub=6, nsteps=2)
    assert b.tuning() == {"substep_waves": waves, "stage_rows": rows}
    got = [m.get_state() for m in b.members]
    b.close()
    for c, g in zip(cases, got):
        ref = _single(vmi, c, vmi.FP64, [c.dt] * 2, 6)
        _assert_match("FP64", g, ref, (c.name, waves, rows))


def test_batch_graph_replays_match_eager(vmi):
    """Several graph replays, with host work and synchronisation between
    them, each equal to the eager launch sequence."""
    cases = _cases(vmi)
    bg = vmi.Batch(cases, vmi.MIXED, use_graph=True)
    be = vmi.Batch(cases, vmi.MIXED, use_graph=False)
    for rep in range(4):
        bg.step(None, nsub=6, nsteps=1 + rep % 2)
        be.step(None, nsub=6, nsteps=1 + rep % 2)
        for mg, me in zip(bg.members, be.members):
            assert _equal(mg.get_state(), me.get_state()), rep
        e = bg.members[0].energy_budget()
        assert np.all(np.isfinite(e))
    bg.close()
    be.close()


def test_batch_active_subset_and_dt(vmi):
    cases = _cases(vmi)
    b = vmi.Batch(cases, vmi.FP64)
    dts = [4.0, 6.0, 3.0]
    before = b.members[1].get_state()
    b.step(dts, nsub=4, nsteps=2, active=[1, 0, 1])
    after = [m.get_state() for m in b.members]
    assert _equal(after[1], before)  # inactive member untouched
    for m in (0, 2):
        ref = _single(vmi, cases[m], vmi.FP64, [dts[m]] * 2, 4)
        assert _equal(after[m], ref), m
    # then all members together, continuing from there
    b.step(dts, nsub=4, nsteps=1)
    ref1 = _single(vmi, cases[1], vmi.FP64, [dts[1]], 4)
    assert _equal(b.members[1].get_state(), ref1)
    tn = b.tuning()
    assert tn["substep_waves"] in (4, 8, 16) and tn["stage_rows"] in (2, 4, 8)
    b.close()


def test_batch_member_state_roundtrip_and_energy(vmi):
    cases = _cases(vmi)
    b = vmi.Batch(cases, vmi.MIXED)
    for m, c in zip(b.members, cases):
        st = m.get_state()
        for f in FIELDS:
            assert np.array_equal(st[f], c.state[f])
    b.step(None, nsub=6, nsteps=2)
    for m, c in zip(b.members, cases):
        d = vmi.Domain(c, vmi.MIXED)
        d.step(c.dt, 6, 2)
        e1, e2 = m.energy_budget(), d.energy_budget()
        # FP32 rounding of the contracted kernels: totals agree to 1e-7; the
        # perturbation mass (a near-cancelling sum) to 1e-7 of the total mass
        scale = np.abs(e2)
        scale[1] = abs(e2[0])
        assert np.all(np.abs(e1 - e2) <= 1e-7 * scale), (c.name, e1, e2)
        d.close()
    assert b.device_bytes() > 0
    b.close()


@pytest.mark.parametrize("kind", ["flat", "ridge"])
@pytest.mark.parametrize("prec_name", ["FP64", "MIXED", "MIXED_COMP"])
def test_batch_geometry_specialisations_match_single(vmi, kind, prec_name):
    """Batches whose members are all flat (or all y-uniform ridges) run the
    kernels specialised for that geometry (zero slope and metric terms
    skipped); results equal the general single-domain kernels (bit for bit
    in FP64)."""
    from vibe_mi import configs
    prec = getattr(vmi, prec_name)
    if kind == "flat":
        cases = [configs.ensemble_member(40, 24, 12, seed=21), configs.ensemble_member(24, 36, 10, seed=22),
                 configs.warm_bubble(nx=20, ny=20, nz=12)]
    else:
        cases = [configs.mountain_wave(nx=36, ny=12, nz=14, noise=0.1),
                 configs.mountain_wave(nx=24, ny=20, nz=10, noise=0.2, seed=5)]
    b = vmi.Batch(cases, prec)
    b.step(None, nsub=6, nsteps=3)
    got = [m.get_state() for m in b.members]
    b.close()
    for c, g in zip(cases, got):
        ref = _single(vmi, c, prec, [c.dt] * 3, 6)
        _assert_match(prec_name, g, ref, (kind, c.name))
