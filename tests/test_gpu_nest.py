"""Nest lateral boundary on the GPU (vibe_mi/nest.h) against the oracle.

Both sides do the same FP64 arithmetic (bilinear lerp form, Davies weights
from the same host cosine), so the child states must match bit for bit:
full interpolation after a parent step, Davies relaxation toward the
parent state interpolated in time between its last two steps, and the
two-way feedback of a child onto the parent cells it covers."""
import numpy as np
import pytest

from oracle_bind import FIELDS, oracle_nest_apply, oracle_nest_restrict

pytestmark = pytest.mark.gpu


def _domains(vmi):
    from vibe_mi import configs
    parent_case = configs.random_terrain_case(nx=32, ny=28, nz=8)
    child_case = configs.random_terrain_case(nx=27, ny=21, nz=8, seed=3)
    return parent_case, child_case


def test_davies_profile_matches_oracle(vmi):
    from oracle_bind import oracle_lib, _pd
    for zone, amax in ((5, 1.0), (3, 0.7), (16, 1.0)):
        o = np.zeros(zone)
        oracle_lib().orc_davies_profile(zone, amax, _pd(o))
        assert np.array_equal(vmi.davies_profile(zone, amax), o)


@pytest.mark.parametrize("prec_name", ["FP64", "MIXED", "MIXED_COMP"])
def test_nest_interpolate_and_relax_match_oracle(vmi, prec_name):
    prec = getattr(vmi, prec_name)
    pc, cc = _domains(vmi)
    parent = vmi.Domain(pc, prec)
    child = vmi.Domain(cc, prec)
    desc = vmi.NestDesc(3, 14, 11, 5, 1.0)
    p_old = parent.get_state()
    parent.step(pc.dt, pc.nsub, 1)
    p_new = parent.get_state()

    vmi.nest_interpolate(child, parent, desc)
    got = child.get_state()
    want = oracle_nest_apply(p_new, None, pc.nx, pc.ny, cc.state, cc.nx, cc.ny, cc.nz, 3, 14, 11, 5, 1.0, 1.0, 1)
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f

    rng = np.random.default_rng(1)
    base = {f: v + 0.01 * rng.normal(size=v.size) for f, v in cc.state.items()}
    for frac, amax, zone in ((0.4, 1.0, 5), (1.0, 0.6, 3), (0.0, 1.0, 4)):
        child.set_state(base)
        d = vmi.NestDesc(3, 14, 11, zone, amax)
        vmi.nest_relax(child, parent, d, frac)
        got = child.get_state()
        want = oracle_nest_apply(p_new, p_old, pc.nx, pc.ny, base, cc.nx, cc.ny, cc.nz, 3, 14, 11, zone, amax,
                                 frac, 0)
        for f in FIELDS:
            assert np.array_equal(got[f], want[f]), (f, frac, zone)
    parent.close()
    child.close()


def test_nest_rejects_bad_arguments(vmi):
    pc, cc = _domains(vmi)
    parent = vmi.Domain(pc, vmi.FP64)
    child = vmi.Domain(cc, vmi.FP64)
    with pytest.raises(vmi.VibeError):
        vmi.nest_relax(child, parent, vmi.NestDesc(3, 0, 0, 17, 1.0), 0.5)  # zone above 16
    with pytest.raises(vmi.VibeError):
        vmi.nest_relax(child, parent, vmi.NestDesc(3, 0, 0, 5, 1.0), 1.5)  # frac outside [0, 1]
    parent.close()
    child.close()


@pytest.mark.parametrize("batched", [False, True])
def test_nest_restrict_matches_oracle(vmi, batched):
    """Device feedback (single domains and batch members) equals the oracle's
    bit for bit on every field; misaligned placements are rejected."""
    pc, cc = _domains(vmi)
    rng = np.random.default_rng(5)
    child_state = {f: v + 0.01 * rng.normal(size=v.size) for f, v in cc.state.items()}
    desc = vmi.NestDesc(3, 15, 12, 5, 1.0)
    if batched:
        b = vmi.Batch([pc, cc], vmi.FP64)
        parent, child = b.members
        child.set_state(child_state)
        vmi.batch_nest_restrict(b, 1, 0, desc)
    else:
        parent, child = vmi.Domain(pc, vmi.FP64), vmi.Domain(cc, vmi.FP64)
        child.set_state(child_state)
        vmi.nest_restrict(child, parent, desc)
    got = parent.get_state()
    want = oracle_nest_restrict(child_state, cc.nx, cc.ny, pc.state, pc.nx, pc.ny, pc.nz, 3, 15, 12)
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f
    assert not np.array_equal(got["theta"], pc.state["theta"])
    with pytest.raises(vmi.VibeError):
        vmi.nest_restrict(child, parent, vmi.NestDesc(3, 14, 12, 5, 1.0))  # i0 not a multiple of r
    if batched:
        b.close()
    else:
        parent.close()
        child.close()


def test_two_way_nested_run_matches_hand_schedule(vmi):
    """NestedRun(two_way=True): one parent step of a parent and one ratio-3
    nest equals the hand-written schedule (3 nest sub-steps with Davies
    relaxations, then the feedback) bit for bit in FP64."""
    from vibe_mi import configs
    from vibe_mi.nesting import NestedRun
    cases, links = configs.nest_hierarchy(pnx=48, pny=42, nz=8, dx=3000.0, nests=[(27, 24, 0, (6, 6))], steps=2)
    run = NestedRun(cases, links, precision=vmi.FP64, two_way=True)
    run.step(2)
    got = [m.get_state() for m in run.members]
    run.close()
    P, A = (vmi.Domain(c, vmi.FP64) for c in cases)
    d = vmi.NestDesc(3, links[0][3], links[0][4], 5, 1.0)
    vmi.nest_interpolate(A, P, d)
    dt0, ns = cases[0].dt, cases[0].nsub
    for _ in range(2):
        P.step(dt0, ns, 1)
        for s in range(3):
            A.step(dt0 / 3, ns, 1)
            vmi.nest_relax(A, P, d, (s + 1) / 3)
        vmi.nest_restrict(A, P, d)
    want = [P.get_state(), A.get_state()]
    P.close()
    A.close()
    for m in range(2):
        for f in FIELDS:
            assert np.array_equal(got[m][f], want[m][f]), (m, f)
