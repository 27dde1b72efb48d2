"""Nest lateral boundary on the GPU (vibe_mi/nest.h) against the oracle.

Both sides do the same FP64 arithmetic (bilinear lerp form, Davies weights
from the same host cosine), so the child states must match bit for bit:
full interpolation after a parent step, and Davies relaxation toward the
parent state interpolated in time between its last two steps."""
import numpy as np
import pytest

from oracle_bind import FIELDS, oracle_nest_apply

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


@pytest.mark.parametrize("prec_name", ["FP64", "MIXED"])
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
