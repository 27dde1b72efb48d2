"""Physics coupling on the GPU (vibe_mi_domain_set_physics_tendency /
_physics_buffers / _set_physics_callback; the reference's slow_tendency with
a PhysicsDriver, src/time/runge_kutta.cpp:46-63) against the oracle.

Bars: FP64 steps with held rates or a per-stage host callback agree with the
oracle to 1e-12 relative L2 per prognostic, slow tendencies to 1e-13; batched
FP64 members match the single-domain path bit for bit; FP32-mixed batches
stay within 1e-4 of FP64 with the same rates.
"""
import ctypes

import numpy as np
import pytest

from oracle_bind import OracleDycore

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _cases(cfg):
    return {
        "mountain_wave": lambda: cfg.mountain_wave(nx=32, ny=12, nz=20, dx=1000.0, a=5000.0, noise_top=3000.0),
        "random_terrain": lambda: cfg.random_terrain_case(nx=24, ny=20, nz=14),
        "warm_bubble": lambda: cfg.warm_bubble(nx=24, ny=20, nz=16, dz=500.0, radius=2500.0),
    }


def _rates(case, seed=5):
    rng = np.random.default_rng(seed)
    c = case.nx * case.ny * case.nz
    return [rng.normal(size=c) * s for s in (2e-3, 1e-2, 1e-2)]


def _relax(tau_t=600.0, tau_u=900.0):
    """Newtonian relaxation of theta' and Rayleigh friction of u, v,
    evaluated on the stage state (a physics package's per-stage role)."""
    def fn(stage, dt, st, rates):
        rates["theta"][:] = -st["theta"] / tau_t
        rates["u"][:] = -st["u"] / tau_u
        rates["v"][:] = -st["v"] / tau_u
    return fn


@pytest.mark.parametrize("name", ["mountain_wave", "random_terrain"])
def test_held_rates_match_oracle(vmi, name):
    from vibe_mi import configs
    case = _cases(configs)[name]()
    F = _rates(case)
    orc = OracleDycore(case)
    orc.set_physics(*F)
    assert orc.step(case.dt, case.nsub, 3) == 0
    dom = vmi.Domain(case, vmi.FP64)
    dom.set_physics_tendency(*F)
    dom.step(case.dt, case.nsub, 3)
    got = dom.get_state()
    R = dom.slow_tendency()
    orc_R = orc.slow_tendency()
    dom.close()
    for k in FIELDS:
        e = rel_l2(got[k], orc.state[k])
        print(name, k, e)
        assert e <= 1e-12, (name, k, e)
        assert rel_l2(R[k], orc_R[k]) <= 1e-13, (name, "R", k)


def test_host_callback_matches_oracle(vmi):
    from vibe_mi import configs
    case = _cases(configs)["warm_bubble"]()
    orc = OracleDycore(case)
    orc.set_physics_callback(_relax())
    assert orc.step(case.dt, case.nsub, 3) == 0
    calls = []
    phys = _relax()

    def fn(stage, dt, st, rates):
        calls.append((stage, dt))
        phys(stage, dt, st, rates)

    dom = vmi.Domain(case, vmi.FP64)
    dom.set_physics_callback(fn, host=True)
    dom.step(case.dt, case.nsub, 3)
    got = dom.get_state()
    dom.close()
    assert calls == [(s, case.dt) for s in (0, 1, 2)] * 3
    for k in FIELDS:
        e = rel_l2(got[k], orc.state[k])
        print("relax", k, e)
        assert e <= 1e-12, (k, e)


def test_device_callback_and_buffers(vmi):
    from vibe_mi import configs
    case = _cases(configs)["warm_bubble"]()
    plain = vmi.Domain(case, vmi.FP64)
    plain.step(case.dt, case.nsub, 2)
    ref = plain.get_state()
    plain.close()

    dom = vmi.Domain(case, vmi.FP64)
    stream = lib_stream = vmi.lib().vibe_mi_domain_stream(dom.h)
    seen = []

    def fn(stage, dt, st, rates):  # device pointers; writes no rates (they stay zero)
        seen.append((stage, rates["stream"], all(isinstance(st[k], int) and st[k] for k in st),
                     all(rates[k] for k in ("theta", "u", "v"))))

    dom.set_physics_callback(fn, host=False)
    dom.step(case.dt, case.nsub, 2)
    got = dom.get_state()
    assert [s[0] for s in seen] == [0, 1, 2, 0, 1, 2]
    assert all(s[1] == lib_stream and s[2] and s[3] for s in seen), seen[:1]
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), k
    # zeroed device buffers switch the coupling on with zero rates: no change
    dom.set_physics_callback(None)
    dom.set_state(case.state)
    p = dom.physics_buffers()
    assert all(p)
    dom.step(case.dt, case.nsub, 2)
    got2 = dom.get_state()
    dom.close()
    for k in FIELDS:
        assert np.array_equal(got2[k], ref[k]), k
    assert stream == lib_stream


def test_failing_callback_and_batch_refusal(vmi):
    from vibe_mi import configs
    case = _cases(configs)["warm_bubble"]()
    dom = vmi.Domain(case, vmi.FP64)
    dom.set_physics_callback(lambda stage, dt, st, rates: 1)
    with pytest.raises(vmi.VibeError) as e:
        dom.step(case.dt, case.nsub, 1)
    assert e.value.code == vmi.VIBE_MI_ERR_PHYSICS
    dom.close()
    b = vmi.Batch([case], vmi.FP64)
    b.members[0].set_physics_callback(_relax())
    with pytest.raises(vmi.VibeError) as e:
        b.step(None, case.nsub, 1)
    assert e.value.code == vmi.VIBE_C_ERR_BAD_ARG
    b.close()


@pytest.mark.parametrize("graph", [True, False])
def test_batched_rates_match_single(vmi, graph):
    from vibe_mi import configs
    mk = _cases(configs)
    cases = [mk["mountain_wave"](), mk["warm_bubble"]()]
    F = _rates(cases[0])
    single = vmi.Domain(cases[0], vmi.FP64)
    single.set_physics_tendency(*F)
    single.step(cases[0].dt, cases[0].nsub, 3)
    ref0 = single.get_state()
    single.close()
    other = vmi.Domain(cases[1], vmi.FP64)
    other.step(cases[1].dt, cases[1].nsub, 3)
    ref1 = other.get_state()
    other.close()
    b = vmi.Batch(cases, vmi.FP64, use_graph=graph)
    b.members[0].set_physics_tendency(*F)
    b.step(None, cases[0].nsub, 3)
    g0, g1 = b.members[0].get_state(), b.members[1].get_state()
    b.close()
    for k in FIELDS:
        assert np.array_equal(g0[k], ref0[k]), ("rates", k)
        assert np.array_equal(g1[k], ref1[k]), ("plain", k)
    # FP32-mixed batch with the same rates stays near FP64
    bm = vmi.Batch(cases[:1], vmi.MIXED, use_graph=graph)
    bm.members[0].set_physics_tendency(*F)
    bm.step(None, cases[0].nsub, 3)
    gm = bm.members[0].get_state()
    bm.close()
    for k in FIELDS:
        e = rel_l2(gm[k], ref0[k])
        assert e <= 1e-4, ("mixed", k, e)
