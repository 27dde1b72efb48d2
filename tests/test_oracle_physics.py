"""Pins of the oracle's physics coupling (orc_dycore_set_physics /
_set_physics_callback; the reference's RungeKutta3Integrator::slow_tendency
with a PhysicsDriver, src/time/runge_kutta.cpp:46-63).

  * the rates enter R linearly: R(F) - R(0) = F for u, v, theta' (to
    rounding), R_pi' moves by lambda F_theta (the linearised equation of
    state), R_rho' and R_w do not move;
  * a callback returning fixed rates reproduces the fixed-rate run bit for
    bit, is called before every stage (0, 1, 2) with the step's dt, sees the
    stage state (stage 0 = the step's initial state) and zeroed rates;
  * a failing callback ends the step with 103.
"""
import numpy as np

from oracle_bind import OracleDycore

RD, CV = 287.05, 717.63  # include/vibe/common/constants.hpp:16-26


def _case():
    from vibe_mi import configs
    return configs.mountain_wave(nx=24, ny=8, nz=12, dx=1000.0, a=5000.0, noise_top=3000.0)


def _rates(case, seed=3):
    rng = np.random.default_rng(seed)
    c = case.nx * case.ny * case.nz
    return [rng.normal(size=c) * s for s in (1e-3, 1e-2, 1e-2)]


def test_rates_enter_the_slow_tendency_linearly():
    case = _case()
    o = OracleDycore(case)
    o.step(case.dt, case.nsub, 2)
    R0 = o.slow_tendency()
    F = _rates(case)
    o.set_physics(*F)
    R1 = o.slow_tendency()
    for k, f in (("theta", F[0]), ("u", F[1]), ("v", F[2])):
        tol = 4 * np.spacing(np.abs(R0[k]) + np.abs(f))
        assert np.all(np.abs((R1[k] - R0[k]) - f) <= tol), k
    lam = (RD / CV) * np.asarray(case.pi0).reshape(-1) / np.asarray(case.theta0).reshape(-1)
    dpi = R1["pi"] - R0["pi"]
    assert np.allclose(dpi, lam * F[0], rtol=1e-9, atol=1e-14 * np.max(np.abs(R0["pi"])))
    assert np.array_equal(R1["rho"], R0["rho"]) and np.array_equal(R1["w"], R0["w"])
    o.set_physics(None, None, None)
    R2 = o.slow_tendency()
    assert all(np.array_equal(R2[k], R0[k]) for k in R0)


def test_callback_matches_fixed_rates_and_sees_each_stage():
    case = _case()
    F = _rates(case)
    fixed = OracleDycore(case)
    fixed.set_physics(*F)
    assert fixed.step(case.dt, case.nsub, 1) == 0
    after1 = {k: v.copy() for k, v in fixed.state.items()}
    assert fixed.step(case.dt, case.nsub, 1) == 0

    calls, seen = [], []

    def cb(stage, dt, st, rates):
        calls.append((stage, dt))
        seen.append(st["theta"].copy())
        assert all(np.all(rates[k] == 0.0) for k in rates)
        rates["theta"][:] = F[0]
        rates["u"][:] = F[1]
        rates["v"][:] = F[2]

    o = OracleDycore(case)
    init_theta = o.state["theta"].copy()
    o.set_physics_callback(cb)
    assert o.step(case.dt, case.nsub, 2) == 0
    assert calls == [(s, case.dt) for s in (0, 1, 2)] * 2
    assert np.array_equal(seen[0], init_theta)
    assert np.array_equal(seen[3], after1["theta"])
    assert not np.array_equal(seen[1], seen[0])  # stage 1 sees the first stage's state
    for k in fixed.state:
        assert np.array_equal(o.state[k], fixed.state[k]), k


def test_failing_callback_ends_the_step():
    case = _case()
    o = OracleDycore(case)
    o.set_physics_callback(lambda stage, dt, st, rates: 1)
    assert o.step(case.dt, case.nsub, 1) == 103
