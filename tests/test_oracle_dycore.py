"""Pins the FP64 CPU oracle of the dynamical core with properties that the
discrete equations must satisfy (CPU only, small grids):

* a resting reference atmosphere is an exact steady state (flat, ridge and
  random terrain) -- catches any unbalanced PGF / buoyancy / metric term;
* total mass is conserved to round-off with the sponge off (flux form);
* mirror symmetry of a symmetric warm bubble is preserved -- catches an
  upwind reconstruction taken from the wrong face (the reference's
  Advection::scalar uses the right-biased WENO value of face i-1/2 at face
  i+1/2, advection.cpp:162-167);
* the HEVI substep's outputs satisfy the implicit vertical equations they
  were solved from (residual of the elimination, independent of Thomas);
* uniform flow over flat ground has zero momentum tendency; a constant
  theta' field is advected without change;
* the sponge is the backward-Euler factor 1/(1 + dt alpha) per level;
* the compensated energy totals match exact (math.fsum) sums;
* the reference state satisfies EOS and hydrostatic balance
  (tests/unit/test_dynamics.cpp:51-68 bar 5e-3);
* the step is the Wicker-Skamarock composition of the slow tendency and the
  substeps, and converges at first order in dt (the reference's
  slow-then-fast stage structure).
"""
import math

import numpy as np
import pytest

from oracle_bind import OracleDycore

RD, CP, CV, G, P0 = 287.05, 1004.64, 717.63, 9.80665, 1e5


def cfgs():
    from vibe_mi import configs
    return configs


@pytest.mark.parametrize("terrain", ["flat", "ridge", "random"])
def test_rest_state_is_exact(terrain):
    c = cfgs()
    if terrain == "flat":
        case = c.warm_bubble(nx=12, ny=10, nz=10, dz=800.0, amp=0.0)
    elif terrain == "ridge":
        case = c.mountain_wave(nx=16, ny=8, nz=12, U=0.0, noise=0.0, h=900.0, a=4000.0)
    else:
        case = c.random_terrain_case(nx=16, ny=12, nz=10)
        for k in case.state:
            case.state[k][:] = 0.0
    orc = OracleDycore(case)
    assert orc.step(case.dt, case.nsub, 3) == 0
    for k, v in orc.state.items():
        assert np.max(np.abs(v)) == 0.0, k


def test_mass_conserved_without_sponge():
    case = cfgs().random_terrain_case(nx=16, ny=12, nz=10)
    case.sponge[:] = 0.0
    orc = OracleDycore(case)
    e0 = orc.energy()
    assert orc.step(case.dt, case.nsub, 5) == 0
    e1 = orc.energy()
    assert abs(e1[0] - e0[0]) <= 1e-14 * e0[0]
    assert abs(e1[1] - e0[1]) <= 1e-14 * e0[0]


def test_mirror_symmetry_of_symmetric_bubble():
    case = cfgs().warm_bubble(nx=16, ny=16, nz=12, dz=500.0, radius=2000.0, zc=2000.0)
    orc = OracleDycore(case)
    assert orc.step(case.dt, case.nsub, 15) == 0
    nx, ny, nz = case.nx, case.ny, case.nz
    th = orc.state["theta"].reshape(nz, ny, nx)
    assert np.max(np.abs(th)) > 0.5
    assert np.max(np.abs(th - th[:, :, ::-1])) < 1e-12
    assert np.max(np.abs(th - th[:, ::-1, :])) < 1e-12
    u = orc.state["u"].reshape(nz, ny, nx)
    # u sits on west faces: mirror of face i is face nx - i (periodic)
    um = -np.roll(u[:, :, ::-1], 1, axis=2)
    assert np.max(np.abs(u)) > 1e-3 and np.max(np.abs(u - um)) < 1e-12


def test_wrong_face_upwind_breaks_symmetry():
    # the same check detects the reference's face mix-up: emulate it on a
    # 1-D WENO flux divergence with u < 0 and compare to the mirrored u > 0
    import ctypes
    from oracle_bind import oracle_lib
    def wl(q):
        l, r = ctypes.c_double(), ctypes.c_double()
        oracle_lib().orc_weno5(*q, ctypes.byref(l), ctypes.byref(r))
        return l.value, r.value
    rng = np.random.default_rng(0)
    q = rng.normal(size=12)
    n = q.size
    def div(qq, U, wrong):
        f = np.zeros(n)
        for i in range(n):  # face i between i-1 and i
            st_l = [qq[(i - 3 + m) % n] for m in range(5)]
            st_r = [qq[(i - 2 + m) % n] for m in range(5)]
            qm = wl(st_l)[0]
            qp = wl(st_l)[1] if wrong else wl(st_r)[1]
            f[i] = U * qm if U > 0 else U * qp
        return np.roll(f, -1) - f
    good = div(q, -1.0, False)
    mirror = div(q[::-1], 1.0, False)[::-1]
    assert np.max(np.abs(good - mirror)) < 1e-12
    bad = div(q, -1.0, True)
    assert np.max(np.abs(bad - mirror)) > 1e-3


def _ops(case):
    nx, ny, nz = case.nx, case.ny, case.nz
    zeta = case.zeta
    dzeta = np.diff(zeta)
    zc = 0.5 * (zeta[1:] + zeta[:-1])
    dzw = np.zeros(nz + 1)
    dzw[1:nz] = zc[1:] - zc[:-1]
    D = (case.ztop - case.zs).reshape(ny, nx)
    return nx, ny, nz, zeta, dzeta, zc, dzw, D


def test_hevi_substep_satisfies_implicit_equations():
    case = cfgs().random_terrain_case(nx=10, ny=8, nz=12)
    orc = OracleDycore(case)
    assert orc.step(case.dt, case.nsub, 2) == 0
    before = {k: v.copy() for k, v in orc.state.items()}
    dtau = 1.0
    orc.substep(dtau)
    after = orc.state
    nx, ny, nz, zeta, dzeta, zc, dzw, D = _ops(case)
    eps = case.eps_s
    an, ao = 0.5 * (1 + eps), 0.5 * (1 - eps)
    r0 = case.rho0.reshape(nz, ny, nx)
    t0 = case.theta0.reshape(nz, ny, nx)
    p0 = case.pi0.reshape(nz, ny, nx)
    kap = (RD / CV) * p0 / r0
    w0 = before["w"].reshape(nz + 1, ny, nx)
    w1 = after["w"].reshape(nz + 1, ny, nx)
    pi0_, pi1 = before["pi"].reshape(nz, ny, nx), after["pi"].reshape(nz, ny, nx)
    rh0, rh1 = before["rho"].reshape(nz, ny, nx), after["rho"].reshape(nz, ny, nx)
    # w equation at interior interfaces (independent of the elimination)
    for k in range(1, nz):
        th0w = 0.5 * (t0[k - 1] + t0[k])
        g = dtau * CP * th0w / (D * dzw[k])
        rhs = w0[k] - g * (an * (pi1[k] - pi1[k - 1]) + ao * (pi0_[k] - pi0_[k - 1]))
        assert np.max(np.abs(w1[k] - rhs)) <= 1e-12 * max(1.0, np.max(np.abs(w1)))
    # pi' and rho' move together: d pi' = kappa d rho' (acoustic EOS)
    assert np.max(np.abs((pi1 - pi0_) - kap * (rh1 - rh0))) <= 1e-13 * np.max(np.abs(pi1 - pi0_)) + 1e-20
    # column mass change equals minus the horizontal convergence only (the
    # vertical flux telescopes to zero at ground and lid)
    u1 = after["u"].reshape(nz, ny, nx)
    v1 = after["v"].reshape(nz, ny, nx)
    zs = case.zs.reshape(ny, nx)
    Du = case.ztop - 0.5 * (np.roll(zs, 1, axis=1) + zs)
    Dv = case.ztop - 0.5 * (np.roll(zs, 1, axis=0) + zs)
    r0u = 0.5 * (np.roll(r0, 1, axis=2) + r0)
    r0v = 0.5 * (np.roll(r0, 1, axis=1) + r0)
    fx = Du * r0u * u1
    fy = Dv * r0v * v1
    hd = (np.roll(fx, -1, axis=2) - fx) / case.dx + (np.roll(fy, -1, axis=1) - fy) / case.dy
    col_change = np.sum((rh1 - rh0) * dzeta[:, None, None], axis=0) * D
    col_expect = -dtau * np.sum(hd * dzeta[:, None, None], axis=0)
    assert np.max(np.abs(col_change - col_expect)) <= 1e-12 * np.max(np.abs(col_expect))


def test_uniform_flow_flat_has_zero_momentum_tendency():
    case = cfgs().warm_bubble(nx=12, ny=10, nz=8, dz=500.0, amp=0.0)
    case.state["u"][:] = 7.25
    case.state["v"][:] = -3.5
    orc = OracleDycore(case)
    R = orc.slow_tendency()
    for k in ("u", "v", "w", "rho", "theta", "pi"):
        assert np.max(np.abs(R[k])) == 0.0, k


def test_constant_theta_perturbation_is_not_changed_by_advection():
    case = cfgs().random_terrain_case(nx=12, ny=10, nz=8, bv=1e-6)
    case.theta0[:] = 300.0  # no reference-theta gradient: only advection acts
    case.state["theta"][:] = 0.75
    case.state["pi"][:] = 0.0
    case.state["rho"][:] = 0.0
    orc = OracleDycore(case)
    R = orc.slow_tendency()
    assert np.max(np.abs(R["theta"])) < 1e-13


def test_sponge_is_backward_euler_factor():
    import ctypes
    from oracle_bind import oracle_lib, OrcState, _pd
    case = cfgs().warm_bubble(nx=8, ny=8, nz=20, dz=500.0, amp=0.0, sponge_depth=4000.0)
    orc = OracleDycore(case)
    for k in orc.state:
        orc.state[k][:] = 1.0
    dt = 2.0
    oracle_lib().orc_dycore_sponge(orc.h, ctypes.byref(orc._st(orc.state)), dt)
    nz, ncol = case.nz, case.nx * case.ny
    f = 1.0 / (1.0 + dt * case.sponge)
    th = orc.state["theta"].reshape(nz, ncol)
    assert np.array_equal(th, np.repeat(np.where(case.sponge > 0, f, 1.0)[:, None], ncol, axis=1))
    w = orc.state["w"].reshape(nz + 1, ncol)
    fw = np.where(case.sponge[np.minimum(np.arange(nz + 1), nz - 1)] > 0,
                  f[np.minimum(np.arange(nz + 1), nz - 1)], 1.0)
    assert np.array_equal(w, np.repeat(fw[:, None], ncol, axis=1))
    assert case.sponge[0] == 0.0 and case.sponge[-1] > 0.15


def test_energy_totals_match_exact_sums():
    case = cfgs().random_terrain_case(nx=10, ny=8, nz=8)
    orc = OracleDycore(case)
    orc.step(case.dt, case.nsub, 1)
    e = orc.energy()
    nx, ny, nz, zeta, dzeta, zc, dzw, D = _ops(case)
    s = {k: v.reshape(-1, ny, nx) for k, v in orc.state.items()}
    zs = case.zs.reshape(ny, nx)
    H = case.ztop
    dv = case.dx * case.dy * ((H - zs) / H)[None] * dzeta[:, None, None] * H
    rho = case.rho0.reshape(nz, ny, nx) + s["rho"]
    uc = 0.5 * (s["u"] + np.roll(s["u"], -1, axis=2))
    vc = 0.5 * (s["v"] + np.roll(s["v"], -1, axis=1))
    wc = 0.5 * (s["w"][:-1] + s["w"][1:])
    T = (case.theta0.reshape(nz, ny, nx) + s["theta"]) * (case.pi0.reshape(nz, ny, nx) + s["pi"])
    z = zs[None] + zc[:, None, None] * (H - zs)[None]
    mass = math.fsum((rho * dv).ravel())
    ke = math.fsum((0.5 * rho * (uc ** 2 + vc ** 2 + wc ** 2) * dv).ravel())
    ie = math.fsum((rho * CV * T * dv).ravel())
    pe = math.fsum((rho * G * z * dv).ravel())
    for got, exp in ((e[0], mass), (e[2], ke), (e[3], ie), (e[4], pe)):
        assert abs(got - exp) <= 1e-13 * abs(exp)


@pytest.mark.parametrize("kind", [0, 1])
def test_reference_state_hydrostatic_and_eos(kind):
    c = cfgs()
    case = (c.warm_bubble(nx=4, ny=4, nz=40) if kind == 0 else
            c.mountain_wave(nx=8, ny=4, nz=40, h=1000.0, a=3000.0, noise=0.0))
    nx, ny, nz, zeta, dzeta, zc, dzw, D = _ops(case)
    r0, t0, p0 = (a.reshape(nz, ny, nx) for a in (case.rho0, case.theta0, case.pi0))
    # EOS: rho0 = p00 pi0^(cv/Rd) / (Rd theta0)
    assert np.max(np.abs(r0 * RD * t0 / (P0 * p0 ** (CV / RD)) - 1)) < 1e-14
    # hydrostatic: d pi0/dz = -g/(cp theta0) (test_dynamics.cpp:51-68, 5e-3)
    dz = (zc[1:] - zc[:-1])[:, None, None] * D[None]
    dpi = (p0[1:] - p0[:-1]) / dz
    rhs = -G / (CP * 0.5 * (t0[1:] + t0[:-1]))
    assert np.max(np.abs(dpi - rhs) / np.abs(rhs)) < 5e-3


def _compose_rk3(orc, dt, nsub, coeffs=(1 / 3, 1 / 2, 1.0), base_from_s0=True):
    """One RK3 step written out from runge_kutta.cpp:77-118 (WS2002): for
    h in {dt/3, dt/2, dt}: R = F_slow(s); s = s0 + h R (w at ground and lid
    kept); nsub substeps of h/nsub; sponge once at the end."""
    import ctypes
    from oracle_bind import oracle_lib
    c = orc.case
    nxy = c.nx * c.ny
    s0 = {k: v.copy() for k, v in orc.state.items()}
    for a in coeffs:
        h = a * dt
        R = orc.slow_tendency()
        base = s0 if base_from_s0 else {k: v.copy() for k, v in orc.state.items()}
        for k in ("u", "v", "rho", "theta", "pi"):
            orc.state[k][:] = base[k] + h * R[k]
        w = base["w"] + h * R["w"]
        w[:nxy] = base["w"][:nxy]
        w[-nxy:] = base["w"][-nxy:]
        orc.state["w"][:] = w
        for _ in range(nsub):
            orc.substep(h / nsub)
    oracle_lib().orc_dycore_sponge(orc.h, ctypes.byref(orc._st(orc.state)), dt)


def test_rk3_step_is_the_wicker_skamarock_composition():
    case = cfgs().random_terrain_case(nx=20, ny=16, nz=8)
    a = OracleDycore(case)
    a.step(case.dt, 3, 1)
    b = OracleDycore(case)
    _compose_rk3(b, case.dt, 3)
    for k in a.state:
        assert np.array_equal(a.state[k], b.state[k]), k
    # plausible mistakes change the result: classical RK3 coefficients, or
    # stages based on the previous stage instead of s^n
    for kw in ({"coeffs": (1 / 3, 2 / 3, 1.0)}, {"base_from_s0": False}):
        c2 = OracleDycore(case)
        _compose_rk3(c2, case.dt, 3, **kw)
        assert max(np.max(np.abs(a.state[k] - c2.state[k])) for k in a.state) > 1e-10, kw


def test_step_converges_at_first_order_in_dt():
    """Convergence in time of the step on a small-amplitude bubble (linear
    regime), Crank-Nicolson HEVI (off-centring 0), 4 substeps per stage, to
    t = 24 s against dt = 3/64 s: the relative error falls by the first-order
    factor 2 per halving of dt (observed rates 1.39, 1.23, 1.18 for dt = 3 ..
    0.375 s).  First order is what the reference's stage structure gives:
    s = s0 + h R_slow once, then the acoustic substeps
    (src/time/runge_kutta.cpp:86-99), a slow-then-fast split.  A last stage
    shorter than dt (a plausible coefficient slip) stops convergence."""
    import dataclasses
    c = cfgs().warm_bubble(nx=16, ny=12, nz=10, dx=500.0, dz=400.0, amp=0.01, radius=2000.0, zc=2000.0,
                           alpha_max=0.0)
    c = dataclasses.replace(c, eps_s=0.0)
    T, nsub = 24.0, 4

    def run(dt, coeffs=None):
        o = OracleDycore(c)
        n = int(round(T / dt))
        if coeffs is None:
            o.step(dt, nsub, n)
        else:
            for _ in range(n):
                _compose_rk3(o, dt, nsub, coeffs=coeffs)
        return {k: np.array(o.state[k]) for k in ("w", "rho", "theta", "pi")}

    ref = run(T / 512)

    def err(s):
        return max(np.linalg.norm(s[k] - ref[k]) / np.linalg.norm(ref[k]) for k in ref)

    e = [err(run(dt)) for dt in (1.5, 0.75, 0.375)]
    rates = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    print("errors", e, "rates", rates)
    assert all(0.9 <= r <= 1.6 for r in rates), rates
    bad = [err(run(dt, (1 / 3, 1 / 2, 0.9))) for dt in (0.75, 0.375)]
    assert math.log2(bad[0] / bad[1]) < 0.2, bad


def test_lamb_wave_frequency_is_the_discrete_one():
    """A standing Lamb wave over an isothermal reference state (closed form in
    tests/lamb_wave.py): over 60 steps (1.3 periods) the oracle's frequency is
    the discrete forward-backward C-grid value to 1e-8 relative, pi' stays
    uniform in z and w, theta' stay at second order in the amplitude."""
    import vibe_mi
    from lamb_wave import cos_amplitude, discrete_frequency, lamb_case, measured_frequency
    c = lamb_case(vibe_mi)
    o = OracleDycore(c)
    a = []
    for _ in range(c.steps):
        a.append(cos_amplitude(o.state["pi"], c))
        o.step(c.dt, c.nsub, 1)
    om, om_d = measured_frequency(a, c.dt), discrete_frequency(c)
    print("omega", om, "discrete", om_d)
    assert abs(om / om_d - 1.0) <= 1e-8, (om, om_d)
    p = np.asarray(o.state["pi"]).reshape(c.nz, c.ny, c.nx)
    assert np.ptp(p, axis=0).max() <= 1e-6 * 1e-5  # relative to the amplitude 1e-5
    assert np.abs(o.state["w"]).max() <= 1e-7 and np.abs(o.state["theta"]).max() <= 1e-7
