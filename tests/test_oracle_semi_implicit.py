"""Pins of the oracle's restatement of the reference's semi-implicit step
(orc_si_assemble / orc_si_step, oracle/literal.c; SemiImplicitIntegrator,
src/time/semi_implicit.cpp:58-226).

The reference's semi_implicit.cpp does not compile from its sources here
(common/timer.hpp:61 through logging.hpp:66, and the State::add_scaled of a
Tendency at semi_implicit.cpp:186), so there is no compiled golden: the
restatement's tendencies are the literal operators pinned bit for bit
against the reference's compiled equations.cpp (test_oracle_literal.py), and
each piece of the step is pinned here by the reference's own test
(tests/unit/test_time.cpp:297-313: a resting state stays finite over five
steps) and by what the mathematics fixes, chosen so that a plausible slip --
a wrong face index, a dropped dt/2, a flipped sign, the solve leaking into
the state, the filter on the first step, a sponge level off by one -- fails
one of them.  CPU only.
"""
import numpy as np
import pytest

from literal_cases import NX, NY, NZ, DX, desc, golden, oracle_si_assemble, oracle_si_step, oracle_tendencies, shape, \
    state

G = golden()


def _rest(grid="flat"):
    """The resting state over the fixture's reference state: rho0, theta0,
    pi0, no wind, no moisture."""
    r0, t0, p0 = (np.ascontiguousarray(G[f"{grid}_{n}"]) for n in ("rho0", "theta0", "pi0"))
    s = [np.zeros(int(np.prod(shape(q)))) for q in range(12)]
    s[3], s[4], s[5] = r0.copy(), t0.copy(), p0.copy()
    return s


def _zero_tendency():
    return [np.zeros(int(np.prod(shape(q)))) for q in range(12)]


def test_resting_state_stays_finite_over_five_steps():
    # tests/unit/test_time.cpp:297-313 (8x8x8 isothermal there; the
    # fixture's 8x6x5 grids here, both with and without terrain)
    for tag in ("c2", "c6"):
        keep = []
        d = desc(G, tag, keep)
        s = _rest("hill" if tag == "c2" else "flat")
        assert oracle_si_step(d, s, 10.0, nsteps=5) == 0
        assert all(np.all(np.isfinite(a)) for a in s)


def test_solve_does_not_reach_the_state_and_scalars_take_an_euler_step():
    # the reference applies pi_new to a copy it drops: u, v, w and pi keep
    # their bits; theta, qv, rho, qc..qg move by dt N with N the literal
    # tendencies, so the increment is linear in dt
    for tag in ("c2", "c4", "c6"):
        keep = []
        d = desc(G, tag, keep)
        s0 = state(G, "hill" if tag != "c6" else "flat")
        n = oracle_tendencies(d, s0)
        s1 = [a.copy() for a in s0]
        s2 = [a.copy() for a in s0]
        assert oracle_si_step(d, s1, 5.0) == 0
        assert oracle_si_step(d, s2, 10.0) == 0
        for q in (0, 1, 2, 5):
            assert np.array_equal(s1[q], s0[q]), (tag, q)
        for q in (3, 4, 6, 7, 8, 9, 10, 11):
            inc1, inc2 = s1[q] - s0[q], s2[q] - s0[q]
            scale = np.max(np.abs(s0[q])) + 1e-300
            assert np.max(np.abs(inc1 - 5.0 * n[q])) <= 4e-16 * scale * 4, (tag, q)
            assert np.max(np.abs(inc2 - 2.0 * inc1)) <= 1e-15 * scale * 8, (tag, q)
        if tag != "c4":  # moist configurations move the hydrometeors
            assert np.any(s1[7] != s0[7])


def test_physics_rates_enter_the_tendency_of_theta_and_qv_only_in_the_state():
    # rates on theta, qv, u, v: theta and qv move by dt * rate more; the
    # u and v rates enter N_u (the right-hand side) but not the state
    keep = []
    d = desc(G, "c2", keep)
    s0 = state(G, "hill")
    rng = np.random.default_rng(5)
    ph = [rng.normal(size=int(np.prod(shape(q)))) * 1e-3 for q in (4, 6, 0, 1)]
    a, b = [x.copy() for x in s0], [x.copy() for x in s0]
    assert oracle_si_step(d, a, 4.0) == 0
    assert oracle_si_step(d, b, 4.0, phys=ph) == 0
    assert np.allclose(b[4] - a[4], 4.0 * ph[0], rtol=0, atol=1e-12)
    assert np.allclose(b[6] - a[6], 4.0 * ph[1], rtol=0, atol=1e-15)
    for q in (0, 1, 2, 3, 5):
        assert np.array_equal(a[q], b[q])


def test_rhs_of_a_uniform_pressure_at_rest_is_the_pressure():
    # no wind, N = 0, pi uniform: G = 0 and div(a grad pi) = 0 exactly
    keep = []
    d = desc(G, "c2", keep)
    s = _rest("hill")
    s[5] = np.full(s[5].size, 0.93)
    a, rhs = oracle_si_assemble(d, s, _zero_tendency(), 6.0)
    assert np.array_equal(rhs, s[5])
    assert np.all(a > 0)


def test_rhs_of_pressure_linear_in_x_sees_only_the_clamped_edges():
    # pi = p0 + c i: the horizontal second difference vanishes inside and
    # the clamped neighbour leaves +-a c / dx^2 at the west / east columns
    keep = []
    d = desc(G, "c6", keep)
    s = _rest("flat")
    c = 2.0**-13  # exact sums: the inner second differences vanish exactly
    i = np.broadcast_to(np.arange(NX, dtype=np.float64), (NZ, NY, NX))
    s[5] = np.ascontiguousarray((0.875 + c * i).ravel())
    a, rhs = oracle_si_assemble(d, s, _zero_tendency(), 6.0)
    a, rhs, p = a.reshape(NZ, NY, NX), rhs.reshape(NZ, NY, NX), s[5].reshape(NZ, NY, NX)
    diff = rhs - p
    assert np.all(diff[:, :, 1:-1] == 0.0)
    assert np.allclose(diff[:, :, 0], a[:, :, 0] * c / DX**2, rtol=1e-9, atol=0)
    assert np.allclose(diff[:, :, -1], -a[:, :, -1] * c / DX**2, rtol=1e-9, atol=0)
    # a = (dt/2)^2 (Rd/cv) pi0 cp theta0: quadrupling dt^2 quadruples a
    a2, _ = oracle_si_assemble(d, s, _zero_tendency(), 12.0)
    assert np.allclose(a2.reshape(NZ, NY, NX), 4.0 * a, rtol=1e-15, atol=0)


def test_rhs_mass_divergence_sign_and_half_dt_of_the_tendency():
    # uniform upward wind W over rho0 falling with height: the mass flux
    # converges at inner levels (G < 0), so rhs > pi there; the same field
    # given as N_w instead enters with the factor dt/2
    keep = []
    d = desc(G, "c6", keep)
    dt, W = 8.0, 0.5
    s = _rest("flat")
    s[2] = np.full(s[2].size, W)
    _, r1 = oracle_si_assemble(d, s, _zero_tendency(), dt)
    s2 = _rest("flat")
    n = _zero_tendency()
    n[2] = np.full(n[2].size, W)
    _, r2 = oracle_si_assemble(d, s2, n, dt)
    p = s[5]
    d1, d2 = (r1 - p).reshape(NZ, NY, NX), (r2 - p).reshape(NZ, NY, NX)
    assert np.all(d1[1:-1] > 0)
    assert np.allclose(d2, 0.5 * dt * d1, rtol=1e-12, atol=0)


def test_rhs_horizontal_divergence_of_a_quadratic_wind():
    # u = U i^2 on faces i = 0..nx over rho0 uniform in x: cell i sees
    # div = U ((i+1)^2 - i^2) rho0 / dx = U (2i + 1) rho0 / dx (a face index
    # off by one shifts it by 2 U rho0 / dx), so rhs - pi = -dt c U (2i+1)
    # rho0 / dx with c = (Rd/cv) pi0 / rho0; the same in y with dy
    keep = []
    d = desc(G, "c6", keep)
    dt, U = 4.0, 0.03
    r0 = G["flat_rho0"].reshape(NZ, NY, NX)
    p0 = G["flat_pi0"].reshape(NZ, NY, NX)
    s = _rest("flat")
    i = np.broadcast_to(np.arange(NX + 1, dtype=np.float64), (NZ, NY, NX + 1))
    s[0] = np.ascontiguousarray((U * i * i).ravel())
    _, rhs = oracle_si_assemble(d, s, _zero_tendency(), dt)
    ic = np.arange(NX, dtype=np.float64)[None, None, :]
    want = -dt * (287.05 / 717.63) * (p0 / r0) * U * (2 * ic + 1) * r0 / DX
    assert np.allclose(rhs.reshape(NZ, NY, NX) - p0, want, rtol=1e-9, atol=0)
    s = _rest("flat")
    j = np.broadcast_to(np.arange(NY + 1, dtype=np.float64)[None, :, None], (NZ, NY + 1, NX))
    s[1] = np.ascontiguousarray((U * j * j).ravel())
    _, rhs = oracle_si_assemble(d, s, _zero_tendency(), dt)
    jc = np.arange(NY, dtype=np.float64)[None, :, None]
    want = -dt * (287.05 / 717.63) * (p0 / r0) * U * (2 * jc + 1) * r0 / 1500.0
    assert np.allclose(rhs.reshape(NZ, NY, NX) - p0, want, rtol=1e-9, atol=0)


def test_time_filter_skips_the_first_step_and_rounds_only():
    keep = []
    d = desc(G, "c2", keep)
    s0 = state(G, "hill")
    a, b, c = ([x.copy() for x in s0] for _ in range(3))
    assert oracle_si_step(d, a, 3.0, time_filter=0.0, first=False) == 0
    assert oracle_si_step(d, b, 3.0, time_filter=0.2, first=True) == 0
    assert oracle_si_step(d, c, 3.0, time_filter=0.2, first=False) == 0
    assert np.array_equal(a[5], b[5])                     # no filter on the first step
    assert np.max(np.abs(c[5] - a[5]) / np.abs(a[5])) <= 2.3e-16  # (1-a) pi + a pi = pi up to rounding
    assert np.any(c[5] != a[5])                           # ... and that rounding shows: the filter ran


@pytest.mark.parametrize("kmin", [NZ - 2, 0])
def test_sponge_levels_and_weights(kmin):
    keep = []
    d = desc(G, "c6", keep)
    s0 = state(G, "flat")
    rates = np.zeros(NZ)
    rates[kmin:] = np.linspace(0.01, 0.05, NZ - kmin)
    dt = 2.0
    a, b = [x.copy() for x in s0], [x.copy() for x in s0]
    assert oracle_si_step(d, a, dt) == 0
    assert oracle_si_step(d, b, dt, sponge=rates) == 0
    for q in range(12):
        sh = shape(q)
        A, B = a[q].reshape(sh), b[q].reshape(sh)
        assert np.array_equal(A[:kmin], B[:kmin]), q            # below the sponge: untouched
        if q >= 6:
            assert np.array_equal(A, B), q                      # moisture is not relaxed
    for q in (0, 1, 2):
        sh = shape(q)
        A, B = a[q].reshape(sh), b[q].reshape(sh)
        for k in range(kmin, sh[0]):
            w = 1.0 / (1.0 + dt * rates[min(k, NZ - 1)])           # w face nz takes level nz-1's rate
            assert np.array_equal(B[k], w * A[k]), (q, k)
    for q, ref in ((3, "rho0"), (4, "theta0"), (5, "pi0")):
        A, B = a[q].reshape(NZ, NY, NX), b[q].reshape(NZ, NY, NX)
        R = G[f"flat_{ref}"].reshape(NZ, NY, NX)
        for k in range(kmin, NZ):
            al = rates[k]
            assert np.allclose(B[k], (A[k] + dt * al * R[k]) / (1.0 + dt * al), rtol=1e-14, atol=0), (q, k)


def test_nonfinite_state_is_reported():
    keep = []
    d = desc(G, "c6", keep)
    s = state(G, "flat")
    s[4][7] = np.nan
    assert oracle_si_step(d, s, 1.0) == 3
