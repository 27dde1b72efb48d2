"""Closed-form pins of the oracle's buoyancy, theta0-gradient, Gal-Chen metric
and Omega terms (CPU only).  Each test also runs the oracle with one planted
mistake (orc_plant, oracle/oracle.h) and shows that its check fails under it:

* gravity wave (tests/gravity_wave.py): in a flat constant-N atmosphere the
  gravest internal gravity mode, started from the eigenvector of the linear
  C-grid operator, advances at that operator's eigen-frequency (Richardson
  extrapolation in dt) and does not grow.  Pins the buoyancy g theta'/theta0
  (oracle/dycore.c R_w) and the -(w dtheta0/dz) term of R_theta' (gam);
  flipped buoyancy grows the mode, a halved gradient moves omega by 20 %.
* pressure along constant height: pi' = a z(x, y, zeta) has no horizontal
  gradient at constant z, so the Gal-Chen PGF (px_op / py_op) vanishes to
  round-off at interior levels over random terrain, in the slow tendency and
  in the substep's momentum update; a flipped metric sign doubles it.
* flow along the coordinate surfaces: w = u dz/dx|zeta + v dz/dy|zeta =
  (1 - zeta)(u dzs/dx + v dzs/dy) crosses no zeta surface, so Omega = 0 and a
  theta' that depends on the level only is not advected: R_theta' is exactly
  -(w_k Gamma_k + w_{k+1} Gamma_{k+1}) / 2.  Pins compute_omega; a flipped
  slope-flux sign or zeta in place of (1 - zeta) advects theta'.
"""
import math

import numpy as np
import pytest

from oracle_bind import OracleDycore, planted

RD, CP, CV, G = 287.05, 1004.64, 717.63, 9.80665


@pytest.fixture(scope="module")
def vm():
    """The package (host-side setup helpers only: no GPU is used here)."""
    import vibe_mi
    return vibe_mi


# ---------------------------------------------------------------- gravity wave

def _oracle_frequency(case, mode, dt, T, plant=0):
    from gravity_wave import run_frequency
    with planted(plant):
        o = OracleDycore(case)
        r = run_frequency(lambda h: o.step(h, case.nsub, 1), lambda: o.state, case, mode, dt, T)
        o.close()
    return r


def test_gravity_wave_frequency_is_the_linear_operator_eigenvalue(vm):
    """16x4x12 cells, dx 1 km, H 6 km, N = 0.01 1/s, theta' 1e-3 K, 600 s (0.56
    of a period).  Observed: omega/omega_eig - 1 = 3.8e-4 (dt 3 s) and 2.1e-4
    (dt 1.5 s), first order; extrapolated 2.8e-5.  Growth rate 2.0e-5 1/s at
    dt 1.5 s, halving with dt.  The continuous anelastic estimate is 1.1 %
    above omega_eig (discrete wavenumbers and compressibility)."""
    from gravity_wave import anelastic_frequency, gw_case
    case, om, mode = gw_case(vm)
    assert abs(anelastic_frequency(case.nx, case.dx, case.ztop, 0.01) / om - 1.0) <= 0.03
    T = 600.0
    w3, g3 = _oracle_frequency(case, mode, 3.0, T)
    w15, g15 = _oracle_frequency(case, mode, 1.5, T)
    print("omega_eig", om, "dt 3", w3, g3, "dt 1.5", w15, g15)
    assert abs(w15 / om - 1.0) <= 5e-4 and abs(w3 / om - 1.0) <= 1e-3
    assert abs((2.0 * w15 - w3) / om - 1.0) <= 1e-4  # first-order Richardson
    assert abs(g15) <= 0.01 * om and abs(g15) < abs(g3)


@pytest.mark.parametrize("plant,what", [(planted.BUOYANCY_SIGN, "buoyancy sign"),
                                        (planted.GAMMA_HALF, "dtheta0/dz halved")])
def test_gravity_wave_check_fails_under_planted_mistake(vm, plant, what):
    from gravity_wave import gw_case
    case, om, mode = gw_case(vm)
    w3, g3 = _oracle_frequency(case, mode, 3.0, 600.0, plant)
    w15, g15 = _oracle_frequency(case, mode, 1.5, 600.0, plant)
    print(what, w3, w15, g15)
    ext_ok = abs((2.0 * w15 - w3) / om - 1.0) <= 1e-4
    grow_ok = abs(g15) <= 0.01 * om
    assert not (ext_ok and grow_ok), what


# ------------------------------------------------- pressure along constant height

def _pi_of_height(case, a=1e-7):
    nx, ny, nz = case.nx, case.ny, case.nz
    zc = 0.5 * (case.zeta[1:] + case.zeta[:-1])
    zs = case.zs.reshape(ny, nx)
    z = zs[None] + zc[:, None, None] * (case.ztop - zs[None])
    return a * z.reshape(-1), a


def _slopes(case):
    nx, ny = case.nx, case.ny
    zs = case.zs.reshape(ny, nx)
    sx = (zs - np.roll(zs, 1, axis=1)) / case.dx
    sy = (zs - np.roll(zs, 1, axis=0)) / case.dy
    return sx, sy


def gal_chen_case(vm):
    """Random terrain, at rest, theta' = 1 (so R_u = -cp PX), pi' = a z."""
    from vibe_mi import configs
    case = configs.random_terrain_case(nx=20, ny=16, nz=12, sigma=400.0, L=12000.0)
    st = {k: np.zeros_like(v) for k, v in case.state.items()}
    st["pi"], _a = _pi_of_height(case)
    st["theta"][:] = 1.0
    case.state = st
    return case


def _interior(x, case):
    return np.asarray(x).reshape(case.nz, case.ny, case.nx)[1:case.nz - 1]


def gal_chen_residuals(case, R_u, R_v, u_sub, v_sub, dtau):
    """Max |PX|, |PY| at interior levels relative to one metric term, from the
    slow tendency (theta' = 1) and from the substep's momentum update."""
    t0 = case.theta0.reshape(case.nz, case.ny, case.nx)[1:case.nz - 1].max()
    sx, sy = _slopes(case)
    a_term = CP * 1e-7 * max(np.abs(sx).max(), np.abs(sy).max())
    r_slow = max(np.abs(_interior(R_u, case)).max(), np.abs(_interior(R_v, case)).max()) / a_term
    r_sub = max(np.abs(_interior(u_sub, case)).max(), np.abs(_interior(v_sub, case)).max()) / (dtau * t0 * a_term)
    return r_slow, r_sub


def _oracle_gal_chen(vm, plant=0):
    case = gal_chen_case(vm)
    dtau = case.dt / 3.0 / case.nsub
    with planted(plant):
        o = OracleDycore(case)
        R = o.slow_tendency()
        o.substep(dtau)
        st = o.state
        o.close()
    return gal_chen_residuals(case, R["u"], R["v"], st["u"], st["v"], dtau)


def test_pressure_of_height_has_no_horizontal_gradient_over_terrain(vm):
    """Random terrain 20x16x12 (slopes up to 0.1): |PX|, |PY| / (cp a |s|) at
    levels 1..nz-2 is at round-off (observed about 1e-12) in both the slow
    tendency and the substep; at levels 0 and nz-1 the one-sided vertical
    gradient leaves a boundary term, as it must."""
    r_slow, r_sub = _oracle_gal_chen(vm)
    print("slow", r_slow, "substep", r_sub)
    assert r_slow <= 1e-10 and r_sub <= 1e-10


def test_gal_chen_check_fails_with_flipped_metric_sign(vm):
    r_slow, r_sub = _oracle_gal_chen(vm, planted.METRIC_SIGN)
    print("planted", r_slow, r_sub)
    assert r_slow > 0.1 and r_sub > 0.1


# ------------------------------------------------ flow along coordinate surfaces

def tangent_flow_case(vm):
    """Random terrain, level-dependent winds U(k), V(k), w from the flow
    along the coordinate surfaces, theta'(k); pi' = rho' = 0.  Returns the
    case and the expected R_theta' (levels 1..nz-1)."""
    from vibe_mi import configs
    case = configs.random_terrain_case(nx=20, ny=16, nz=12, sigma=400.0, L=12000.0)
    nx, ny, nz = case.nx, case.ny, case.nz
    kk = np.arange(nz)
    U = 5.0 + 0.3 * kk
    V = -2.0 + 0.2 * kk
    th = 0.3 * kk - 1.0
    sx, sy = _slopes(case)
    # u dzs/dx at cell centres from the two faces, likewise v
    mcu = 0.5 * (sx + np.roll(sx, -1, axis=1))
    mcv = 0.5 * (sy + np.roll(sy, -1, axis=0))
    mc = U[:, None, None] * mcu[None] + V[:, None, None] * mcv[None]  # [nz][ny][nx]
    w = np.zeros((nz + 1, ny, nx))
    for k in range(1, nz):
        w[k] = (1.0 - case.zeta[k]) * 0.5 * (mc[k - 1] + mc[k])  # dz/dt along zeta = const
    w[0] = mc[0]  # kinematic lower boundary
    ones = np.ones((nz, ny, nx))
    case.state = {"u": (U[:, None, None] * ones).reshape(-1), "v": (V[:, None, None] * ones).reshape(-1),
                  "w": w.reshape(-1), "rho": np.zeros(nx * ny * nz), "theta": (th[:, None, None] * ones).reshape(-1),
                  "pi": np.zeros(nx * ny * nz)}
    # dtheta0/dz at interior interfaces, physical height differences
    t0 = case.theta0.reshape(nz, ny, nx)
    zs = case.zs.reshape(ny, nx)
    zc = 0.5 * (case.zeta[1:] + case.zeta[:-1])
    zcell = zs[None] + zc[:, None, None] * (case.ztop - zs[None])
    gam = np.zeros((nz + 1, ny, nx))
    gam[1:nz] = (t0[1:] - t0[:-1]) / (zcell[1:] - zcell[:-1])
    expect = -0.5 * (w[:nz] * gam[:nz] + w[1:] * gam[1:])
    scale = np.abs(expect[1:]).max()
    return case, expect[1:], scale


def _oracle_tangent(vm, plant=0):
    case, expect, scale = tangent_flow_case(vm)
    with planted(plant):
        o = OracleDycore(case)
        R = o.slow_tendency()
        o.close()
    got = R["theta"].reshape(case.nz, case.ny, case.nx)[1:]
    return np.abs(got - expect).max() / scale


def test_flow_along_coordinate_surfaces_does_not_advect_theta(vm):
    """R_theta' equals -(w_k Gamma_k + w_{k+1} Gamma_{k+1})/2 at levels 1..nz-1
    to 1e-10 of its size (observed about 1e-14): Omega is zero for the flow
    along the zeta surfaces, and the flux form cancels the horizontal
    divergence of a horizontally uniform theta'.  (Level 0 uses the ground
    interface's Gamma, a convention, and is left out.)"""
    err = _oracle_tangent(vm)
    print("rel err", err)
    assert err <= 1e-10


@pytest.mark.parametrize("plant", [planted.OMEGA_SIGN, planted.OMEGA_ZETA])
def test_tangent_flow_check_fails_with_planted_omega_mistake(vm, plant):
    err = _oracle_tangent(vm, plant)
    print("planted", plant, err)
    assert err > 1e-3
