"""The library's host setup helpers against golden vectors computed by the
reference's own code (tests/golden/setup_reference.json, written by
tests/golden/make_setup_golden.py from src/grid/geometry.cpp,
src/dyn/damping.cpp and src/dyn/reference_state.cpp): stretched vertical
coordinate, radial Witch of Agnesi, terrain smoothing, sponge profile and
the isothermal reference state, bit for bit.  Host code only (no GPU)."""
import ctypes
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setup_reference.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as fh:
        return {k: np.array(v) for k, v in json.load(fh).items()}


@pytest.fixture(scope="module")
def L():
    import vibe_mi
    return vibe_mi.lib(), vibe_mi._pd


def test_stretched_zeta(L, gold):
    lib, pd = L
    for nz, h, s, key in ((16, 1.0, 1.08, "zeta_16_1_1.08"), (40, 1.0, 1.05, "zeta_40_1_1.05")):
        z = np.empty(nz + 1)
        assert lib.vibe_mi_zeta_stretched(nz, h, s, pd(z)) == 0
        assert np.array_equal(z, gold[key]), key


def test_agnesi_bell_and_smoothing(L, gold):
    lib, pd = L
    zs = np.empty(8 * 6)
    assert lib.vibe_mi_terrain_agnesi_bell(8, 6, 1000.0, 1500.0, 800.0, 3000.0, 3000.0, 4000.0, pd(zs)) == 0
    assert np.array_equal(zs, gold["agnesi_8x6"])
    assert lib.vibe_mi_terrain_smooth(pd(zs), 8, 6, 2, 50.0) == 0
    assert np.array_equal(zs, gold["agnesi_8x6_smoothed_2_50"])


def test_sponge_profile(L, gold):
    lib, pd = L
    z = gold["zeta_16_1_1.08"].copy()
    a = np.empty(16)
    # Damping(grid, z_sponge_start = 14 km, z_top = 20 km, alpha_max = 0.2) on flat terrain
    assert lib.vibe_mi_sponge(16, 20000.0, pd(z), 0.0, 0.7, 0.2, pd(a)) == 0
    assert np.array_equal(a, gold["sponge_16_14km_0.2"])


def test_isothermal_reference_state(L, gold):
    lib, pd = L
    nx, ny, nz = 8, 6, 16
    z = gold["zeta_16_1_1.08"].copy()
    zs = gold["agnesi_8x6"].copy()
    r0, t0, p0 = (np.empty(nx * ny * nz) for _ in range(3))
    assert lib.vibe_mi_reference_state(0, 300.0, 0.0, 100000.0, nx, ny, nz, 20000.0, pd(z), pd(zs),
                                       pd(r0), pd(t0), pd(p0)) == 0
    assert np.array_equal(t0, gold["isothermal_theta0"])
    assert np.array_equal(p0, gold["isothermal_pi0"])
    assert np.array_equal(r0, gold["isothermal_rho0"])


@pytest.mark.parametrize("tag,p_sfc", [("n001", 100000.0), ("uneven", 98500.0)])
def test_from_profile_reference_state(L, gold, tag, p_sfc):
    """vibe_mi_reference_state_from_profile against the reference's own
    ReferenceState::from_profile (reference_state.cpp:61-110) on the 8x6x16
    Agnesi grid, bit for bit: a constant-N sounding and an uneven one with
    an inversion that starts above part of the terrain and ends below the
    lid.  Beside the golden vectors: each cell's pi0 drop is the hydrostatic
    increment -g dz / (cp theta0) across the cell (pi0 is stored at the
    cell's upper interface), and rho0 obeys the reference's equation of
    state."""
    lib, pd = L
    nx, ny, nz = 8, 6, 16
    z = gold["zeta_16_1_1.08"].copy()
    zs = gold["agnesi_8x6"].copy()
    zp = gold[f"profile_{tag}_z"].astype(np.float64)
    thp = gold[f"profile_{tag}_theta"].astype(np.float64)
    r0, t0, p0 = (np.empty(nx * ny * nz) for _ in range(3))
    assert lib.vibe_mi_reference_state_from_profile(nx, ny, nz, 20000.0, pd(z), pd(zs), len(zp), pd(zp), pd(thp),
                                                    p_sfc, pd(r0), pd(t0), pd(p0)) == 0
    assert np.array_equal(t0, gold[f"from_profile_{tag}_theta0"])
    assert np.array_equal(p0, gold[f"from_profile_{tag}_pi0"])
    assert np.array_equal(r0, gold[f"from_profile_{tag}_rho0"])
    RD, CP, CV, G, P0 = 287.05, 1004.64, 717.63, 9.80665, 1e5
    # EOS in the reference's form rho0 = p00 pi0^(cv/Rd) / (Rd theta0)
    # (reference_state.cpp:26-28; its cp - cv differs from Rd by 0.04)
    assert np.max(np.abs(r0 * RD * t0 / (P0 * p0 ** (CV / RD)) - 1.0)) < 1e-13
    # hydrostatic integration as stored: pi0 of cell k is pi at its upper
    # interface, so pi0(k) - pi0(k-1) = -g dz_k / (cp theta0(k)) exactly (to
    # rounding), with pi_s = (p_s/p00)^(Rd/cp) below the first cell
    pi3, th3 = p0.reshape(nz, ny, nx), t0.reshape(nz, ny, nx)
    H = 20000.0 - zs.reshape(ny, nx)
    dz = np.diff(z)[:, None, None] * H[None]
    below = np.concatenate([np.full((1, ny, nx), (p_sfc / P0) ** (RD / CP)), pi3[:-1]])
    step = -G * dz / (CP * th3)
    assert np.max(np.abs((pi3 - below) - step) / np.abs(step)) < 1e-9
    # bad arguments: fewer than two profile points
    assert lib.vibe_mi_reference_state_from_profile(nx, ny, nz, 20000.0, pd(z), pd(zs), 1, pd(zp), pd(thp),
                                                    p_sfc, pd(r0), pd(t0), pd(p0)) == 2


def test_energy_budget_definitions_match_reference(gold):
    """The oracle's compensated totals (the GPU's checker) on the state of
    oracle/ref_setup.cpp against the reference's energy_budget
    (diagnostics.cpp:255-277; plain sums, hence 1e-13 relative)."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import vibe_mi
    from oracle_bind import OracleDycore
    nx, ny, nz = 8, 6, 16
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kw, jw, iw = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx), indexing="ij")
    st = {"u": (5.0 + 0.25 * i + 0.5 * j - 0.125 * k).reshape(-1),
          "v": (-2.0 + 0.1 * i - 0.2 * j + 0.05 * k).reshape(-1),
          "w": (0.01 * (iw - jw) + 0.002 * kw).reshape(-1),
          "rho": (1e-3 * (i - j + k) / 16.0).reshape(-1),
          "theta": ((i + 2 * j - k) / 8.0).reshape(-1),
          "pi": (1e-4 * (2 * i - j + k) / 16.0).reshape(-1)}
    case = vibe_mi.Case("golden", nx, ny, nz, 1000.0, 1500.0, 20000.0, gold["zeta_16_1_1.08"],
                        gold["agnesi_8x6"], gold["isothermal_rho0"], gold["isothermal_theta0"],
                        gold["isothermal_pi0"], np.zeros(nz), st, 1.0, 1, 1)
    e = OracleDycore(case).energy()
    ref = gold["energy_mass_kinetic_internal_potential_total"]
    got = np.array([e[0], e[2], e[3], e[4], e[5]])
    assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(ref)), (got, ref)


def _nest_states():
    """Parent (16x14x3) and child (12x9x3) states of oracle/ref_setup.cpp's
    nest case: every field of the parent holds p, of the child c."""
    k, j, i = np.meshgrid(np.arange(3), np.arange(14), np.arange(16), indexing="ij")
    p = 0.5 * i - 0.25 * j + 0.125 * k + 0.01 * i * j
    k2, j2, i2 = np.meshgrid(np.arange(3), np.arange(9), np.arange(12), indexing="ij")
    c = 10.0 + 0.1 * i2 - 0.2 * j2 + k2
    def st(a, nxy):
        out = {f: a.reshape(-1).copy() for f in ("u", "v", "rho", "theta", "pi")}
        out["w"] = np.concatenate([a.reshape(-1), a[-1].reshape(-1)])
        return out
    return st(p, 16 * 14), st(c, 12 * 9)


def test_oracle_nest_operators_match_reference(gold):
    """oracle/nest.c against the reference's Nest::interpolate_from_parent
    and apply_lateral_boundary (nest.cpp:52-128) on the cell-centred fields
    (u and v sit at their own faces here, at cell centres in the reference)."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle_bind import oracle_nest_apply
    par, chi = _nest_states()
    interp = oracle_nest_apply(par, None, 16, 14, chi, 12, 9, 3, 3, 12, 15, 4, 1.0, 1.0, 1)
    relax = oracle_nest_apply(par, None, 16, 14, chi, 12, 9, 3, 3, 12, 15, 4, 1.0, 1.0, 0)
    for f in ("rho", "theta", "pi"):
        assert np.array_equal(interp[f], gold["nest_interpolated_cell"]), f
        assert np.array_equal(relax[f], gold["nest_relaxed_cell_zone4"]), f
    assert np.array_equal(interp["w"][:12 * 9 * 3], gold["nest_interpolated_cell"])


def test_davies_profile_matches_reference(L, gold):
    lib, pd = L
    for zone, am, key in ((4, 1.0, "davies_4_1"), (5, 0.8, "davies_5_0.8")):
        a = np.empty(zone)
        assert lib.vibe_mi_davies_profile(zone, am, pd(a)) == 0
        assert np.array_equal(a, gold[key]), key
