"""Device diagnostics against values computed by the reference's own code
(tests/golden/setup_reference.json, see tests/golden/make_setup_golden.py):
the advective CFL number (Advection::cfl_number, advection.cpp:325-340) and
the stability report (TimeStepController::inspect, timestep_control.cpp:
23-56) bit for bit, and the energy budget (energy_budget, diagnostics.cpp:
255-277) within 1e-13 (the reference sums plainly, the device with
compensation)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setup_reference.json")


def _case(vmi, gold):
    nx, ny, nz = 8, 6, 16
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kw, jw, iw = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx), indexing="ij")
    st = {"u": (5.0 + 0.25 * i + 0.5 * j - 0.125 * k).reshape(-1),
          "v": (-2.0 + 0.1 * i - 0.2 * j + 0.05 * k).reshape(-1),
          "w": (0.01 * (iw - jw) + 0.002 * kw).reshape(-1),
          "rho": (1e-3 * (i - j + k) / 16.0).reshape(-1),
          "theta": ((i + 2 * j - k) / 8.0).reshape(-1),
          "pi": (1e-4 * (2 * i - j + k) / 16.0).reshape(-1)}
    return vmi.Case("golden", nx, ny, nz, 1000.0, 1500.0, 20000.0, gold["zeta_16_1_1.08"], gold["agnesi_8x6"],
                    gold["isothermal_rho0"], gold["isothermal_theta0"], gold["isothermal_pi0"], np.zeros(nz),
                    st, 6.0, 6, 1)


def test_cfl_and_energy_match_the_reference(vmi):
    with open(GOLD) as fh:
        gold = {k: np.array(v) for k, v in json.load(fh).items()}
    case = _case(vmi, gold)
    for prec in (vmi.FP64, vmi.MIXED):
        dom = vmi.Domain(case, prec)
        cfl, _ = dom.cfl(6.0)
        e = dom.energy_budget()
        rep = dom.stability(6.0)
        dom.close()
        want = gold["inspect_dt6_adv_acoustic_vertical_maxwind_mindz_stable"]
        got_rep = [rep["cfl_advection"], rep["cfl_acoustic"], rep["cfl_vertical"], rep["max_wind"], rep["min_dz"],
                   1.0 if rep["stable"] else 0.0]
        assert got_rep == list(want), (prec, got_rep, list(want))
        assert cfl == gold["cfl_dt6"][0], (prec, cfl, gold["cfl_dt6"][0])
        ref = gold["energy_mass_kinetic_internal_potential_total"]
        got = np.array([e[0], e[2], e[3], e[4], e[5]])
        assert np.all(np.abs(got - ref) <= 1e-13 * np.abs(ref)), (prec, got, ref)


def _flat_case(vmi, nx, ny, nz, dx, state):
    zeta = np.linspace(0.0, 1.0, nz + 1)
    zs = np.zeros(nx * ny)
    r0, t0, p0 = (np.empty(nx * ny * nz) for _ in range(3))
    assert vmi.lib().vibe_mi_reference_state(0, 300.0, 0.0, 1e5, nx, ny, nz, 10000.0, vmi._pd(zeta), vmi._pd(zs),
                                             vmi._pd(r0), vmi._pd(t0), vmi._pd(p0)) == 0
    return vmi.Case("nest", nx, ny, nz, dx, dx, 10000.0, zeta, zs, r0, t0, p0, np.zeros(nz), state, 1.0, 1, 1)


def test_device_nest_operators_match_reference(vmi):
    """vibe_mi_nest_interpolate / vibe_mi_nest_relax (frac 1) against the
    reference's Nest::interpolate_from_parent and apply_lateral_boundary
    (nest.cpp:52-128) on the cell-centred fields, bit for bit."""
    from test_setup_golden import _nest_states
    with open(GOLD) as fh:
        gold = {k: np.array(v) for k, v in json.load(fh).items()}
    par, chi = _nest_states()
    P = vmi.Domain(_flat_case(vmi, 16, 14, 3, 3000.0, par), vmi.FP64)
    C = vmi.Domain(_flat_case(vmi, 12, 9, 3, 1000.0, chi), vmi.FP64)
    d = vmi.NestDesc(3, 12, 15, 4, 1.0)
    vmi.nest_relax(C, P, d, 1.0)
    relaxed = C.get_state()
    vmi.nest_interpolate(C, P, d)
    interp = C.get_state()
    P.close()
    C.close()
    for f in ("rho", "theta", "pi"):
        assert np.array_equal(interp[f], gold["nest_interpolated_cell"]), f
        assert np.array_equal(relaxed[f], gold["nest_relaxed_cell_zone4"]), f
