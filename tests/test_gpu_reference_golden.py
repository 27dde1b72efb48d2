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
