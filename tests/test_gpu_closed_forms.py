"""The oracle's closed-form pins (tests/test_oracle_closed_forms.py) run on
the device, through the C ABI:

* gravity wave: the gravest internal gravity mode of a flat constant-N
  atmosphere advances at the eigen-frequency of the linear C-grid operator.
  FP64 single-domain and batched (hipGraph) paths give the oracle's
  frequency to 1e-12 relative; the FP32-mixed and FP32-compensated batched
  paths stay within 1e-4 of it (observed about 1e-7).
* pi' = a z over random terrain: the device slow tendency (theta' = 1) and
  one device substep leave |PX|, |PY| at round-off at interior levels (FP64).
* flow along the coordinate surfaces: the device R_theta' equals
  -(w_k Gamma_k + w_{k+1} Gamma_{k+1}) / 2 at levels 1..nz-1 (FP64).
"""
import numpy as np
import pytest

from oracle_bind import OracleDycore
from test_oracle_closed_forms import gal_chen_case, gal_chen_residuals, tangent_flow_case

pytestmark = pytest.mark.gpu

T_WAVE = 600.0
DT_WAVE = 3.0


def _device_frequency(vmi, case, mode, precision, batched):
    from gravity_wave import run_frequency
    if batched:
        b = vmi.Batch([case], precision, use_graph=True)
        m = b.members[0]
        r = run_frequency(lambda h: b.step([h], case.nsub, 1), m.get_state, case, mode, DT_WAVE, T_WAVE)
        b.close()
    else:
        d = vmi.Domain(case, precision)
        r = run_frequency(lambda h: d.step(h, case.nsub, 1), d.get_state, case, mode, DT_WAVE, T_WAVE)
        d.close()
    return r


@pytest.fixture(scope="module")
def wave(vmi):
    from gravity_wave import gw_case, run_frequency
    case, om, mode = gw_case(vmi)
    o = OracleDycore(case)
    ref = run_frequency(lambda h: o.step(h, case.nsub, 1), lambda: o.state, case, mode, DT_WAVE, T_WAVE)
    o.close()
    return case, om, mode, ref


@pytest.mark.parametrize("batched", [False, True])
def test_gravity_wave_fp64_matches_oracle(vmi, wave, batched):
    case, om, mode, ref = wave
    w, g = _device_frequency(vmi, case, mode, vmi.FP64, batched)
    print("device", w, g, "oracle", ref, "eigen", om)
    assert abs(w / ref[0] - 1.0) <= 1e-12 and abs(g - ref[1]) <= 1e-12 * om
    assert abs(w / om - 1.0) <= 1e-3


@pytest.mark.parametrize("prec", ["MIXED", "MIXED_COMP"])
def test_gravity_wave_fp32_modes(vmi, wave, prec):
    case, om, mode, ref = wave
    w, g = _device_frequency(vmi, case, mode, getattr(vmi, prec), True)
    print(prec, w, g, "oracle", ref)
    assert abs(w / ref[0] - 1.0) <= 1e-4 and abs(g - ref[1]) <= 1e-4 * om


def test_gal_chen_pgf_of_height_vanishes_on_device(vmi):
    case = gal_chen_case(vmi)
    dtau = case.dt / 3.0 / case.nsub
    d = vmi.Domain(case, vmi.FP64)
    R = d.slow_tendency()
    d.acoustic_substep(dtau)
    st = d.get_state()
    d.close()
    r_slow, r_sub = gal_chen_residuals(case, R["u"], R["v"], st["u"], st["v"], dtau)
    print("slow", r_slow, "substep", r_sub)
    assert r_slow <= 1e-10 and r_sub <= 1e-10


def test_tangent_flow_does_not_advect_theta_on_device(vmi):
    case, expect, scale = tangent_flow_case(vmi)
    d = vmi.Domain(case, vmi.FP64)
    R = d.slow_tendency()
    d.close()
    got = R["theta"].reshape(case.nz, case.ny, case.nx)[1:]
    err = np.abs(got - expect).max() / scale
    print("rel err", err)
    assert err <= 1e-10
