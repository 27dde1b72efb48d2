"""GPU checks of the compensated FP32-mixed precision (VIBE_MI_MIXED_COMP).

The mode keeps u in the FP32 substep state as an FP32 value plus an 8-bit
remainder, adds the substeps' FP32 increments of u to it with an exact
two-sum, and forms the x part of the substep's horizontal mass-flux
divergence in FP64.
Bars: after 30 steps on the small flat, ridge and random-terrain cases
every prognostic is within 3e-5 relative L2 of the FP64 path (the plain
FP32-mixed path is held to 1e-4; at these lengths the FP32 stage arithmetic,
common to both modes, is a large part of what remains) and no field is more
than 10 % further from it than in the plain FP32-mixed path; mass agrees with
FP64 to 1e-10 relative.  Batched members are checked against single domains in
test_gpu_batch.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _cases(cfg):
    return {
        "warm_bubble": lambda: cfg.warm_bubble(nx=24, ny=20, nz=16, dz=500.0, radius=2500.0),
        "mountain_wave": lambda: cfg.mountain_wave(nx=32, ny=12, nz=20, dx=1000.0, a=5000.0, noise_top=3000.0),
        "random_terrain": lambda: cfg.random_terrain_case(nx=24, ny=20, nz=14),
    }


@pytest.mark.parametrize("name", ["warm_bubble", "mountain_wave", "random_terrain"])
def test_compensated_close_to_fp64(vmi, name):
    from vibe_mi import configs
    case = _cases(configs)[name]()
    runs = {}
    for prec in (vmi.FP64, vmi.MIXED, vmi.MIXED_COMP):
        d = vmi.Domain(case, prec)
        d.step(case.dt, case.nsub, 30)
        runs[prec] = (d.get_state(), d.energy_budget())
        d.close()
    ref, e64 = runs[vmi.FP64]
    for k in FIELDS:
        em = rel_l2(runs[vmi.MIXED][0][k], ref[k])
        ec = rel_l2(runs[vmi.MIXED_COMP][0][k], ref[k])
        print(name, k, "mixed", em, "comp", ec)
        assert np.all(np.isfinite(runs[vmi.MIXED_COMP][0][k]))
        assert ec <= 3e-5, (k, ec)
        assert ec <= 1.1 * em, (k, ec, em)
    ec_mass = runs[vmi.MIXED_COMP][1][0]
    assert abs(ec_mass - e64[0]) <= 1e-10 * abs(e64[0])
