"""Edge sizes on the GPU against the FP64 oracle: the smallest grid the
kernels take (4x4x3), a ragged one (one column past a 64-column tile, few
rows), and the tallest FP64 column (nz = 103); FP64 steps agree to 1e-12
relative L2 per prognostic, single-domain and batched.  In FP32-mixed and
FP32-compensated, the tallest FP32 column (nz = 128), the smallest and the
ragged grid stay within 1e-4 of the oracle, single-domain and batched."""
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
        "smallest_4x4x3": lambda: cfg.warm_bubble(nx=4, ny=4, nz=3, dz=500.0, radius=1000.0),
        "ragged_65x5x7": lambda: cfg.mountain_wave(nx=65, ny=5, nz=7, dx=1000.0, a=5000.0, noise_top=3000.0),
        "tall_8x4x103": lambda: cfg.warm_bubble(nx=8, ny=4, nz=103, dz=150.0, radius=1500.0),
    }


@pytest.mark.parametrize("name", ["smallest_4x4x3", "ragged_65x5x7", "tall_8x4x103"])
def test_edge_sizes_match_oracle(vmi, name):
    from vibe_mi import configs
    case = _cases(configs)[name]()
    print(name, case.nx, case.ny, case.nz, flush=True)
    orc = OracleDycore(case)
    assert orc.step(case.dt, case.nsub, 2) == 0
    dom = vmi.Domain(case, vmi.FP64)
    dom.step(case.dt, case.nsub, 2)
    got = dom.get_state()
    dom.close()
    b = vmi.Batch([case], vmi.FP64)
    b.step(None, case.nsub, 2)
    gb = b.members[0].get_state()
    b.close()
    for k in FIELDS:
        assert rel_l2(got[k], orc.state[k]) <= 1e-12, (name, "single", k)
        assert np.array_equal(gb[k], got[k]), (name, "batched", k)


def _fp32_modes_match_oracle(vmi, case, prec_name):
    prec = getattr(vmi, prec_name)
    orc = OracleDycore(case)
    assert orc.step(case.dt, case.nsub, 2) == 0
    for batched in (False, True):
        print(prec_name, case.nx, case.ny, case.nz, "batched" if batched else "single", flush=True)
        if batched:
            b = vmi.Batch([case], prec)
            b.step(None, case.nsub, 2)
            got = b.members[0].get_state()
            b.close()
        else:
            dom = vmi.Domain(case, prec)
            dom.step(case.dt, case.nsub, 2)
            got = dom.get_state()
            dom.close()
        for k in FIELDS:
            e = rel_l2(got[k], orc.state[k])
            assert e <= 1e-4, (prec_name, batched, k, e)


@pytest.mark.parametrize("prec_name", ["MIXED", "MIXED_COMP"])
def test_tallest_fp32_column_matches_oracle(vmi, prec_name):
    from vibe_mi import configs
    _fp32_modes_match_oracle(vmi, configs.warm_bubble(nx=8, ny=4, nz=128, dz=150.0, radius=1500.0), prec_name)


@pytest.mark.parametrize("name", ["smallest_4x4x3", "ragged_65x5x7"])
@pytest.mark.parametrize("prec_name", ["MIXED", "MIXED_COMP"])
def test_edge_sizes_fp32_modes_match_oracle(vmi, name, prec_name):
    from vibe_mi import configs
    _fp32_modes_match_oracle(vmi, _cases(configs)[name](), prec_name)
