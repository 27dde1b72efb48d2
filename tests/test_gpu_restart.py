"""Restart through .vibebin files on the GPU (write_restart / read_restart
roles, src/io/field_io.cpp:207-250): an FP64 run that stops, writes a restart,
and continues in a fresh domain (or batch) from the file matches the
uninterrupted run bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def _cases(cfg):
    return {
        "warm_bubble": lambda: cfg.warm_bubble(nx=24, ny=20, nz=16, dz=500.0, radius=2500.0),
        "mountain_wave": lambda: cfg.mountain_wave(nx=32, ny=12, nz=20, dx=1000.0, a=5000.0, noise_top=3000.0),
    }


@pytest.mark.parametrize("name", ["warm_bubble", "mountain_wave"])
def test_domain_restart_continues_bit_exact(vmi, tmp_path, name):
    from vibe_mi import configs
    case = _cases(configs)[name]()
    full = vmi.Domain(case, vmi.FP64)
    full.step(case.dt, case.nsub, 5)
    ref = full.get_state()
    full.close()

    first = vmi.Domain(case, vmi.FP64)
    first.step(case.dt, case.nsub, 3)
    path = tmp_path / "restart.vibebin"
    first.write_restart(path, time=3 * case.dt, step=3)
    first.close()
    print(name, "restart written after 3 steps,", path.stat().st_size, "bytes", flush=True)

    cont = vmi.Domain(case, vmi.FP64)
    t, s = cont.read_restart(path)
    assert (t, s) == (3 * case.dt, 3)
    cont.step(case.dt, case.nsub, 2)
    got = cont.get_state()
    cont.close()
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), (name, k)


def test_batch_member_restart_continues_bit_exact(vmi, tmp_path):
    from vibe_mi import configs
    mk = _cases(configs)
    cases = [mk["warm_bubble"](), mk["mountain_wave"]()]
    full = vmi.Batch(cases, vmi.FP64, use_graph=True)
    full.step(None, cases[0].nsub, 4)
    ref = [m.get_state() for m in full.members]
    full.close()

    first = vmi.Batch(cases, vmi.FP64, use_graph=True)
    first.step(None, cases[0].nsub, 2)
    paths = [tmp_path / f"m{i}.vibebin" for i in range(2)]
    for m, p in zip(first.members, paths):
        m.write_restart(p, time=2.0, step=2)
    first.close()

    cont = vmi.Batch(cases, vmi.FP64, use_graph=True)
    for m, p in zip(cont.members, paths):
        assert m.read_restart(p) == (2.0, 2)
    cont.step(None, cases[0].nsub, 2)
    for i, m in enumerate(cont.members):
        got = m.get_state()
        for k in FIELDS:
            assert np.array_equal(got[k], ref[i][k]), (i, k)
    cont.close()
