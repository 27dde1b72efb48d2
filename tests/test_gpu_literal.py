"""The reference's literal discrete equations on the GPU
(vibe_mi_literal_*, include/vibe_mi/literal.h) against the reference's own
compiled Equations (golden vectors tests/golden/equations_reference.npz,
written by oracle/ref_equations.cpp) and the oracle's restatement
(oracle/literal.c): tendencies of all 12 species, acoustic tendencies and
the maximum sound speed, bit for bit, in the three numerics configurations
of the fixture; then on a larger random grid (40 x 33 x 17, Agnesi terrain,
every configuration) against the oracle, bit for bit.
"""
import ctypes

import numpy as np
import pytest

from literal_cases import CASES, SPECIES, LiteralDesc, desc, golden, oracle_max_sound_speed, oracle_tendencies, ptrs, \
    shape, state
from oracle_bind import _PD

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, vmi, d):
        self.L = vmi.lib()
        self.h = ctypes.c_void_p()
        rc = self.L.vibe_mi_literal_create(ctypes.cast(ctypes.byref(d), ctypes.c_void_p), ctypes.byref(self.h))
        assert rc == 0, rc

    def tendencies(self, s, acoustic=False):
        t = [np.full(a.size, np.nan) for a in s]
        fn = self.L.vibe_mi_literal_acoustic if acoustic else self.L.vibe_mi_literal_tendencies
        assert fn(self.h, ptrs(s), ptrs(t)) == 0
        return t

    def max_sound_speed(self, s):
        out = ctypes.c_double()
        assert self.L.vibe_mi_literal_max_sound_speed(self.h, ptrs(s), ctypes.byref(out)) == 0
        return out.value

    def close(self):
        self.L.vibe_mi_literal_destroy(self.h)


@pytest.fixture(scope="module")
def gold():
    return golden()


@pytest.mark.parametrize("tag", sorted(CASES))
def test_device_literal_equals_reference(vmi, gold, tag):
    print("[literal]", tag, flush=True)
    keep = []
    d = desc(gold, tag, keep)
    s = state(gold, CASES[tag][0])
    dev = Dev(vmi, d)
    t = dev.tendencies(s)
    for q, name in enumerate(SPECIES):
        want = gold[f"{tag}_tend_{name}"]
        assert np.array_equal(t[q], want), (tag, name, float(np.nanmax(np.abs(t[q] - want))))
    a = dev.tendencies(s, acoustic=True)
    for q in range(12):
        if q in (0, 1, 2, 3, 5):
            assert np.array_equal(a[q], gold[f"{tag}_acoustic_{SPECIES[q]}"]), (tag, SPECIES[q])
        else:
            assert np.all(a[q] == 0.0)
    assert dev.max_sound_speed(s) == gold[f"{tag}_max_sound_speed"][0]
    dev.close()


def test_device_literal_equals_oracle_on_a_larger_grid(vmi):
    from vibe_mi import configs
    nx, ny, nz = 40, 33, 17
    rng = np.random.default_rng(21)
    case = configs.mountain_wave(nx=nx, ny=ny, nz=nz, dx=900.0, a=6000.0, noise_top=3000.0)
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (case.zeta, case.zs, case.rho0, case.theta0,
                                                                  case.pi0)]
    base = [np.ascontiguousarray(a) for a in keep[2:]]
    s = []
    for q in range(12):
        n = int(np.prod(shape(q, nx, ny, nz)))
        if q in (3, 4, 5):
            a = base[q - 3] * (1.0 + 0.01 * rng.normal(size=n))
        elif q < 3:
            a = rng.normal(size=n) * (10.0 if q < 2 else 1.0)
        else:
            a = rng.normal(size=n) * 1e-3
        s.append(np.ascontiguousarray(a))
    for order, cor, lat, damp, dord, moist in [(2, 1, 45.0, 0.0, 2, 1), (4, 2, 20.0, 1e5, 2, 0),
                                               (6, 0, 0.0, 5e9, 4, 1)]:
        print("[literal] large", order, cor, dord, flush=True)
        p = [a.ctypes.data_as(_PD) for a in keep]
        d = LiteralDesc(nx, ny, nz, case.dx, case.dy, case.ztop, *p, order, cor, lat, damp, dord, moist)
        dev = Dev(vmi, d)
        for acoustic in (False, True):
            t = dev.tendencies(s, acoustic)
            o = oracle_tendencies(d, s, acoustic)
            for q in range(12):
                assert np.array_equal(t[q], o[q]), (order, acoustic, SPECIES[q])
        assert dev.max_sound_speed(s) == oracle_max_sound_speed(d, s)
        dev.close()
