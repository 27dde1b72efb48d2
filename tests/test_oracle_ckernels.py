"""Pins the oracle's restatement of the reference C kernels (CPU only).

Checks: golden vectors written from the reference's own c_kernels.c
(tests/golden/), the reference's known-answer unit tests
(tests/unit/test_common.cpp:163-266), the compiled reference itself when
oracle/_ref is present, and mathematics (closed-form Thomas solutions, dense
solves, the linear-weight limit of WENO5, exact rational sums).
"""
import ctypes
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle_bind import oracle_lib, ref_lib, _pd

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def weno(q):
    l, r = ctypes.c_double(), ctypes.c_double()
    oracle_lib().orc_weno5(*q, ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value


def thomas(a, b, c, d):
    n = len(b)
    x, ws = np.zeros(n), np.zeros(n)
    a = None if a is None else np.ascontiguousarray(a, float)
    c = None if c is None else np.ascontiguousarray(c, float)
    rc = oracle_lib().orc_tridiagonal_solve(_pd(a), _pd(np.ascontiguousarray(b, float)), _pd(c),
                                            _pd(np.ascontiguousarray(d, float)), n, _pd(x), _pd(ws))
    return rc, x


def test_weno5_golden():
    data = json.load(open(os.path.join(GOLD, "weno5.json")))
    assert data["cases"][0]["q_left"] == 1.1306425090328729  # the fixed probe stencil
    for cs in data["cases"]:
        assert weno(cs["q"]) == (cs["q_left"], cs["q_right"])


def test_weno5_constant_and_linear_limit():
    # exact on constants (test_common.cpp:187-192)
    assert weno([3.5] * 5) == (3.5, 3.5)
    # for a quadratic the three smoothness indicators are equal, so the
    # weights are the linear ones and q- is the 5th-order linear stencil
    x = np.arange(-2, 3, dtype=float)
    q = 0.3 * x * x - 1.7 * x + 2.0
    lin = (2 * q[0] - 13 * q[1] + 47 * q[2] + 27 * q[3] - 3 * q[4]) / 60.0
    assert abs(weno(q)[0] - lin) < 1e-13
    # mirror property: q+ of (a..e) is q- of (e..a)
    rng = np.random.default_rng(0)
    for _ in range(50):
        q = rng.normal(size=5)
        assert weno(q)[1] == weno(q[::-1])[0]


def test_weno5_plausible_mistake_fails():
    # swapping the ideal weights d0 <-> d2 (the variant in the reference's
    # kernels_cpu.cpp:109-128) gives 1.2679..., not the C ABI's value
    assert abs(weno([0.2, -0.4, 1.1, 0.9, 2.3])[0] - 1.267931829545877) > 0.1


@pytest.mark.parametrize("n", [1, 2, 3, 24, 96])
def test_thomas_dense_solve(n):
    rng = np.random.default_rng(n)
    a, c = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - 1)
    b, d = 3 + rng.uniform(0, 1, n), rng.normal(size=n)
    rc, x = thomas(a if n > 1 else None, b, c if n > 1 else None, d)
    A = np.diag(b) + np.diag(a, -1) + np.diag(c, 1)
    assert rc == 0
    assert np.max(np.abs(x - np.linalg.solve(A, d))) <= 1e-13 * np.max(np.abs(x))


def test_thomas_reference_unit_tests():
    # test_common.cpp:163-175: diag(1,-4,1), residual 1e-10 -- our oracle passes
    n = 24
    rc, x = thomas(np.ones(n - 1), -4 * np.ones(n), np.ones(n - 1), np.ones(n))
    lhs = np.concatenate(([0], x[:-1])) - 4 * x + np.concatenate((x[1:], [0]))
    assert rc == 0 and np.max(np.abs(lhs - 1)) < 1e-14
    # ... and the reference's own C implementation does not (its forward sweep
    # uses c[k] in place of c[k-1]); recorded in the golden fixture
    g = json.load(open(os.path.join(GOLD, "thomas_reference_defect.json")))
    xr = np.array(g["x_reference"])
    lhs_r = np.concatenate(([0], xr[:-1])) - 4 * xr + np.concatenate((xr[1:], [0]))
    assert np.max(np.abs(lhs_r - 1)) > 1e-2
    # return codes, test_common.cpp:177-185
    assert thomas(None, [0.0, 0.0], None, [1.0, 1.0])[0] == 1
    L = oracle_lib()
    x, ws = np.zeros(2), np.zeros(2)
    assert L.orc_tridiagonal_solve(None, None, None, None, 2, _pd(x), _pd(ws)) == 2
    assert L.orc_tridiagonal_solve(None, _pd(np.zeros(2)), None, _pd(np.ones(2)), 0, _pd(x), _pd(ws)) == 2


def test_thomas_sine_closed_form():
    # (1,-4,1) with x_j = sin(j pi/(n+1)) (test_time.cpp:53-68 construction)
    n = 32
    j = np.arange(1, n + 1)
    xt = np.sin(j * np.pi / (n + 1))
    d = -4 * xt + np.concatenate(([0], xt[:-1])) + np.concatenate((xt[1:], [0]))
    rc, x = thomas(np.ones(n - 1), -4 * np.ones(n), np.ones(n - 1), d)
    assert rc == 0 and np.max(np.abs(x - xt)) < 1e-14


def test_thomas_bad_index_mistake_fails():
    # a plausible mistake (using c[k] instead of c[k-1], as the reference's C
    # implementation does) is caught by the dense check above
    L = ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(9)
    n = 10
    a, c = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - 1)
    b, d = 3 + rng.uniform(0, 1, n), rng.normal(size=n)
    x, ws = np.zeros(n), np.zeros(n)
    L.vibe_tridiagonal_solve_f64(_pd(a), _pd(b), _pd(c), _pd(d), n, _pd(x), _pd(ws))
    A = np.diag(b) + np.diag(a, -1) + np.diag(c, 1)
    assert np.max(np.abs(x - np.linalg.solve(A, d))) > 1e-6


def test_cfl_flux_golden():
    g = json.load(open(os.path.join(GOLD, "cfl_flux.json")))
    L = oracle_lib()
    arr = {k: np.array(g[k]) for k in ("u", "v", "w", "q", "rho", "dzeta", "zs")}
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    for res in g["results"]:
        zz = arr["zs"] if res["zs"] else None
        mw, md = ctypes.c_double(), ctypes.c_double()
        c = L.orc_cfl_scan(_pd(arr["u"]), _pd(arr["v"]), _pd(arr["w"]), _pd(arr["dzeta"]), _pd(zz), nx, ny, nz,
                           g["dx"], g["dy"], g["z_top"], g["dt"], ctypes.byref(mw), ctypes.byref(md))
        assert (c, mw.value, md.value) == (res["cfl"], res["max_wind"], res["min_dz"])
        out = np.zeros(nx * ny * nz)
        L.orc_advect_scalar_flux(_pd(arr["q"]), _pd(arr["rho"]), _pd(arr["u"]), _pd(arr["v"]), _pd(arr["w"]),
                                 _pd(arr["dzeta"]), _pd(zz), nx, ny, nz, g["dx"], g["dy"], g["z_top"], _pd(out))
        assert np.array_equal(out, np.array(res["dqdt"]))


def test_cfl_known_answers():
    # test_common.cpp:227-240
    nx, ny, nz = 4, 3, 3
    u = np.full(nx * ny * nz, 20.0)
    z = np.zeros(nx * ny * nz)
    dz = np.array([0.05, 0.10, 0.20])
    mw, md = ctypes.c_double(), ctypes.c_double()
    L = oracle_lib()
    c1 = L.orc_cfl_scan(_pd(u), _pd(z), _pd(z), _pd(dz), None, nx, ny, nz, 1000., 1000., 20000., 1.0,
                        ctypes.byref(mw), ctypes.byref(md))
    c2 = L.orc_cfl_scan(_pd(u), _pd(z), _pd(z), _pd(dz), None, nx, ny, nz, 1000., 1000., 20000., 2.0,
                        ctypes.byref(mw), ctypes.byref(md))
    assert c1 == 0.02 and c2 == 2 * c1 and mw.value == 20.0 and md.value == 1000.0


def test_flux_advection_constant_field_and_conservation():
    # test_common.cpp:255-266: constants give exactly 0
    nx, ny, nz = 6, 5, 4
    n = nx * ny * nz
    L = oracle_lib()
    q, rho = np.full(n, 3.0), np.full(n, 1.2)
    u, v, w = np.full(n, 7.0), np.full(n, -3.0), np.full(n, 0.5)
    dz = np.array([0.05, 0.08, 0.12, 0.2])
    out = np.full(n, 999.0)
    L.orc_advect_scalar_flux(_pd(q), _pd(rho), _pd(u), _pd(v), _pd(w), _pd(dz), None, nx, ny, nz, 1000., 1000.,
                             20000., _pd(out))
    assert np.max(np.abs(out)) == 0.0
    # periodic telescoping: sum(rho*dqdt) over one level with w = 0 vanishes
    rng = np.random.default_rng(3)
    q, rho, u, v = (rng.uniform(1, 2, n) for _ in range(4))
    w = np.zeros(n)
    L.orc_advect_scalar_flux(_pd(q), _pd(rho), _pd(u), _pd(v), _pd(w), _pd(dz), None, nx, ny, nz, 1000., 1000.,
                             20000., _pd(out))
    assert abs(np.sum((rho * out).reshape(nz, -1), axis=1)).max() < 1e-12


def test_neumaier_golden_and_exact():
    g = json.load(open(os.path.join(GOLD, "neumaier.json")))
    L = oracle_lib()
    for cs in g["cases"]:
        x = np.array(cs["x"])
        assert L.orc_neumaier_sum(_pd(x), x.size) == cs["sum"]
    rng = np.random.default_rng(11)
    x = rng.normal(size=5000) * 10.0 ** rng.integers(-8, 8, 5000)
    exact = float(sum(Fraction(v) for v in x))
    assert abs(L.orc_neumaier_sum(_pd(x), x.size) - exact) <= 2 * np.spacing(abs(exact))


def test_bilinear_golden():
    g = json.load(open(os.path.join(GOLD, "bilinear.json")))
    f = np.array(g["f"])
    L = oracle_lib()
    for p in g["points"]:
        assert L.orc_bilinear(_pd(f), g["halo"], g["nsx"], g["nsy"], g["nz"], p["xi"], p["yj"], p["k"]) == p["value"]
    # test_common.cpp:198-216 with the correct expectation (the reference test
    # omits the 100*k term): f = 10 i + j + 100 k
    halo, nx, ny, nz = 2, 5, 4, 2
    nsx, nsy = nx + 2 * halo, ny + 2 * halo
    f = np.zeros(nsx * nsy * (nz + 2 * halo))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                f[((k + halo) * nsy + (j + halo)) * nsx + (i + halo)] = 10.0 * i + j + 100.0 * k
    assert L.orc_bilinear(_pd(f), halo, nsx, nsy, nz, 2.0, 1.0, 1) == 121.0
    assert L.orc_bilinear(_pd(f), halo, nsx, nsy, nz, 1.5, 1.0, 0) == 16.0
    assert L.orc_bilinear(_pd(f), halo, nsx, nsy, nz, -50.0, -50.0, 0) == 0.0


def test_oracle_matches_compiled_reference():
    L = ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(21)
    O = oracle_lib()
    for _ in range(300):
        q = rng.normal(size=5) * float(rng.choice([1e-5, 1.0, 1e3]))
        l, r = ctypes.c_double(), ctypes.c_double()
        L.vibe_weno5_reconstruct_f64(*q, ctypes.byref(l), ctypes.byref(r))
        assert weno(q) == (l.value, r.value)
    x = rng.normal(size=20000)
    assert O.orc_neumaier_sum(_pd(x), x.size) == L.vibe_neumaier_sum_f64(_pd(x), x.size)
