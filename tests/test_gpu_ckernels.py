"""GPU C-ABI kernels against the oracle restatement and the reference's own
compiled c_kernels.c (golden fixtures)."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle_bind import oracle_lib, _pd

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_weno5_bit_exact_vs_oracle(vmi):
    import ctypes
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rng.normal(size=5) * rng.choice([1e-3, 1.0, 300.0])
        g = vmi.weno5_reconstruct(*q)
        l, r = ctypes.c_double(), ctypes.c_double()
        oracle_lib().orc_weno5(*q, ctypes.byref(l), ctypes.byref(r))
        assert g == (l.value, r.value)


def test_weno5_golden_from_reference(vmi):
    data = json.load(open(os.path.join(GOLD, "weno5.json")))
    for case in data["cases"]:
        g = vmi.weno5_reconstruct(*case["q"])
        assert g[0] == case["q_left"] and g[1] == case["q_right"]


def test_thomas_matches_oracle_and_numpy(vmi):
    rng = np.random.default_rng(2)
    for n in (1, 2, 3, 17, 96):
        a = rng.uniform(-1, 1, n - 1)
        c = rng.uniform(-1, 1, n - 1)
        b = 3.0 + rng.uniform(0, 1, n)
        d = rng.normal(size=n)
        rc, x = vmi.tridiagonal_solve(a, b, c, d)
        assert rc == 0
        xo = np.zeros(n); ws = np.zeros(n)
        assert oracle_lib().orc_tridiagonal_solve(_pd(a) if n > 1 else None, _pd(b), _pd(c) if n > 1 else None,
                                                  _pd(d), n, _pd(xo), _pd(ws)) == 0
        assert np.array_equal(x, xo)
        A = np.diag(b) + np.diag(a, -1) + np.diag(c, 1)
        ref = np.linalg.solve(A, d)
        assert np.max(np.abs(x - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_thomas_return_codes(vmi):
    # tests/unit/test_common.cpp:177-185 of the reference
    rc, _ = vmi.tridiagonal_solve(None, [0.0, 0.0], None, [1.0, 1.0])
    assert rc == 1
    import ctypes
    L = vmi.lib()
    x = np.zeros(2); ws = np.zeros(2)
    assert L.vibe_tridiagonal_solve_f64(None, None, None, None, 2, vmi._pd(x), vmi._pd(ws)) == 2
    b = np.zeros(2); d = np.ones(2)
    assert L.vibe_tridiagonal_solve_f64(None, vmi._pd(b), None, vmi._pd(d), 0, vmi._pd(x), vmi._pd(ws)) == 2


def test_thomas_sine_closed_form(vmi):
    # diag(-1, 2, -1) x = d with x_j = sin(j pi/(n+1)) exactly solvable
    n = 64
    j = np.arange(1, n + 1)
    xt = np.sin(j * np.pi / (n + 1))
    d = 2 * xt - np.concatenate(([0], xt[:-1])) - np.concatenate((xt[1:], [0]))
    rc, x = vmi.tridiagonal_solve(-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1), d)
    assert rc == 0 and np.max(np.abs(x - xt)) < 1e-12


def test_batched_thomas_device(vmi):
    import ctypes
    # device buffers via the library's own stream: use hip runtime directly
    hip = ctypes.CDLL("libamdhip64.so")
    rng = np.random.default_rng(3)
    n, nsys = 40, 1000
    a = rng.uniform(-1, 1, (n - 1, nsys)); c = rng.uniform(-1, 1, (n - 1, nsys))
    b = 3 + rng.uniform(0, 1, (n, nsys)); d = rng.normal(size=(n, nsys))
    bufs = []
    def dev(arr):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(arr.nbytes)) == 0
        assert hip.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0
        bufs.append(p)
        return p
    da, db, dc, dd = (dev(np.ascontiguousarray(q)) for q in (a, b, c, d))
    dx, dws = dev(np.zeros((n, nsys))), dev(np.zeros((n, nsys)))
    rc = vmi.lib().vibe_mi_tridiagonal_solve_batched_f64(da, db, dc, dd, n, nsys, dx, dws, None)
    assert rc == 0
    x = np.zeros((n, nsys))
    assert hip.hipMemcpy(x.ctypes.data_as(ctypes.c_void_p), dx, ctypes.c_size_t(x.nbytes), 2) == 0
    for s in range(0, nsys, 97):
        xo = np.zeros(n); ws = np.zeros(n)
        aa, bb, cc, dd_ = (np.ascontiguousarray(q[:, s]) for q in (a, b, c, d))
        oracle_lib().orc_tridiagonal_solve(_pd(aa), _pd(bb), _pd(cc), _pd(dd_), n, _pd(xo), _pd(ws))
        assert np.array_equal(x[:, s], xo)
    for p in bufs:
        hip.hipFree(p)


def test_neumaier_exact(vmi):
    rng = np.random.default_rng(4)
    x = rng.normal(size=100001) * 10.0 ** rng.integers(-8, 8, size=100001)
    exact = float(sum(Fraction(v) for v in x))
    g = vmi.neumaier_sum(x)
    assert abs(g - exact) <= 4 * np.spacing(abs(exact))
    # c_kernels.c / test_common.cpp:242-253 cancellation sequence
    y = np.ones(10000); y[0] = 1e16; y[-1] = -1e16
    assert vmi.neumaier_sum(y) == 9998.0
    assert vmi.neumaier_sum(np.ones(10 ** 6)) == 1e6


def test_cfl_and_flux_bit_exact(vmi):
    L = oracle_lib()
    rng = np.random.default_rng(5)
    nx, ny, nz = 13, 9, 7
    n = nx * ny * nz
    u, v, w = rng.normal(size=n) * 10, rng.normal(size=n) * 10, rng.normal(size=n)
    q, rho = rng.uniform(1, 2, n), rng.uniform(0.5, 1.2, n)
    dz = rng.uniform(0.05, 0.2, nz)
    zs = rng.uniform(0, 500, nx * ny)
    import ctypes
    for zz in (None, zs):
        mw, md = ctypes.c_double(), ctypes.c_double()
        o = L.orc_cfl_scan(_pd(u), _pd(v), _pd(w), _pd(dz), _pd(zz), nx, ny, nz, 1000., 900., 20000., 3.,
                           ctypes.byref(mw), ctypes.byref(md))
        g = vmi.cfl_scan(u, v, w, dz, zz, nx, ny, nz, 1000., 900., 20000., 3.)
        assert g == (o, mw.value, md.value)
        out = np.zeros(n)
        L.orc_advect_scalar_flux(_pd(q), _pd(rho), _pd(u), _pd(v), _pd(w), _pd(dz), _pd(zz), nx, ny, nz,
                                 1000., 900., 20000., _pd(out))
        gf = vmi.advect_scalar_flux(q, rho, u, v, w, dz, zz, nx, ny, nz, 1000., 900., 20000.)
        assert np.array_equal(gf, out)


def test_bilinear_matches_oracle(vmi):
    L = oracle_lib()
    halo, nx, ny, nz = 2, 5, 4, 2
    nsx, nsy = nx + 2 * halo, ny + 2 * halo
    f = np.random.default_rng(6).normal(size=nsx * nsy * (nz + 2 * halo))
    for xi, yj, k in [(2.0, 1.0, 1), (1.5, 1.0, 0), (-50., -50., 0), (3.7, 2.2, 5), (0.3, 9.0, -1)]:
        assert vmi.bilinear(f, halo, nsx, nsy, nz, xi, yj, k) == L.orc_bilinear(_pd(f), halo, nsx, nsy, nz, xi, yj, k)
