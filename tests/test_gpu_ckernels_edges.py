"""Edge cases of the six C-ABI drop-ins on the GPU against the reference's
own compiled c_kernels.c (oracle/_ref/libref_ckernels.so, built from
src/common/c_kernels.c unmodified), bit for bit: empty and degenerate grids
(nx, ny or nz zero or negative), NULL optional and required arguments,
ragged shapes (1 x 1 x 1, widths that are not multiples of a wavefront),
level and position clamps of the bilinear stencil, and sums of zero, one
and cancelling values.  Behaviour follows c_kernels.c:22-24, 71, 115-119,
158, 183-184, 194, 219-222.
"""
import ctypes

import numpy as np
import pytest

from oracle_bind import _pd, ref_lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    L = ref_lib()
    if L is None:
        pytest.skip("oracle/_ref/libref_ckernels.so is not built")
    return L


def _cfl(L, u, v, w, dz, zs, nx, ny, nz, dx=1000.0, dy=900.0, zt=20000.0, dt=3.0):
    mw, md = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
    c = L.vibe_cfl_scan_f64(_pd(u), _pd(v), _pd(w), _pd(dz), _pd(zs), nx, ny, nz, dx, dy, zt, dt,
                            ctypes.byref(mw), ctypes.byref(md))
    return c, mw.value, md.value


def test_cfl_degenerate_and_ragged_grids(vmi, R):
    L = vmi.lib()
    rng = np.random.default_rng(31)
    for nx, ny, nz in [(0, 4, 3), (4, 0, 3), (4, 3, 0), (-2, 3, 2), (1, 1, 1), (3, 65, 1), (65, 1, 2), (7, 9, 5)]:
        m = max(nx, 0) * max(ny, 0) * max(nz, 0)
        u, v, w = (rng.normal(size=max(m, 1)) * 20 for _ in range(3))
        dz = rng.uniform(1e-5, 0.2, max(nz, 1))  # includes dz < 1 m, clamped to 1
        zs = rng.uniform(0, 500, max(nx * ny, 1))
        for z in (None, zs):
            g, r = _cfl(L, u, v, w, dz, z, nx, ny, nz), _cfl(R, u, v, w, dz, z, nx, ny, nz)
            assert g == r, (nx, ny, nz, z is None, g, r)
    # required NULL arguments: 0.0 and the outputs untouched
    mw, md = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
    for L_ in (L, R):
        assert L_.vibe_cfl_scan_f64(None, None, None, None, None, 2, 2, 2, 1.0, 1.0, 1.0, 1.0,
                                    ctypes.byref(mw), ctypes.byref(md)) == 0.0
        assert mw.value == -1.0 and md.value == -1.0
    # NULL outputs are allowed
    x = np.ones(8)
    assert L.vibe_cfl_scan_f64(_pd(x), _pd(x), _pd(x), _pd(np.ones(2)), None, 2, 2, 2, 1.0, 1.0, 1.0, 1.0,
                               None, None) == R.vibe_cfl_scan_f64(_pd(x), _pd(x), _pd(x), _pd(np.ones(2)), None, 2, 2,
                                                                  2, 1.0, 1.0, 1.0, 1.0, None, None)


def test_flux_degenerate_and_ragged_grids(vmi, R):
    L = vmi.lib()
    rng = np.random.default_rng(32)
    for nx, ny, nz in [(0, 3, 2), (3, 0, 2), (3, 2, 0), (1, 1, 1), (1, 5, 3), (5, 1, 3), (2, 2, 2), (65, 3, 4),
                       (3, 65, 1)]:
        m = max(nx, 0) * max(ny, 0) * max(nz, 0)
        q, rho = rng.uniform(1, 2, max(m, 1)), rng.uniform(0.5, 1.2, max(m, 1))
        u, v, w = (rng.normal(size=max(m, 1)) for _ in range(3))
        dz = rng.uniform(1e-5, 0.1, max(nz, 1))
        zs = rng.uniform(0, 500, max(nx * ny, 1))
        for z in (None, zs):
            outs = []
            for L_ in (L, R):
                out = np.full(max(m, 1), 7.0)
                L_.vibe_advect_scalar_flux_f64(_pd(q), _pd(rho), _pd(u), _pd(v), _pd(w), _pd(dz), _pd(z), nx, ny, nz,
                                               1000.0, 900.0, 20000.0, _pd(out))
                outs.append(out)
            assert np.array_equal(outs[0], outs[1]), (nx, ny, nz, z is None)
            if m == 0:
                assert np.all(outs[0] == 7.0)  # nothing written
    assert L.vibe_mi_last_error() == 0
    # a NULL required argument writes nothing
    out = np.full(8, 7.0)
    x = np.ones(8)
    L.vibe_advect_scalar_flux_f64(None, _pd(x), _pd(x), _pd(x), _pd(x), _pd(np.ones(2)), None, 2, 2, 2, 1.0, 1.0,
                                  1.0, _pd(out))
    assert np.all(out == 7.0)


def test_neumaier_empty_single_and_cancelling(vmi, R):
    L = vmi.lib()
    x = np.array([3.5])
    assert L.vibe_neumaier_sum_f64(_pd(x), 0) == R.vibe_neumaier_sum_f64(_pd(x), 0) == 0.0
    assert L.vibe_neumaier_sum_f64(None, 5) == R.vibe_neumaier_sum_f64(None, 5) == 0.0
    assert L.vibe_neumaier_sum_f64(_pd(x), 1) == R.vibe_neumaier_sum_f64(_pd(x), 1) == 3.5
    for y in (np.array([1e300, 1.0, -1e300]), np.array([1.0, 1e100, 1.0, -1e100]),
              np.array([0.1] * 10 + [-1.0]), np.array([-0.0, -0.0])):
        y = np.ascontiguousarray(y)
        assert L.vibe_neumaier_sum_f64(_pd(y), y.size) == R.vibe_neumaier_sum_f64(_pd(y), y.size), y


def test_bilinear_clamps_and_null(vmi, R):
    L = vmi.lib()
    rng = np.random.default_rng(33)
    halo, nx, ny, nz = 3, 6, 5, 4
    nsx, nsy = nx + 2 * halo, ny + 2 * halo
    f = rng.normal(size=nsx * nsy * (nz + 2 * halo))
    for xi, yj, k in [(0.0, 0.0, 0), (nx - 1.0, ny - 1.0, nz - 1), (2.5, 3.25, 2), (-7.3, 2.0, 1),
                      (100.0, -100.0, 3), (1.999999, 0.000001, -4), (3.0, 2.0, 99), (-0.5, -0.5, 0)]:
        assert L.vibe_bilinear_f64(_pd(f), halo, nsx, nsy, nz, xi, yj, k) == \
            R.vibe_bilinear_f64(_pd(f), halo, nsx, nsy, nz, xi, yj, k), (xi, yj, k)
    for L_ in (L, R):
        assert L_.vibe_bilinear_f64(None, halo, nsx, nsy, nz, 1.0, 1.0, 0) == 0.0
        assert L_.vibe_bilinear_f64(_pd(f), halo, nsx, nsy, 0, 1.0, 1.0, 0) == 0.0


def test_weno_and_thomas_arguments(vmi, R):
    L = vmi.lib()
    # NULL outputs: nothing happens (no write, no fault)
    L.vibe_weno5_reconstruct_f64(1.0, 2.0, 3.0, 4.0, 5.0, None, None)
    assert L.vibe_mi_last_error() == 0
    for q in ([0.0] * 5, [1e-300, -1e-300, 0.0, 1e-300, 0.0], [1e8, -1e8, 1e8, -1e8, 1e8]):
        a = [ctypes.c_double(), ctypes.c_double()]
        b = [ctypes.c_double(), ctypes.c_double()]
        L.vibe_weno5_reconstruct_f64(*q, ctypes.byref(a[0]), ctypes.byref(a[1]))
        R.vibe_weno5_reconstruct_f64(*q, ctypes.byref(b[0]), ctypes.byref(b[1]))
        assert (a[0].value, a[1].value) == (b[0].value, b[1].value), q
    # Thomas return codes: n = 1 (a and c ignored), a zero pivot, bad sizes
    x, ws = np.zeros(3), np.zeros(3)
    one, zero, two = np.array([2.0]), np.array([0.0]), np.array([4.0])
    for L_ in (L, R):
        assert L_.vibe_tridiagonal_solve_f64(None, _pd(one), None, _pd(two), 1, _pd(x), _pd(ws)) == 0
        assert x[0] == 2.0
        assert L_.vibe_tridiagonal_solve_f64(None, _pd(zero), None, _pd(two), 1, _pd(x), _pd(ws)) == 1
        assert L_.vibe_tridiagonal_solve_f64(None, _pd(one), None, _pd(two), -3, _pd(x), _pd(ws)) == 2
        assert L_.vibe_tridiagonal_solve_f64(None, _pd(one), None, _pd(two), 1, _pd(x), None) == 2
