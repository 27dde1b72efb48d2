"""Pins of the Helmholtz operator restatement (oracle/helmholtz.py) against
the reference's unit-test values and closed forms."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import helmholtz as orc  # noqa: E402


def test_residual_of_zero_matches_reference_test():
    # tests/unit/test_time.cpp:173-183: a = b = f = 1, x = 0 on 4x4x4 -> sqrt(64)
    sh = (4, 4, 4)
    dz = orc.dz_of(np.linspace(0, 1, 5), np.zeros((4, 4)), 20000.0)
    r = orc.residual_l2(np.ones(sh), np.ones(sh), np.ones(sh), np.zeros(sh), 2000.0, 2000.0, dz)
    assert abs(r - 8.0) <= 1e-12


def test_constant_field_sees_only_b():
    rng = np.random.default_rng(0)
    sh = (6, 5, 7)
    a, b = rng.uniform(0.5, 2, sh), rng.uniform(0.1, 1, sh)
    dz = orc.dz_of(np.cumsum(np.r_[0, rng.uniform(1, 2, 6)]) / 9.0, rng.uniform(0, 500, (5, 7)), 20000.0)
    y = orc.apply(a, b, np.full(sh, 3.0), 1000.0, 1500.0, dz)
    assert np.array_equal(y, 3.0 * b)  # every flux vanishes exactly


def test_quadratic_in_x_has_the_closed_form_second_difference():
    nx = 9
    sh = (3, 2, nx)
    i = np.arange(nx, dtype=np.float64)
    x = np.broadcast_to(i * i, sh).copy()
    dz = orc.dz_of(np.linspace(0, 1, 4), np.zeros((2, nx)), 3000.0)
    dx = 250.0
    y = orc.apply(np.ones(sh), np.zeros(sh), x, dx, 300.0, dz)
    want = np.full(nx, -2.0 / dx**2)
    want[0] = -1.0 / dx**2                       # zero flux through the west boundary
    want[-1] = (2 * nx - 3) / dx**2              # zero flux through the east boundary
    assert np.allclose(y, np.broadcast_to(want, sh), rtol=1e-12, atol=0)


def test_vertical_manufactured_solution_of_the_reference_test():
    # tests/unit/test_time.cpp:116-152: a = 1, b = 0.01, x = sin(pi (k+1)/(nz+1))
    # uniform in x, y; f from the reference's own tridiagonal formula
    nz, ztop = 16, 20000.0
    dzeta = 1.08 ** np.arange(nz)
    zeta = np.r_[0.0, np.cumsum(dzeta)] / np.sum(dzeta)
    sh = (nz, 4, 4)
    dz = orc.dz_of(zeta, np.zeros((4, 4)), ztop)
    xt = np.broadcast_to(np.sin(np.pi * (np.arange(nz) + 1) / (nz + 1))[:, None, None], sh).copy()
    f = np.empty(sh)
    for k in range(nz):
        d = (zeta[k + 1] - zeta[k]) * ztop
        diag, off = 2.0 / (d * d) + 0.01, 1.0 / (d * d)
        f[k] = diag * xt[k] - off * (xt[max(k - 1, 0)] + xt[min(k + 1, nz - 1)])
    y = orc.apply(np.ones(sh), np.full(sh, 0.01), xt, 2000.0, 2000.0, dz)
    assert np.allclose(y, f, rtol=1e-12, atol=1e-20)
    x = orc.solve(np.ones(sh), np.full(sh, 0.01), f, 2000.0, 2000.0, dz)
    assert np.max(np.abs(x - xt)) < 1e-8


def test_direct_solution_has_tiny_residual():
    rng = np.random.default_rng(4)
    sh = (5, 4, 6)
    a, b = rng.uniform(0.5, 2, sh), rng.uniform(0.1, 1, sh)
    dz = orc.dz_of(np.linspace(0, 1, 6), rng.uniform(0, 300, (4, 6)), 5000.0)
    f = rng.normal(size=sh)
    x = orc.solve(a, b, f, 800.0, 900.0, dz)
    assert orc.residual_l2(a, b, f, x, 800.0, 900.0, dz) <= 1e-12 * np.linalg.norm(f)
