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


# ---- multigrid_restrict / multigrid_prolong (helmholtz.cpp:131-164) ------

def test_restrict_prolong_norm_bound_of_the_reference_test():
    # tests/unit/test_time.cpp:101-114: 8x8x4 -> 4x4x4 -> 8x8x4,
    # fine = sin(0.7 i) + cos(0.5 j); the round trip must not grow the norm
    # by more than 1.5 and stays finite
    i = np.arange(8.0)
    fine = np.broadcast_to(np.sin(0.7 * i)[None, None, :] + np.cos(0.5 * i)[None, :, None], (4, 8, 8)).copy()
    coarse = orc.restrict(fine, (4, 4, 4))
    back = orc.prolong(coarse, (4, 8, 8))
    assert np.all(np.isfinite(coarse))
    assert np.linalg.norm(back) <= 1.5 * np.linalg.norm(fine)


def test_transfers_reproduce_constants():
    # the restriction weights sum to one (1/4 + 4/8 + 4/16), the bilinear
    # weights too: a constant is reproduced exactly by both
    c = np.full((3, 9, 11), 2.5)
    assert np.array_equal(orc.restrict(c, (3, 5, 6)), np.full((3, 5, 6), 2.5))
    assert np.array_equal(orc.prolong(c, (3, 17, 21)), np.full((3, 17, 21), 2.5))


def test_restriction_of_a_linear_field_is_its_value_at_the_coarse_point():
    # full weighting is symmetric about (2i, 2j): a field linear in (i, j)
    # restricts to its value there wherever no neighbour is clamped; a
    # weight moved off-centre or a transposed index breaks it
    nz, ny, nx = 2, 13, 15
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fine = 3.0 * i - 2.0 * j + 7.0 * k + 1.0
    cs = (nz, 6, 7)
    co = orc.restrict(fine, cs)
    kc, jc, ic = np.meshgrid(*[np.arange(n) for n in cs], indexing="ij")
    want = 3.0 * (2 * ic) - 2.0 * (2 * jc) + 7.0 * kc + 1.0
    inner = (ic >= 1) & (2 * ic + 1 <= nx - 1) & (jc >= 1) & (2 * jc + 1 <= ny - 1)
    assert np.array_equal(co[inner], want[inner])
    # at the clamped west edge the i-1 neighbour is the edge value itself
    assert co[0, 3, 0] == 0.25 * fine[0, 6, 0] + 0.125 * (fine[0, 6, 0] + fine[0, 6, 1] + fine[0, 5, 0] +
                                                          fine[0, 7, 0]) + 0.0625 * (
        fine[0, 5, 0] + fine[0, 5, 1] + fine[0, 7, 0] + fine[0, 7, 1])


def test_prolongation_is_exact_for_bilinear_fields():
    # bilinear interpolation reproduces a + b x + c y + d x y exactly at
    # fine points inside the coarse extent; even fine points take the coarse
    # value itself
    nz, cy, cx = 2, 5, 6
    kc, jc, ic = np.meshgrid(np.arange(nz), np.arange(cy), np.arange(cx), indexing="ij")
    g = lambda x, y, z: 0.5 + 2.0 * x - 3.0 * y + 0.25 * x * y + z  # noqa: E731
    co = g(ic * 1.0, jc * 1.0, kc * 1.0)
    fs = (nz, 2 * cy - 1, 2 * cx - 1)
    fi = orc.prolong(co, fs)
    k, j, i = np.meshgrid(*[np.arange(n) for n in fs], indexing="ij")
    assert np.allclose(fi, g(0.5 * i, 0.5 * j, k * 1.0), rtol=0, atol=1e-13)
    assert np.array_equal(fi[:, ::2, ::2], co)
    # past the coarse extent the clamped neighbour is the edge value, so the
    # last fine column repeats the one on the last coarse column
    f2 = orc.prolong(co, (nz, 2 * cy, 2 * cx))
    assert np.array_equal(f2[:, :, -1], f2[:, :, -2])
    assert np.array_equal(f2[:, -1, :], f2[:, -2, :])


def test_restrict_of_larger_coarse_field_clamps_into_the_fine_one():
    # the coarse field's extents set the loop (the reference calls it with a
    # field of the fine grid's size): points past nx/2 read clamped values
    fine = np.arange(2 * 4 * 6, dtype=np.float64).reshape(2, 4, 6)
    full = orc.restrict(fine, (2, 4, 6))
    assert np.array_equal(full[:, :, 3:], np.broadcast_to(full[:, :, 3:4], full[:, :, 3:].shape))
