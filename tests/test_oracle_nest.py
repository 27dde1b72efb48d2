"""Pins the oracle's nest lateral boundary (CPU only).

Reference: src/grid/nest.cpp:34-128 (davies_profile, interpolate_from_parent,
apply_lateral_boundary) and the bilinear arithmetic of c_kernels.c:108-144.
Checks: the Davies profile's closed form and end values, bilinear
interpolation reproducing linear parent fields exactly at every child point
(each field at its own staggered position), the relaxation formula applied
only inside the zone, and the time interpolation between the parent's two
states.
"""
import math

import numpy as np

from oracle_bind import FIELDS, oracle_lib, oracle_nest_apply, _pd


def davies(zone, amax):
    out = np.zeros(zone)
    oracle_lib().orc_davies_profile(zone, amax, _pd(out))
    return out


def test_davies_profile_closed_form():
    a = davies(5, 1.0)
    want = [0.5 * (1 - math.cos(math.pi * (5 - n) / 5)) for n in range(5)]
    assert np.allclose(a, want, rtol=0, atol=1e-15)
    assert a[0] == 1.0  # full replacement on the outer boundary
    assert np.all(np.diff(a) < 0) and a[-1] > 0
    assert np.allclose(davies(4, 0.5), 0.5 * davies(4, 1.0), rtol=0, atol=1e-16)


def _linear_parent(pnx, pny, nz, coef):
    """Fields linear in the parent index coordinates (x = i, y = j) of each
    field's own grid: value = a + b x + c y + d k."""
    st = {}
    for f, (a, b, c, d) in zip(FIELDS, coef):
        nk = nz + 1 if f == "w" else nz
        k, j, i = np.meshgrid(np.arange(nk), np.arange(pny), np.arange(pnx), indexing="ij")
        st[f] = (a + b * i + c * j + d * k).astype(np.float64).ravel()
    return st


def _empty(nx, ny, nz, val=0.0):
    return {f: np.full((nz + 1 if f == "w" else nz) * nx * ny, val) for f in FIELDS}


def test_interpolation_exact_on_linear_fields_with_staggering():
    pnx, pny, nz, r = 40, 36, 3, 3
    rng = np.random.default_rng(3)
    coef = [tuple(rng.normal(size=4)) for _ in FIELDS]
    parent = _linear_parent(pnx, pny, nz, coef)
    cnx, cny, i0, j0 = 30, 27, 21, 18  # covers parent cells 7..17 x 6..15, away from the wrap
    out = oracle_nest_apply(parent, None, pnx, pny, _empty(cnx, cny, nz), cnx, cny, nz, r, i0, j0, 5, 1.0, 1.0, 1)
    for f, (a, b, c, d) in zip(FIELDS, coef):
        nk = nz + 1 if f == "w" else nz
        k, j, i = np.meshgrid(np.arange(nk), np.arange(cny), np.arange(cnx), indexing="ij")
        xp = (i0 + i) / r if f == "u" else (i0 + i + 0.5) / r - 0.5
        yp = (j0 + j) / r if f == "v" else (j0 + j + 0.5) / r - 0.5
        want = a + b * xp + c * yp + d * k
        assert np.max(np.abs(out[f] - want.ravel())) < 1e-12, f


def test_relaxation_only_in_zone_and_formula():
    pnx, pny, nz, r, zone = 20, 20, 2, 3, 4
    rng = np.random.default_rng(9)
    pnew = {f: rng.normal(size=(nz + 1 if f == "w" else nz) * pnx * pny) for f in FIELDS}
    pold = {f: rng.normal(size=v.size) for f, v in pnew.items()}
    cnx, cny = 24, 18
    child = {f: rng.normal(size=(nz + 1 if f == "w" else nz) * cnx * cny) for f in FIELDS}
    full_new = oracle_nest_apply(pnew, None, pnx, pny, _empty(cnx, cny, nz), cnx, cny, nz, r, 6, 3, zone, 1.0, 1.0, 1)
    full_old = oracle_nest_apply(pold, None, pnx, pny, _empty(cnx, cny, nz), cnx, cny, nz, r, 6, 3, zone, 1.0, 1.0, 1)
    frac, amax = 0.3, 0.8
    out = oracle_nest_apply(pnew, pold, pnx, pny, child, cnx, cny, nz, r, 6, 3, zone, amax, frac, 0)
    al = davies(zone, amax)
    j, i = np.meshgrid(np.arange(cny), np.arange(cnx), indexing="ij")
    d = np.minimum(np.minimum(i, cnx - 1 - i), np.minimum(j, cny - 1 - j)).ravel()
    for f in FIELDS:
        nk = nz + 1 if f == "w" else nz
        dd = np.tile(d, nk)
        p = full_old[f] + frac * (full_new[f] - full_old[f])
        a = np.where(dd < zone, al[np.minimum(dd, zone - 1)], 0.0)
        want = np.where(dd < zone, (1 - a) * child[f] + a * p, child[f])
        assert np.max(np.abs(out[f] - want)) < 1e-14, f
        assert np.array_equal(out[f][dd >= zone], child[f][dd >= zone])
        # the outer ring takes the parent value (alpha = alpha_max)
        ring = dd == 0
        assert np.allclose(out[f][ring], (1 - amax) * child[f][ring] + amax * p[ring], rtol=0, atol=1e-14)
