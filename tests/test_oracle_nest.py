"""Pins the oracle's nest lateral boundary (CPU only).

Reference: src/grid/nest.cpp:34-128 (davies_profile, interpolate_from_parent,
restrict_to_parent, apply_lateral_boundary) and the bilinear arithmetic of
c_kernels.c:108-144.  Checks: the Davies profile's closed form and end
values, bilinear interpolation reproducing linear parent fields exactly at
every child point (each field at its own staggered position), the relaxation
formula applied only inside the zone, the time interpolation between the
parent's two states, and the feedback: bit-exact against the reference's
compiled restrict_to_parent on a cell field, exact means of linear fields at
each field's parent position, and conservation of the covered mass.
"""
import math

import numpy as np

from oracle_bind import FIELDS, oracle_lib, oracle_nest_apply, oracle_nest_restrict, _pd


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


def test_restriction_matches_reference_on_cell_field():
    """orc_nest_restrict against the reference's Nest::restrict_to_parent
    (nest.cpp:85-108, compiled by oracle/ref.mk; tests/golden) on the 16x14x3
    parent p = i/2 - j/4 + k/8 + i j/100 and the relaxed 12x9x3 child at
    offset (12, 15) child cells, ratio 3 (parent cells 4..7 x 5..7): bit for
    bit on every parent cell, and the cells outside the nest untouched."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "setup_reference.json")))
    pnx, pny, nz, cnx, cny, r = 16, 14, 3, 12, 9, 3
    k, j, i = np.meshgrid(np.arange(nz), np.arange(pny), np.arange(pnx), indexing="ij")
    p = (0.5 * i - 0.25 * j + 0.125 * k + 0.01 * i * j).ravel()
    child = np.array(gold["nest_relaxed_cell_zone4"])
    cst = {f: (np.zeros(cnx * cny * (nz + 1)) if f == "w" else child.copy()) for f in FIELDS}
    pst = {f: (np.zeros(pnx * pny * (nz + 1)) if f == "w" else p.copy()) for f in FIELDS}
    out = oracle_nest_restrict(cst, cnx, cny, pst, pnx, pny, nz, r, 12, 15)
    want = np.array(gold["nest_restricted_parent_cell"])
    for f in ("rho", "theta", "pi"):
        assert np.array_equal(out[f], want), f
    assert not np.array_equal(want, p)  # the feedback changed the covered cells


def test_restriction_of_linear_fields_and_mass():
    """Linear child fields (in child index coordinates) restrict to the value
    at each field's parent position: the parent cell centre for cell fields
    and w, the parent west-face midpoint for u, the south-face midpoint for v
    (to rounding); the covered parent mass equals the child's mass."""
    pnx, pny, nz, r = 20, 18, 3, 3
    cnx, cny, i0, j0 = 15, 12, 9, 6
    rng = np.random.default_rng(11)
    coef = {f: rng.normal(size=4) for f in FIELDS}
    child = {}
    for f in FIELDS:
        nk = nz + 1 if f == "w" else nz
        k, j, i = np.meshgrid(np.arange(nk), np.arange(cny), np.arange(cnx), indexing="ij")
        a, b, c, d = coef[f]
        child[f] = (a + b * i + c * j + d * k).ravel()
    parent = _empty(pnx, pny, nz, 7.0)
    out = oracle_nest_restrict(child, cnx, cny, parent, pnx, pny, nz, r, i0, j0)
    I0, J0, iext, jext = i0 // r, j0 // r, cnx // r, cny // r
    for f in FIELDS:
        nk = nz + 1 if f == "w" else nz
        got = out[f].reshape(nk, pny, pnx)
        k, jj, ii = np.meshgrid(np.arange(nk), np.arange(jext), np.arange(iext), indexing="ij")
        # child index coordinate of the parent position of each field
        xc = ii * r + (0.0 if f == "u" else (r - 1) / 2.0)
        yc = jj * r + (0.0 if f == "v" else (r - 1) / 2.0)
        if f == "u":
            yc = jj * r + (r - 1) / 2.0
        if f == "v":
            xc = ii * r + (r - 1) / 2.0
        a, b, c, d = coef[f]
        want = a + b * xc + c * yc + d * k
        sub = got[:, J0:J0 + jext, I0:I0 + iext]
        assert np.max(np.abs(sub - want)) < 1e-12, f
        mask = np.ones((pny, pnx), bool)
        mask[J0:J0 + jext, I0:I0 + iext] = False
        assert np.all(got[:, mask] == 7.0), f  # outside the nest untouched
    # mass of the covered region (equal cell areas: parent area = r^2 child areas)
    rho_c = child["rho"].reshape(nz, cny, cnx)[:, :jext * r, :iext * r]
    rho_p = out["rho"].reshape(nz, pny, pnx)[:, J0:J0 + jext, I0:I0 + iext]
    assert abs(rho_p.sum() * r * r - rho_c.sum()) <= 1e-12 * np.abs(rho_c).sum()
