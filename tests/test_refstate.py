"""Conversion between the reference's State layout (12 species, halo'd
fields; include/vibe/grid/field.hpp:34-96, include/vibe/dyn/state.hpp:53-56)
and the library's dense perturbation state (vibe_mi/dycore.h).  Host code
only (no GPU).

* The layout is the reference's own: tests/golden holds the raw storage of a
  State whose species were filled through the reference's (i, j, k)
  indexing (oracle/ref_setup.cpp, 5 x 4 x 3, halo 2); from_reference must
  read back exactly the values the reference wrote, and the field sizes must
  equal the reference's storage sizes.
* to_reference then from_reference is the identity bit for bit; the
  duplicated periodic faces u(nx) = u(0), v(ny) = v(0) are written; with
  fill_halo the lateral halo holds the periodic values (the reference state
  taken at the periodic interior point); without it halos, and the moisture
  species always, are left untouched.
"""
import ctypes
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setup_reference.json")
SPECIES = ("u", "v", "w", "rho", "theta", "pi", "qv", "qc", "qr", "qi", "qs", "qg")


@pytest.fixture(scope="module")
def L():
    import vibe_mi
    return vibe_mi


def _pp(arrs):
    P = ctypes.POINTER(ctypes.c_double)
    return (P * 12)(*[a.ctypes.data_as(P) if a is not None else P() for a in arrs])


def _dense(nx, ny, nz):
    c = nx * ny * nz
    return {k: np.zeros(c + (nx * ny if k == "w" else 0)) for k in ("u", "v", "w", "rho", "theta", "pi")}


def _from_ref(L, nx, ny, nz, h, species, r0, t0, p0):
    d = _dense(nx, ny, nz)
    pd = L._pd
    rc = L.lib().vibe_mi_state_from_reference(nx, ny, nz, h, _pp(species), pd(r0), pd(t0), pd(p0),
                                              *[pd(d[k]) for k in ("u", "v", "w", "rho", "theta", "pi")])
    return rc, d


def _to_ref(L, nx, ny, nz, h, d, r0, t0, p0, fill, species):
    pd = L._pd
    return L.lib().vibe_mi_state_to_reference(nx, ny, nz, h, *[pd(d[k]) for k in ("u", "v", "w", "rho", "theta", "pi")],
                                              pd(r0), pd(t0), pd(p0), fill, _pp(species))


def test_layout_is_the_references(L):
    gold = json.load(open(GOLD))
    nx, ny, nz, h = 5, 4, 3, 2
    species = [np.array(gold[f"layout_h2_{s}"], dtype=np.float64) for s in SPECIES]
    for q, a in enumerate(species):
        assert L.lib().vibe_mi_reference_field_size(nx, ny, nz, h, q) == a.size, SPECIES[q]
    zero = np.zeros(species[3].size)
    rc, d = _from_ref(L, nx, ny, nz, h, species, zero, zero, zero)
    assert rc == 0
    for q, key in enumerate(("u", "v", "w", "rho", "theta", "pi")):
        nk = nz + 1 if key == "w" else nz
        k, j, i = np.meshgrid(np.arange(nk), np.arange(ny), np.arange(nx), indexing="ij")
        want = (1000.0 * q + 100.0 * k + 10.0 * j + i).ravel()
        assert np.array_equal(d[key], want), key


def _ref_arrays(L, nx, ny, nz, h, fill=-7.0):
    return [np.full(L.lib().vibe_mi_reference_field_size(nx, ny, nz, h, q), fill) for q in range(12)]


def test_round_trip_faces_halos_and_moisture(L):
    nx, ny, nz, h = 7, 5, 4, 4
    rng = np.random.default_rng(2)
    d = {k: rng.normal(size=v.size) for k, v in _dense(nx, ny, nz).items()}
    nc = L.lib().vibe_mi_reference_field_size(nx, ny, nz, h, 3)
    r0, t0, p0 = (np.full(nc, 0.0) for _ in range(3))
    # reference state on the interior only (halo zero, as ReferenceState leaves it)
    def lay(sp):
        nsx = nx + 2 * h + (1 if sp == 0 else 0)
        nsy = ny + 2 * h + (1 if sp == 1 else 0)
        nsz = nz + 2 * h + (1 if sp == 2 else 0)
        return nsx, nsy, nsz
    nsx, nsy, nsz = lay(3)
    R0 = r0.reshape(nsz, nsy, nsx)
    T0 = t0.reshape(nsz, nsy, nsx)
    P0 = p0.reshape(nsz, nsy, nsx)
    R0[h:h + nz, h:h + ny, h:h + nx] = 1.0 + rng.uniform(size=(nz, ny, nx))
    T0[h:h + nz, h:h + ny, h:h + nx] = 300.0 + rng.uniform(size=(nz, ny, nx))
    P0[h:h + nz, h:h + ny, h:h + nx] = 0.9 + 0.01 * rng.uniform(size=(nz, ny, nx))
    for fill in (0, 1):
        sp = _ref_arrays(L, nx, ny, nz, h)
        assert _to_ref(L, nx, ny, nz, h, d, r0, t0, p0, fill, sp) == 0
        rc, back = _from_ref(L, nx, ny, nz, h, sp, r0, t0, p0)
        assert rc == 0
        for k in ("u", "v", "w"):
            assert np.array_equal(back[k], d[k]), (fill, k)
        # the full value rho0 + rho' is an FP64 rounding of the sum: the
        # perturbation comes back as (rho0 + rho') - rho0, bit for bit
        for k, base in (("rho", R0), ("theta", T0), ("pi", P0)):
            b0 = base[h:h + nz, h:h + ny, h:h + nx].ravel()
            assert np.array_equal(back[k], (b0 + d[k]) - b0), (fill, k)
            assert np.max(np.abs(back[k] - d[k])) <= 4e-16 * np.max(np.abs(b0 + d[k])), (fill, k)
        for q in range(6, 12):
            assert np.all(sp[q] == -7.0)  # moisture untouched
        U = sp[0].reshape(*lay(0)[::-1])
        V = sp[1].reshape(*lay(1)[::-1])
        Th = sp[4].reshape(nsz, nsy, nsx)
        # duplicated periodic faces
        assert np.array_equal(U[h:h + nz, h:h + ny, h + nx], U[h:h + nz, h:h + ny, h])
        assert np.array_equal(V[h:h + nz, h + ny, h:h + nx], V[h:h + nz, h, h:h + nx])
        if fill:
            # west halo column = last interior column, south halo row = last row
            assert np.array_equal(U[h:h + nz, h:h + ny, h - 1], U[h:h + nz, h:h + ny, h + nx - 1])
            assert np.array_equal(Th[h:h + nz, h:h + ny, 0], Th[h:h + nz, h:h + ny, nx])
            assert np.array_equal(Th[h:h + nz, 0, h:h + nx], Th[h:h + nz, ny, h:h + nx])
            assert np.array_equal(Th[h:h + nz, 0, 0], Th[h:h + nz, ny, nx])  # corner
            assert np.all(Th[:h] == -7.0) and np.all(Th[h + nz:] == -7.0)  # vertical halo untouched
        else:
            assert np.all(Th[h:h + nz, :h] == -7.0) and np.all(U[h:h + nz, h:h + ny, :h] == -7.0)


def test_perturbations_exact_and_bad_arguments(L):
    """rho = rho0 + rho' is a rounding of the sum; from_reference returns
    rho - rho0, which equals rho' exactly when the sum is exact (here: small
    dyadic values); NULL dry species are rejected, NULL moisture accepted."""
    nx, ny, nz, h = 4, 4, 3, 1
    d = {k: np.full(v.size, 0.25) for k, v in _dense(nx, ny, nz).items()}
    nc = L.lib().vibe_mi_reference_field_size(nx, ny, nz, h, 3)
    r0 = np.full(nc, 1.5)
    sp = _ref_arrays(L, nx, ny, nz, h)
    assert _to_ref(L, nx, ny, nz, h, d, r0, r0, r0, 0, sp) == 0
    sp2 = list(sp[:6]) + [None] * 6
    rc, back = _from_ref(L, nx, ny, nz, h, sp2, r0, r0, r0)
    assert rc == 0 and all(np.all(back[k] == 0.25) for k in back)
    bad = list(sp)
    bad[4] = None
    rc, _ = _from_ref(L, nx, ny, nz, h, bad, r0, r0, r0)
    assert rc == L.VIBE_C_ERR_BAD_ARG
    assert L.lib().vibe_mi_reference_field_size(0, 4, 3, 1, 0) == 0
