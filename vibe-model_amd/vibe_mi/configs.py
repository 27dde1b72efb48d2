"""Builders of the benchmark workloads (geometry, reference state, initial
state), using the library's host-side setup helpers.

  warm_bubble     dry 3-D warm bubble, 64x64x40, 250 m, neutral theta0 = 300 K,
                  cos^2 bubble 2 K / r = 2 km at z = 2 km, Rayleigh sponge in
                  the top 3 km, dt = 2 s, 6 substeps, 150 steps
                  (src/driver/initial_conditions.cpp:136-159 form)
  mountain_wave   512x512x80, 1 km, tanh-stretched levels to 20 km
                  (dz 50 -> 500 m), Witch-of-Agnesi ridge h = 1 km, a = 10 km,
                  U = 10 m/s, N = 0.01 1/s, theta_sfc = 300 K, seeded
                  +/-0.1 K noise below 1 km, dt = 6 s, 6 substeps, 600 steps
  mountain_wave_large  the same at 1024x1024x96
  ensemble        48 members of uneven size (nx, ny in 96..320 step 32,
                  nz in {40, 60, 80}), dx = 2 km, N = 0.01 1/s,
                  U(z) = 15 tanh(z / 3 km), 1-5 seeded warm bubbles each
                  (returns a list of cases for vibe_mi.Batch)
  nest_hierarchy  384x384x60 parent at 3 km, three ratio-3 nests at 1 km and
                  one at 333 m, random terrain, dt refined per level
                  (returns cases and nest links for vibe_mi.nesting.NestedRun)

Every size can be overridden so that tests run the same physics on small
grids.
"""
from __future__ import annotations

import numpy as np

from . import Case, _check, _pd, lib

P_SFC = 100000.0


def _levels(nz, ztop, kind, dz_bottom=None, dz_top=None):
    zeta = np.zeros(nz + 1)
    if kind == "uniform":
        _check(lib().vibe_mi_zeta_uniform(nz, _pd(zeta)), "zeta_uniform")
    else:
        _check(lib().vibe_mi_zeta_tanh(nz, ztop, dz_bottom, dz_top, _pd(zeta)), "zeta_tanh")
    return zeta


def _reference(kind, theta_sfc, bv, nx, ny, nz, ztop, zeta, zs):
    n = nx * ny * nz
    r0, t0, p0 = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(lib().vibe_mi_reference_state(kind, theta_sfc, bv, P_SFC, nx, ny, nz, ztop, _pd(zeta), _pd(zs),
                                         _pd(r0), _pd(t0), _pd(p0)), "reference_state")
    return r0, t0, p0


def _sponge(nz, ztop, zeta, zs00, start_fraction, alpha_max):
    a = np.zeros(nz)
    _check(lib().vibe_mi_sponge(nz, ztop, _pd(zeta), zs00, start_fraction, alpha_max, _pd(a)), "sponge")
    return a


def _zero_state(nx, ny, nz):
    cell = nx * ny * nz
    return {"u": np.zeros(cell), "v": np.zeros(cell), "w": np.zeros(nx * ny * (nz + 1)),
            "rho": np.zeros(cell), "theta": np.zeros(cell), "pi": np.zeros(cell)}


def _balance(st, r0, t0):
    _check(lib().vibe_mi_ic_balance_pi(r0.size, _pd(r0), _pd(t0), _pd(st["rho"]), _pd(st["theta"]),
                                       _pd(st["pi"])), "balance_pi")


def warm_bubble(nx=64, ny=64, nz=40, dx=250.0, dz=250.0, amp=2.0, radius=2000.0, zc=2000.0,
                sponge_depth=3000.0, alpha_max=0.2, dt=2.0, nsub=6, steps=150) -> Case:
    ztop = nz * dz
    zeta = _levels(nz, ztop, "uniform")
    zs = np.zeros(nx * ny)
    r0, t0, p0 = _reference(0, 300.0, 0.0, nx, ny, nz, ztop, zeta, zs)
    sp = _sponge(nz, ztop, zeta, 0.0, max(0.0, (ztop - sponge_depth) / ztop), alpha_max)
    st = _zero_state(nx, ny, nz)
    _check(lib().vibe_mi_ic_bubble(nx, ny, nz, dx, dx, ztop, _pd(zeta), _pd(zs), 0.5 * nx * dx, 0.5 * ny * dx, zc,
                                   radius, radius, amp, _pd(st["theta"])), "ic_bubble")
    _balance(st, r0, t0)
    return Case("warm_bubble", nx, ny, nz, dx, dx, ztop, zeta, zs, r0, t0, p0, sp, st, dt, nsub, steps)


def mountain_wave(nx=512, ny=512, nz=80, dx=1000.0, ztop=20000.0, dz_bottom=50.0, dz_top=500.0, h=1000.0,
                  a=10000.0, U=10.0, bv=0.01, theta_sfc=300.0, noise=0.1, noise_top=1000.0, seed=20240611,
                  sponge_start=0.8, alpha_max=0.2, dt=6.0, nsub=6, steps=600) -> Case:
    mean = ztop / nz
    if not dz_bottom < mean < dz_top:  # small test grids: keep the 1:10 stretching
        dz_bottom, dz_top = 0.2 * mean, 2.0 * mean
    zeta = _levels(nz, ztop, "tanh", dz_bottom, dz_top)
    zs = np.zeros(nx * ny)
    _check(lib().vibe_mi_terrain_agnesi_ridge(nx, ny, dx, h, a, _pd(zs)), "agnesi")
    r0, t0, p0 = _reference(1, theta_sfc, bv, nx, ny, nz, ztop, zeta, zs)
    sp = _sponge(nz, ztop, zeta, float(zs[0]), sponge_start, alpha_max)
    st = _zero_state(nx, ny, nz)
    _check(lib().vibe_mi_ic_wind(nx, ny, nz, ztop, _pd(zeta), _pd(zs), U, 0.0, 0.0, _pd(st["u"])), "ic_wind")
    if noise > 0:
        _check(lib().vibe_mi_ic_noise(nx, ny, nz, ztop, _pd(zeta), _pd(zs), noise, noise_top, seed,
                                      _pd(st["theta"])), "ic_noise")
    _balance(st, r0, t0)
    c = Case("mountain_wave", nx, ny, nz, dx, dx, ztop, zeta, zs, r0, t0, p0, sp, st, dt, nsub, steps)
    c.meta["seed"] = seed
    return c


def mountain_wave_large(**kw) -> Case:
    args = dict(nx=1024, ny=1024, nz=96)
    args.update(kw)
    c = mountain_wave(**args)
    c.name = "mountain_wave_large"
    return c


def random_terrain_case(nx=64, ny=48, nz=24, dx=2000.0, ztop=16000.0, sigma=300.0, L=20000.0, seed=7,
                        U=8.0, bv=0.01, dt=6.0, nsub=6, steps=10) -> Case:
    """Small stable-stratified case over seeded 2-D random terrain (exercises
    both metric slopes)."""
    zeta = _levels(nz, ztop, "tanh", 0.4 * ztop / nz, 2.0 * ztop / nz)
    zs = np.zeros(nx * ny)
    _check(lib().vibe_mi_terrain_gaussian(nx, ny, dx, dx, sigma, L, seed, _pd(zs)), "terrain_gaussian")
    r0, t0, p0 = _reference(1, 300.0, bv, nx, ny, nz, ztop, zeta, zs)
    sp = _sponge(nz, ztop, zeta, float(zs[0]), 0.75, 0.1)
    st = _zero_state(nx, ny, nz)
    _check(lib().vibe_mi_ic_wind(nx, ny, nz, ztop, _pd(zeta), _pd(zs), U, 0.0, 0.0, _pd(st["u"])), "ic_wind")
    # v also blows so that both horizontal metric terms act
    st["v"][:] = 0.5 * U
    _check(lib().vibe_mi_ic_bubble(nx, ny, nz, dx, dx, ztop, _pd(zeta), _pd(zs), 0.5 * nx * dx, 0.5 * ny * dx,
                                   3000.0, 8000.0, 2000.0, 1.5, _pd(st["theta"])), "ic_bubble")
    _check(lib().vibe_mi_ic_noise(nx, ny, nz, ztop, _pd(zeta), _pd(zs), 0.2, 2000.0, seed + 1,
                                  _pd(st["theta"])), "ic_noise")
    _balance(st, r0, t0)
    return Case("random_terrain", nx, ny, nz, dx, dx, ztop, zeta, zs, r0, t0, p0, sp, st, dt, nsub, steps)


def ensemble_member(nx, ny, nz, seed, dx=2000.0, ztop=20000.0, bv=0.01, theta_sfc=300.0, umax=15.0,
                    hshear=3000.0, dt=8.0, nsub=6, steps=300, name="ensemble_member") -> Case:
    """One ensemble member: flat terrain, uniform levels, constant-N base
    state, U(z) = umax tanh(z / hshear), 1-5 cos^2 warm bubbles of 1-3 K and
    2-6 km radius at seeded positions (numpy PCG64 seeded with `seed`)."""
    rng = np.random.default_rng(seed)
    zeta = _levels(nz, ztop, "uniform")
    zs = np.zeros(nx * ny)
    r0, t0, p0 = _reference(1, theta_sfc, bv, nx, ny, nz, ztop, zeta, zs)
    sp = _sponge(nz, ztop, zeta, 0.0, 0.75, 0.1)
    st = _zero_state(nx, ny, nz)
    _check(lib().vibe_mi_ic_wind(nx, ny, nz, ztop, _pd(zeta), _pd(zs), 0.0, umax, hshear, _pd(st["u"])), "ic_wind")
    nb = int(rng.integers(1, 6))
    bubbles = []
    for _ in range(nb):
        amp = float(rng.uniform(1.0, 3.0))
        r = float(rng.uniform(2000.0, 6000.0))
        xc = float(rng.uniform(0.0, nx * dx))
        yc = float(rng.uniform(0.0, ny * dx))
        zc = float(rng.uniform(2000.0, 6000.0))
        bubbles.append((xc, yc, zc, r, amp))
        _check(lib().vibe_mi_ic_bubble(nx, ny, nz, dx, dx, ztop, _pd(zeta), _pd(zs), xc, yc, zc, r, r, amp,
                                       _pd(st["theta"])), "ic_bubble")
    _balance(st, r0, t0)
    c = Case(name, nx, ny, nz, dx, dx, ztop, zeta, zs, r0, t0, p0, sp, st, dt, nsub, steps)
    c.meta.update(seed=seed, bubbles=bubbles)
    return c


def ensemble(n=48, seed=20240611, sizes=tuple(range(96, 321, 32)), levels=(40, 60, 80), **kw):
    """n members of uneven size: nx, ny drawn per axis from `sizes`, nz from
    `levels` (PCG64 seeded with `seed`; member m uses seed + 1 + m)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        nx = int(rng.choice(sizes))
        ny = int(rng.choice(sizes))
        nz = int(rng.choice(levels))
        out.append(ensemble_member(nx, ny, nz, seed + 1 + m, name=f"ensemble{m:02d}", **kw))
    return out


def _bilinear_terrain(zs_p, pnx, pny, cnx, cny, r, i0, j0):
    """Child terrain at cell centres, bilinear in the (periodic) parent
    terrain, with the nest placement of vibe_mi/nest.h."""
    Z = zs_p.reshape(pny, pnx)
    i = np.arange(cnx)
    j = np.arange(cny)
    xp = (i0 + i + 0.5) / r - 0.5
    yp = (j0 + j + 0.5) / r - 0.5
    ia, ja = np.floor(xp).astype(int), np.floor(yp).astype(int)
    wx, wy = xp - ia, yp - ja
    i0w, i1w = ia % pnx, (ia + 1) % pnx
    j0w, j1w = ja % pny, (ja + 1) % pny
    q00 = Z[np.ix_(j0w, i0w)]
    q10 = Z[np.ix_(j0w, i1w)]
    q01 = Z[np.ix_(j1w, i0w)]
    q11 = Z[np.ix_(j1w, i1w)]
    WX = wx[None, :]
    WY = wy[:, None]
    lo = q00 + WX * (q10 - q00)
    hi = q01 + WX * (q11 - q01)
    return np.ascontiguousarray((lo + WY * (hi - lo)).ravel())


def _nest_domain(name, nx, ny, nz, dx, ztop, zeta, zs, theta_sfc, bv, U, dt, nsub, steps, alpha):
    r0, t0, p0 = _reference(1, theta_sfc, bv, nx, ny, nz, ztop, zeta, zs)
    sp = _sponge(nz, ztop, zeta, float(zs[0]), 0.8, alpha)
    st = _zero_state(nx, ny, nz)
    _check(lib().vibe_mi_ic_wind(nx, ny, nz, ztop, _pd(zeta), _pd(zs), U, 0.0, 0.0, _pd(st["u"])), "ic_wind")
    return Case(name, nx, ny, nz, dx, dx, ztop, zeta, zs, r0, t0, p0, sp, st, dt, nsub, steps)


def nest_hierarchy(pnx=384, pny=384, nz=60, dx=3000.0, ztop=20000.0, sigma=300.0, L=20000.0, seed=20240611,
                   theta_amp=0.5, U=10.0, bv=0.01, theta_sfc=300.0, dt=18.0, nsub=6, steps=300, nests=None,
                   sponge_alpha=0.1):
    """Parent 384x384x60 at 3 km with three ratio-3 nests (271x271, 211x331,
    151x151 at 1 km) and one second-level nest (181x181 at 333 m inside the
    first), sharing one 60-level stretched vertical coordinate.  Seeded
    Gaussian random terrain (sigma 300 m, correlation length 20 km) on the
    parent; nest terrain is the bilinear interpolation of their parent's.
    Stable base state N = 0.01 1/s, uniform U, random-phase theta' of
    amplitude theta_amp below 4 km on the parent.  Nest states start as the
    parent's interpolation (vibe_mi.nesting.NestedRun does that on the
    device).  Time steps refine with the grid: dt, dt/3, dt/9.

    Returns (cases, links) with links = [(child, parent, ratio, i0, j0)],
    i0 / j0 counted in child cells."""
    if nests is None:
        # (nx, ny, parent, offset in parent cells)
        nests = [(271, 271, 0, (40, 40)), (211, 331, 0, (200, 30)), (151, 151, 0, (150, 250)),
                 (181, 181, 1, (45, 45))]
    r = 3
    zeta = _levels(nz, ztop, "tanh", 0.3 * ztop / nz, 1.8 * ztop / nz)
    zs = np.zeros(pnx * pny)
    _check(lib().vibe_mi_terrain_gaussian(pnx, pny, dx, dx, sigma, L, seed, _pd(zs)), "terrain_gaussian")
    parent = _nest_domain("nest_parent", pnx, pny, nz, dx, ztop, zeta, zs, theta_sfc, bv, U, dt, nsub, steps,
                          sponge_alpha)
    # random-phase theta' on the parent: 16 modes, integer wavenumbers 1..8
    rng = np.random.default_rng(seed)
    x = (np.arange(pnx) + 0.5) / pnx
    y = (np.arange(pny) + 0.5) / pny
    X, Y = np.meshgrid(x, y)
    plane = np.zeros_like(X)
    for _ in range(16):
        kx, ky = rng.integers(1, 9, size=2)
        plane += np.cos(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
    plane *= theta_amp / np.max(np.abs(plane))
    zc = 0.5 * (zeta[:-1] + zeta[1:])
    Z = zs.reshape(pny, pnx)[None, :, :] + zc[:, None, None] * (ztop - zs.reshape(pny, pnx))[None, :, :]
    env = np.where(Z < 2000.0, 1.0, np.where(Z < 4000.0, np.cos(0.5 * np.pi * (Z - 2000.0) / 2000.0) ** 2, 0.0))
    parent.state["theta"] = np.ascontiguousarray((env * plane[None, :, :]).ravel())
    _balance(parent.state, parent.rho0, parent.theta0)
    parent.meta.update(seed=seed, level=0)
    cases, links = [parent], []
    for q, (nx, ny, par, (pi0, pj0)) in enumerate(nests):
        P = cases[par]
        czs = _bilinear_terrain(P.zs, P.nx, P.ny, nx, ny, r, r * pi0, r * pj0)
        level = P.meta["level"] + 1
        c = _nest_domain(f"nest{q + 1}", nx, ny, nz, P.dx / r, ztop, zeta, czs, theta_sfc, bv, U, P.dt / r, nsub,
                         steps * r ** level, sponge_alpha)
        c.meta.update(level=level)
        cases.append(c)
        links.append((q + 1, par, r, r * pi0, r * pj0))
    return cases, links


CONFIGS = {
    "warm_bubble": warm_bubble,
    "mountain_wave": mountain_wave,
    "mountain_wave_large": mountain_wave_large,
    "random_terrain": random_terrain_case,
}
