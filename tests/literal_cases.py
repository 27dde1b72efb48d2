"""Inputs of the reference's literal-equation golden vectors
(tests/golden/equations_reference.npz, written by
tests/golden/make_equations_golden.py from oracle/ref_equations.cpp), rebuilt
bit for bit from the integer hash the driver uses, and ctypes bindings of the
oracle's literal restatement (oracle/literal.c).

Test infrastructure: used only by tests/.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from oracle_bind import _D, _I, _PD, oracle_lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "equations_reference.npz")
SPECIES = ("u", "v", "w", "rho", "theta", "pi", "qv", "qc", "qr", "qi", "qs", "qg")
NX, NY, NZ = 8, 6, 5
DX, DY, ZTOP = 1000.0, 1500.0, 12000.0
# tag: (grid, advection order, coriolis mode, latitude, damping, damping order, moist)
CASES = {
    "c2": ("hill", 2, 1, 45.0, 0.0, 2, 1),
    "c4": ("hill", 4, 0, 45.0, 2e5, 2, 0),
    "c6": ("flat", 6, 2, 30.0, 3e9, 4, 1),
}


class LiteralDesc(ctypes.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I), ("dx", _D), ("dy", _D), ("z_top", _D), ("zeta", _PD),
                ("zs", _PD), ("rho0", _PD), ("theta0", _PD), ("pi0", _PD), ("order", _I), ("coriolis", _I),
                ("latitude", _D), ("div_coeff", _D), ("div_order", _I), ("moist", _I)]


def shape(q, nx=NX, ny=NY, nz=NZ):
    """Dense (nz, ny, nx) shape of species q with its stagger."""
    return (nz + (q == 2), ny + (q == 1), nx + (q == 0))


def golden():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def hash_r(q, sh):
    """r(i, j, k, s) = h / 2^20 - 0.5 of oracle/ref_equations.cpp over a
    (nz, ny, nx) index box."""
    k, j, i = np.meshgrid(*[np.arange(n, dtype=np.uint64) for n in sh], indexing="ij")
    h = (((i + np.uint64(7)) * np.uint64(73856093)) ^ ((j + np.uint64(11)) * np.uint64(19349663)) ^
         ((k + np.uint64(13)) * np.uint64(83492791)) ^ (np.uint64(q + 3) * np.uint64(2654435761))) & np.uint64(0xFFFFF)
    return h.astype(np.float64) / 1048576.0 - 0.5


def state(g, grid):
    """The 12 species of the driver's hashed state on grid 'hill' or 'flat'."""
    r0, t0, p0 = (g[f"{grid}_{n}"].reshape(NZ, NY, NX) for n in ("rho0", "theta0", "pi0"))
    s = []
    for q in range(12):
        r = hash_r(q, shape(q))
        if q == 0:
            a = 8.0 + 6.0 * r
        elif q == 1:
            a = -3.0 + 5.0 * r
        elif q == 2:
            a = 0.8 * r
        elif q == 3:
            a = r0 + 0.02 * r
        elif q == 4:
            a = t0 + 3.0 * r
        elif q == 5:
            a = p0 + 0.002 * r
        elif q == 6:
            a = np.maximum(0.0, 0.01 * (r + 0.2))
        else:
            a = 0.002 * (r + 0.1)
        s.append(np.ascontiguousarray(a.ravel()))
    return s


def desc(g, tag, keep):
    grid, order, cor, lat, damp, dord, moist = CASES[tag]
    arrs = [np.ascontiguousarray(g["zeta"]), np.ascontiguousarray(g["zs_hill"]) if grid == "hill" else None,
            np.ascontiguousarray(g[f"{grid}_rho0"]), np.ascontiguousarray(g[f"{grid}_theta0"]),
            np.ascontiguousarray(g[f"{grid}_pi0"])]
    keep.extend(a for a in arrs if a is not None)
    p = [None if a is None else a.ctypes.data_as(_PD) for a in arrs]
    return LiteralDesc(NX, NY, NZ, DX, DY, ZTOP, *p, order, cor, lat, damp, dord, moist)


def ptrs(arrs):
    return (_PD * 12)(*[a.ctypes.data_as(_PD) for a in arrs])


def _bind():
    L = oracle_lib()
    if not getattr(L, "_literal_bound", False):
        L.orc_literal_tendencies.restype = _I
        L.orc_literal_tendencies.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD), ctypes.POINTER(_PD)]
        L.orc_literal_acoustic.restype = _I
        L.orc_literal_acoustic.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD), ctypes.POINTER(_PD)]
        L.orc_literal_max_sound_speed.restype = _D
        L.orc_literal_max_sound_speed.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD)]
        L.orc_literal_step.restype = _I
        L.orc_literal_step.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD), _D, _I, _PD]
        L.orc_literal_diagnose.restype = None
        L.orc_literal_diagnose.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD), _PD]
        L.orc_si_assemble.restype = None
        L.orc_si_assemble.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD), ctypes.POINTER(_PD), _D, _PD,
                                      _PD]
        L.orc_si_step.restype = _I
        L.orc_si_step.argtypes = [ctypes.POINTER(LiteralDesc), ctypes.POINTER(_PD), _D, _PD, _D, _I,
                                  ctypes.POINTER(_PD)]
        L._literal_bound = True
    return L


def oracle_tendencies(d, s, acoustic=False):
    L = _bind()
    t = [np.zeros(a.size) for a in s]
    fn = L.orc_literal_acoustic if acoustic else L.orc_literal_tendencies
    assert fn(ctypes.byref(d), ptrs(s), ptrs(t)) == 0
    return t


def oracle_max_sound_speed(d, s):
    return _bind().orc_literal_max_sound_speed(ctypes.byref(d), ptrs(s))


def oracle_diagnose(d, s):
    out = np.zeros(s[5].size)
    _bind().orc_literal_diagnose(ctypes.byref(d), ptrs(s), out.ctypes.data_as(_PD))
    return out


def oracle_step(d, s, dt, nsub, sponge=None, nsteps=1):
    """nsteps literal RK3 steps of s in place; returns the last return code."""
    L = _bind()
    sp = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float64)
    rc = 0
    for _ in range(nsteps):
        rc = L.orc_literal_step(ctypes.byref(d), ptrs(s), dt, nsub, None if sp is None else sp.ctypes.data_as(_PD))
        if rc:
            break
    return rc


def oracle_si_assemble(d, s, n, dt):
    """(a, rhs) of SemiImplicitIntegrator::assemble_rhs for state s and
    tendency n (oracle/literal.c)."""
    nc = s[3].size
    a, r = np.zeros(nc), np.zeros(nc)
    _bind().orc_si_assemble(ctypes.byref(d), ptrs(s), ptrs(n), dt, a.ctypes.data_as(_PD), r.ctypes.data_as(_PD))
    return a, r


def oracle_si_step(d, s, dt, sponge=None, time_filter=0.0, first=True, phys=None, nsteps=1):
    """nsteps state updates of SemiImplicitIntegrator::step on s in place
    (the first of them with the filter's first-step flag `first`); returns
    the last return code."""
    L = _bind()
    sp = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float64)
    ph = None
    if phys is not None:
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in phys]
        ph = (_PD * 4)(*[None if a is None else a.ctypes.data_as(_PD) for a in keep])
    rc = 0
    for n in range(nsteps):
        rc = L.orc_si_step(ctypes.byref(d), ptrs(s), dt, None if sp is None else sp.ctypes.data_as(_PD), time_filter,
                           1 if (first and n == 0) else 0, ph)
        if rc:
            break
    return rc
