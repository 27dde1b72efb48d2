"""ctypes bindings of the CPU oracle (oracle/build/liboracle.so) and of the
reference's own compiled C kernels (oracle/_ref/libref_ckernels.so).

Test infrastructure: used only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_ckernels.so")

_D = ctypes.c_double
_I = ctypes.c_int
_SZ = ctypes.c_size_t
_PD = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


class OrcDomain(ctypes.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I), ("dx", _D), ("dy", _D), ("ztop", _D),
                ("zeta", _PD), ("zs", _PD), ("rho0", _PD), ("theta0", _PD), ("pi0", _PD),
                ("sponge", _PD), ("eps_s", _D)]


class OrcState(ctypes.Structure):
    _fields_ = [("u", _PD), ("v", _PD), ("w", _PD), ("rho", _PD), ("theta", _PD), ("pi", _PD)]


ORC_PHYSICS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                  ctypes.POINTER(OrcState), _PD, _PD, _PD)


def _pd(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


_olib = None


def oracle_lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError("oracle not built: make -C oracle")
        L = ctypes.CDLL(ORACLE_LIB)
        L.orc_tridiagonal_solve.restype = _I
        L.orc_tridiagonal_solve.argtypes = [_PD, _PD, _PD, _PD, _I, _PD, _PD]
        L.orc_weno5.restype = None
        L.orc_weno5.argtypes = [_D] * 5 + [_PD, _PD]
        L.orc_bilinear.restype = _D
        L.orc_bilinear.argtypes = [_PD, _I, _I, _I, _I, _D, _D, _I]
        L.orc_cfl_scan.restype = _D
        L.orc_cfl_scan.argtypes = [_PD, _PD, _PD, _PD, _PD, _I, _I, _I, _D, _D, _D, _D, _PD, _PD]
        L.orc_neumaier_sum.restype = _D
        L.orc_neumaier_sum.argtypes = [_PD, _SZ]
        L.orc_advect_scalar_flux.restype = None
        L.orc_advect_scalar_flux.argtypes = [_PD] * 7 + [_I, _I, _I, _D, _D, _D, _PD]
        L.orc_dycore_create.restype = _VP
        L.orc_dycore_create.argtypes = [ctypes.POINTER(OrcDomain)]
        L.orc_dycore_destroy.restype = None
        L.orc_dycore_destroy.argtypes = [_VP]
        L.orc_dycore_step.restype = _I
        L.orc_dycore_step.argtypes = [_VP, ctypes.POINTER(OrcState), _D, _I]
        L.orc_dycore_set_physics.restype = None
        L.orc_dycore_set_physics.argtypes = [_VP, _PD, _PD, _PD]
        L.orc_dycore_set_physics_callback.restype = None
        L.orc_dycore_set_physics_callback.argtypes = [_VP, ORC_PHYSICS_FN, _VP]
        L.orc_dycore_slow_tendency.restype = None
        L.orc_dycore_slow_tendency.argtypes = [_VP, ctypes.POINTER(OrcState), ctypes.POINTER(OrcState)]
        L.orc_dycore_substep.restype = None
        L.orc_dycore_substep.argtypes = [_VP, ctypes.POINTER(OrcState), _D]
        L.orc_dycore_sponge.restype = None
        L.orc_dycore_sponge.argtypes = [_VP, ctypes.POINTER(OrcState), _D]
        L.orc_dycore_energy.restype = None
        L.orc_dycore_energy.argtypes = [_VP, ctypes.POINTER(OrcState), _PD]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [_I]
        L.orc_plant.restype = None
        L.orc_plant.argtypes = [_I]
        L.orc_weno_face.restype = _D
        L.orc_weno_face.argtypes = [_D] * 5
        L.orc_upwind_weno.restype = _D
        L.orc_upwind_weno.argtypes = [_D] * 7
        L.orc_up5.restype = _D
        L.orc_up5.argtypes = [_D] * 7
        L.orc_davies_profile.restype = None
        L.orc_davies_profile.argtypes = [_I, _D, _PD]
        L.orc_nest_restrict.restype = None
        L.orc_nest_restrict.argtypes = [ctypes.POINTER(OrcState), _I, _I, ctypes.POINTER(OrcState), _I, _I, _I, _I,
                                        _I, _I]
        L.orc_nest_apply.restype = None
        L.orc_nest_apply.argtypes = [ctypes.POINTER(OrcState), ctypes.POINTER(OrcState), _I, _I,
                                     ctypes.POINTER(OrcState), _I, _I, _I, _I, _I, _I, _I, _D, _D, _I]
        _olib = L
    return _olib


def ref_lib():
    """The reference's own c_kernels.c compiled by oracle/ref.mk, or None."""
    if not os.path.exists(REF_LIB):
        return None
    L = ctypes.CDLL(REF_LIB)
    L.vibe_tridiagonal_solve_f64.restype = _I
    L.vibe_tridiagonal_solve_f64.argtypes = [_PD, _PD, _PD, _PD, _I, _PD, _PD]
    L.vibe_weno5_reconstruct_f64.restype = None
    L.vibe_weno5_reconstruct_f64.argtypes = [_D] * 5 + [_PD, _PD]
    L.vibe_bilinear_f64.restype = _D
    L.vibe_bilinear_f64.argtypes = [_PD, _I, _I, _I, _I, _D, _D, _I]
    L.vibe_cfl_scan_f64.restype = _D
    L.vibe_cfl_scan_f64.argtypes = [_PD, _PD, _PD, _PD, _PD, _I, _I, _I, _D, _D, _D, _D, _PD, _PD]
    L.vibe_neumaier_sum_f64.restype = _D
    L.vibe_neumaier_sum_f64.argtypes = [_PD, _SZ]
    L.vibe_advect_scalar_flux_f64.restype = None
    L.vibe_advect_scalar_flux_f64.argtypes = [_PD] * 7 + [_I, _I, _I, _D, _D, _D, _PD]
    return L


FIELDS = ("u", "v", "w", "rho", "theta", "pi")


class planted:
    """Context manager: the oracle runs with one planted mistake (orc_plant
    codes in oracle/oracle.h), restored to the true equations on exit."""

    BUOYANCY_SIGN, GAMMA_HALF, METRIC_SIGN, OMEGA_SIGN, OMEGA_ZETA = 1, 2, 3, 4, 5

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        oracle_lib().orc_plant(self.which)
        return self

    def __exit__(self, *exc):
        oracle_lib().orc_plant(0)
        return False


class OracleDycore:
    """FP64 CPU oracle of one domain (mirrors vibe_mi.Domain)."""

    def __init__(self, case, threads: int = 0):
        L = oracle_lib()
        self.case = case
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (case.zeta, case.zs, case.rho0, case.theta0, case.pi0, case.sponge)]
        z, zs, r0, t0, p0, sp = self._keep
        self.dom = OrcDomain(case.nx, case.ny, case.nz, case.dx, case.dy, case.ztop, _pd(z), _pd(zs), _pd(r0),
                             _pd(t0), _pd(p0), _pd(sp), case.eps_s)
        self.h = L.orc_dycore_create(ctypes.byref(self.dom))
        if threads:
            L.orc_set_threads(threads)
        self.state = case.copy_state()

    def close(self):
        if self.h:
            oracle_lib().orc_dycore_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _st(self, st):
        return OrcState(*[_pd(st[k]) for k in FIELDS])

    def step(self, dt, nsub, nsteps=1):
        s = self._st(self.state)
        for _ in range(nsteps):
            rc = oracle_lib().orc_dycore_step(self.h, ctypes.byref(s), dt, nsub)
            if rc:
                return rc
        return 0

    def slow_tendency(self):
        c = self.case
        cell = c.nx * c.ny * c.nz
        R = {k: np.zeros(c.nx * c.ny * (c.nz + 1) if k == "w" else cell) for k in FIELDS}
        oracle_lib().orc_dycore_slow_tendency(self.h, ctypes.byref(self._st(self.state)), ctypes.byref(self._st(R)))
        return R

    def substep(self, dtau):
        oracle_lib().orc_dycore_substep(self.h, ctypes.byref(self._st(self.state)), dtau)

    def energy(self):
        out = np.zeros(6)
        oracle_lib().orc_dycore_energy(self.h, ctypes.byref(self._st(self.state)), _pd(out))
        return out

    def set_physics(self, theta=None, u=None, v=None):
        """Fixed physics rates (orc_dycore_set_physics); all None: off."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (theta, u, v)]
        self._phys_keep = arrs
        oracle_lib().orc_dycore_set_physics(self.h, *[_pd(a) for a in arrs])

    def set_physics_callback(self, fn):
        """fn(stage, dt, state dict, rates dict) before each stage's slow
        tendency (orc_dycore_set_physics_callback); None removes it."""
        L = oracle_lib()
        if fn is None:
            self._phys_cb = None
            L.orc_dycore_set_physics_callback(self.h, ORC_PHYSICS_FN(), None)
            return
        c = self.case
        cell = c.nx * c.ny * c.nz
        nwf = c.nx * c.ny * (c.nz + 1)

        def arr(p, n):
            return np.ctypeslib.as_array(ctypes.cast(p, _PD), shape=(n,))

        def tramp(_user, stage, dt, sp, rt, ru, rv):
            try:
                s = sp.contents
                st = {"u": arr(s.u, cell), "v": arr(s.v, cell), "w": arr(s.w, nwf), "rho": arr(s.rho, cell),
                      "theta": arr(s.theta, cell), "pi": arr(s.pi, cell)}
                rates = {"theta": arr(rt, cell), "u": arr(ru, cell), "v": arr(rv, cell)}
                return int(fn(stage, dt, st, rates) or 0)
            except Exception:
                return 1

        self._phys_cb = ORC_PHYSICS_FN(tramp)
        L.orc_dycore_set_physics_callback(self.h, self._phys_cb, None)


FIELDS = ("u", "v", "w", "rho", "theta", "pi")


def _ostate(st):
    arrs = {k: np.ascontiguousarray(st[k], dtype=np.float64) for k in FIELDS}
    return OrcState(*[_pd(arrs[k]) for k in FIELDS]), arrs


def oracle_nest_apply(pnew, pold, pnx, pny, child, cnx, cny, nz, ratio, i0, j0, zone, alpha_max, frac, full):
    """Oracle nest operator on dict states; returns the new child state."""
    out = {k: np.array(child[k], dtype=np.float64, copy=True) for k in FIELDS}
    sn, _kn = _ostate(pnew)
    so, _ko = _ostate(pold if pold is not None else pnew)
    sc = OrcState(*[_pd(out[k]) for k in FIELDS])
    oracle_lib().orc_nest_apply(ctypes.byref(sn), ctypes.byref(so), pnx, pny, ctypes.byref(sc), cnx, cny, nz,
                                ratio, i0, j0, zone, alpha_max, frac, full)
    return out


def oracle_nest_restrict(child, cnx, cny, parent, pnx, pny, nz, ratio, i0, j0):
    """Oracle two-way feedback on dict states; returns the new parent state."""
    out = {k: np.array(parent[k], dtype=np.float64, copy=True) for k in FIELDS}
    sc, _kc = _ostate(child)
    sp = OrcState(*[_pd(out[k]) for k in FIELDS])
    oracle_lib().orc_nest_restrict(ctypes.byref(sc), cnx, cny, ctypes.byref(sp), pnx, pny, nz, ratio, i0, j0)
    return out
