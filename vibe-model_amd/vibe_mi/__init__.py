"""vibe_mi -- Python host side of the MI355X dynamical-core library.

Thin ctypes bindings over libvibe_mi.so (the C ABI declared in
include/vibe_mi/c_kernels.h and include/vibe_mi/dycore.h).  The Python
names mirror the reference's interfaces:

  * ``tridiagonal_solve``, ``weno5_reconstruct``, ``bilinear``, ``cfl_scan``,
    ``neumaier_sum``, ``advect_scalar_flux`` -- the C11 kernels of the
    reference's include/vibe/common/c_kernels.h:54-113, same argument meaning
    and return codes;
  * ``Domain`` with ``step(dt, substeps)`` -- the role of
    ``RungeKutta3Integrator::step(State&, StepContext)``
    (include/vibe/time/integrator.hpp:62-82, src/time/runge_kutta.cpp:77-118),
    ``slow_tendency`` (Equations::tendencies, src/dyn/equations.cpp:60-150),
    ``acoustic_substep`` (src/time/runge_kutta.cpp:65-75) and
    ``energy_budget`` (src/dyn/diagnostics.cpp:255-277);
  * ``Helmholtz`` -- the 3-D Helmholtz solvers of the semi-implicit path
    (HelmholtzSolver, src/time/helmholtz.cpp:173-476): BiCGSTAB and FP32
    solves with FP64 iterative refinement (include/vibe_mi/helmholtz.h);
  * ``vibebin_*`` and ``Domain.write_restart`` / ``read_restart`` -- the
    reference's .vibebin BinaryWriter / BinaryReader and write_restart /
    read_restart (src/io/binary_writer.cpp, src/io/field_io.cpp:207-250),
    same bytes on disk (include/vibe_mi/vibebin.h).

Every compute call runs on the GPU; if the library is missing or no GPU is
visible the calls raise instead of computing anything on the CPU.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VIBE_MI_LIB names another build of the library (e.g. one compiled with other
# tile shapes); the default is the in-tree build
LIB_PATH = os.environ.get("VIBE_MI_LIB") or os.path.join(_HERE, "lib", "libvibe_mi.so")

VIBE_C_OK = 0
VIBE_C_ERR_SINGULAR = 1
VIBE_C_ERR_BAD_ARG = 2
VIBE_MI_ERR_HIP = 100
VIBE_MI_ERR_NO_DEVICE = 101
VIBE_MI_ERR_NONFINITE = 102
VIBE_MI_ERR_PHYSICS = 103
VIBE_MI_ERR_TOO_LARGE = 104
VIBE_MI_ERR_IO = 110
VIBE_MI_ERR_FORMAT = 111
VIBE_MI_ERR_NOT_FOUND = 112
FP64 = 0
MIXED = 1
MIXED_COMP = 2  # FP32-mixed with u compensated (FP32 + 8-bit remainder) and the x part of the horizontal divergence in FP64


class VibeError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


class NumericalError(VibeError):
    """The state became non-finite (runge_kutta.cpp:103-106 role)."""


class IoError(VibeError):
    """A .vibebin file could not be opened, parsed or written, or lacks a
    record (vibe::IoError role, include/vibe/common/error.hpp:41)."""


_D = ctypes.c_double
_I = ctypes.c_int
_SZ = ctypes.c_size_t
_PD = ctypes.POINTER(ctypes.c_double)
_PF = ctypes.POINTER(ctypes.c_float)
_VP = ctypes.c_void_p
# int fn(user, stage, dt, stream, u, v, w, rho', theta', pi', theta_rate, u_rate, v_rate)
PHYSICS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                              *([ctypes.c_void_p] * 9))


class DomainDesc(ctypes.Structure):
    _fields_ = [
        ("nx", _I), ("ny", _I), ("nz", _I),
        ("dx", _D), ("dy", _D), ("ztop", _D),
        ("zeta", _PD), ("zs", _PD), ("rho0", _PD), ("theta0", _PD), ("pi0", _PD),
        ("sponge", _PD), ("eps_s", _D), ("precision", _I),
    ]


class HelmholtzDesc(ctypes.Structure):
    """vibe_mi_helmholtz_desc (vibe_mi/helmholtz.h)."""
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I), ("dx", _D), ("dy", _D), ("ztop", _D),
                ("zeta", _PD), ("zs", _PD)]


class LiteralDesc(ctypes.Structure):
    """vibe_mi_literal_desc (vibe_mi/literal.h)."""
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I), ("dx", _D), ("dy", _D), ("z_top", _D), ("zeta", _PD),
                ("zs", _PD), ("rho0", _PD), ("theta0", _PD), ("pi0", _PD), ("order", _I), ("coriolis", _I),
                ("latitude", _D), ("div_coeff", _D), ("div_order", _I), ("moist", _I)]


class SiDesc(ctypes.Structure):
    """vibe_mi_si_desc (vibe_mi/semi_implicit.h)."""
    _fields_ = [("eq", LiteralDesc), ("sponge", _PD), ("solver", _I), ("tol", _D), ("max_iter", _I),
                ("time_filter", _D)]


class SiStats(ctypes.Structure):
    """vibe_mi_si_stats (IntegratorStats, integrator.hpp:53-59)."""
    _fields_ = [("steps", ctypes.c_longlong), ("acoustic_substeps", ctypes.c_longlong), ("wall_time", _D),
                ("helmholtz_iterations", _I), ("helmholtz_residual", _D), ("diverged", _I)]


class NestDesc(ctypes.Structure):
    """vibe_mi_nest_desc (vibe_mi/nest.h)."""
    _fields_ = [("ratio", _I), ("i0", _I), ("j0", _I), ("zone", _I), ("alpha_max", _D)]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libvibe_mi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "vibe_tridiagonal_solve_f64": (_I, [_PD, _PD, _PD, _PD, _I, _PD, _PD]),
        "vibe_tridiagonal_solve_f32": (_I, [_PF, _PF, _PF, _PF, _I, _PF, _PF]),
        "vibe_weno5_reconstruct_f64": (None, [_D] * 5 + [_PD, _PD]),
        "vibe_weno5_reconstruct_f32": (None, [ctypes.c_float] * 5 + [_PF, _PF]),
        "vibe_bilinear_f64": (_D, [_PD, _I, _I, _I, _I, _D, _D, _I]),
        "vibe_cfl_scan_f64": (_D, [_PD, _PD, _PD, _PD, _PD, _I, _I, _I, _D, _D, _D, _D, _PD, _PD]),
        "vibe_neumaier_sum_f64": (_D, [_PD, _SZ]),
        "vibe_neumaier_sum_f32": (_D, [_PF, _SZ]),
        "vibe_advect_scalar_flux_f64": (None, [_PD, _PD, _PD, _PD, _PD, _PD, _PD, _I, _I, _I, _D, _D, _D, _PD]),
        "vibe_mi_tridiagonal_solve_batched_f64": (_I, [_VP, _VP, _VP, _VP, _I, _I, _VP, _VP, _VP]),
        "vibe_mi_tridiagonal_solve_batched_noalloc_f64": (_I, [_VP, _VP, _VP, _VP, _I, _I, _VP, _VP]),
        "vibe_mi_noalloc_scratch_capacity": (_SZ, []),
        "vibe_mi_weno5_reconstruct_batched_f64": (_I, [_VP, _SZ, _VP, _VP, _VP]),
        "vibe_mi_neumaier_sum_device_f64": (_I, [_VP, _SZ, _PD, _VP]),
        "vibe_mi_last_error": (_I, []),
        "vibe_mi_domain_create": (_I, [ctypes.POINTER(DomainDesc), ctypes.POINTER(_VP)]),
        "vibe_mi_domain_destroy": (_I, [_VP]),
        "vibe_mi_domain_set_state": (_I, [_VP] + [_PD] * 6),
        "vibe_mi_domain_get_state": (_I, [_VP] + [_PD] * 6),
        "vibe_mi_domain_step": (_I, [_VP, _D, _I, _I]),
        "vibe_mi_domain_step_timed": (_I, [_VP, _D, _I, _I, _PD]),
        "vibe_mi_domain_kernel_times": (_I, [_VP, _PD, _PD, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "vibe_mi_domain_slow_tendency": (_I, [_VP] + [_PD] * 6),
        "vibe_mi_domain_substep": (_I, [_VP, _D]),
        "vibe_mi_domain_set_physics_tendency": (_I, [_VP, _PD, _PD, _PD]),
        "vibe_mi_domain_physics_buffers": (_I, [_VP] + [ctypes.POINTER(_VP)] * 3),
        "vibe_mi_domain_set_physics_callback": (_I, [_VP, PHYSICS_FN, _VP, _I]),
        "vibe_mi_domain_energy": (_I, [_VP, _PD]),
        "vibe_mi_domain_cfl": (_I, [_VP, _D, _PD, _PD]),
        "vibe_mi_domain_stability": (_I, [_VP, _D, _PD]),
        "vibe_mi_domain_stream": (_VP, [_VP]),
        "vibe_mi_domain_device_bytes": (_SZ, [_VP]),
        "vibe_mi_batch_create": (_I, [ctypes.POINTER(DomainDesc), _I, ctypes.POINTER(_VP)]),
        "vibe_mi_batch_destroy": (_I, [_VP]),
        "vibe_mi_batch_member": (_I, [_VP, _I, ctypes.POINTER(_VP)]),
        "vibe_mi_batch_step": (_I, [_VP, _PD, _I, _I, _VP]),
        "vibe_mi_batch_step_timed": (_I, [_VP, _PD, _I, _I, _VP, _PD]),
        "vibe_mi_batch_set_graph": (_I, [_VP, _I]),
        "vibe_mi_batch_tuning": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "vibe_mi_batch_kernel_times": (_I, [_VP, _PD, _PD, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "vibe_mi_batch_device_bytes": (_SZ, [_VP]),
        "vibe_mi_davies_profile": (_I, [_I, _D, _PD]),
        "vibe_mi_nest_interpolate": (_I, [_VP, _VP, ctypes.POINTER(NestDesc)]),
        "vibe_mi_nest_relax": (_I, [_VP, _VP, ctypes.POINTER(NestDesc), _D]),
        "vibe_mi_batch_nest_relax": (_I, [_VP, _I, _I, ctypes.POINTER(NestDesc), _D]),
        "vibe_mi_nest_restrict": (_I, [_VP, _VP, ctypes.POINTER(NestDesc)]),
        "vibe_mi_batch_nested_step": (_I, [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(NestDesc),
                                          _I, _D, _I, _I]),
        "vibe_mi_batch_nested_step_timed": (_I, [_VP, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                                ctypes.POINTER(NestDesc), _I, _D, _I, _I, _PD]),
        "vibe_mi_batch_nest_restrict": (_I, [_VP, _I, _I, ctypes.POINTER(NestDesc)]),
        "vibe_mi_helmholtz_create": (_I, [ctypes.POINTER(HelmholtzDesc), ctypes.POINTER(_VP)]),
        "vibe_mi_helmholtz_destroy": (_I, [_VP]),
        "vibe_mi_helmholtz_set_coefficients": (_I, [_VP, _PD, _PD]),
        "vibe_mi_helmholtz_apply": (_I, [_VP, _PD, _PD]),
        "vibe_mi_helmholtz_residual": (_I, [_VP, _PD, _PD, _PD]),
        "vibe_mi_helmholtz_solve": (_I, [_VP, _PD, _PD, _I, _D, _I, ctypes.POINTER(_I), _PD]),
        "vibe_mi_multigrid_restrict_f64": (_I, [_VP, _I, _I, _I, _VP, _I, _I, _I, _VP]),
        "vibe_mi_si_create": (_I, [ctypes.POINTER(SiDesc), ctypes.POINTER(_VP)]),
        "vibe_mi_si_destroy": (_I, [_VP]),
        "vibe_mi_si_set_state": (_I, [_VP, ctypes.POINTER(_PD)]),
        "vibe_mi_si_get_state": (_I, [_VP, ctypes.POINTER(_PD)]),
        "vibe_mi_si_set_physics_tendency": (_I, [_VP, _PD, _PD, _PD, _PD]),
        "vibe_mi_si_step": (_I, [_VP, _D, _I]),
        "vibe_mi_si_set_time_filter": (_I, [_VP, _D]),
        "vibe_mi_si_get_stats": (_I, [_VP, ctypes.POINTER(SiStats)]),
        "vibe_mi_si_reset_stats": (_I, [_VP]),
        "vibe_mi_si_max_stable_dt": (_I, [_VP, _PD]),
        "vibe_mi_si_rhs": (_I, [_VP, _D, _PD, _PD]),
        "vibe_mi_si_pressure": (_I, [_VP, _PD]),
        "vibe_mi_multigrid_prolong_f64": (_I, [_VP, _I, _I, _I, _VP, _I, _I, _I, _VP]),
        "vibe_mi_zeta_uniform": (_I, [_I, _PD]),
        "vibe_mi_zeta_tanh": (_I, [_I, _D, _D, _D, _PD]),
        "vibe_mi_terrain_agnesi_ridge": (_I, [_I, _I, _D, _D, _D, _PD]),
        "vibe_mi_zeta_stretched": (_I, [_I, _D, _D, _PD]),
        "vibe_mi_terrain_agnesi_bell": (_I, [_I, _I, _D, _D, _D, _D, _D, _D, _PD]),
        "vibe_mi_terrain_smooth": (_I, [_PD, _I, _I, _I, _D]),
        "vibe_mi_terrain_gaussian": (_I, [_I, _I, _D, _D, _D, _D, ctypes.c_uint64, _PD]),
        "vibe_mi_reference_state": (_I, [_I, _D, _D, _D, _I, _I, _I, _D, _PD, _PD, _PD, _PD, _PD]),
        "vibe_mi_reference_field_size": (_SZ, [_I, _I, _I, _I, _I]),
        "vibe_mi_state_from_reference": (_I, [_I, _I, _I, _I, ctypes.POINTER(_PD), _PD, _PD, _PD] + [_PD] * 6),
        "vibe_mi_state_to_reference": (_I, [_I, _I, _I, _I] + [_PD] * 9 + [_I, ctypes.POINTER(_PD)]),
        "vibe_mi_literal_create": (_I, [_VP, ctypes.POINTER(_VP)]),
        "vibe_mi_literal_destroy": (_I, [_VP]),
        "vibe_mi_literal_tendencies": (_I, [_VP, ctypes.POINTER(_PD), ctypes.POINTER(_PD)]),
        "vibe_mi_literal_acoustic": (_I, [_VP, ctypes.POINTER(_PD), ctypes.POINTER(_PD)]),
        "vibe_mi_literal_max_sound_speed": (_I, [_VP, ctypes.POINTER(_PD), _PD]),
        "vibe_mi_domain_physics_enable": (_I, [_VP, _I]),
        "vibe_mi_domain_max_stable_dt": (_I, [_VP, _PD]),
        "vibe_mi_reference_state_from_profile": (_I, [_I, _I, _I, _D, _PD, _PD, _I, _PD, _PD, _D, _PD, _PD, _PD]),
        "vibe_mi_sponge": (_I, [_I, _D, _PD, _D, _D, _D, _PD]),
        "vibe_mi_ic_bubble": (_I, [_I, _I, _I, _D, _D, _D, _PD, _PD, _D, _D, _D, _D, _D, _D, _PD]),
        "vibe_mi_ic_noise": (_I, [_I, _I, _I, _D, _PD, _PD, _D, _D, ctypes.c_uint64, _PD]),
        "vibe_mi_ic_wind": (_I, [_I, _I, _I, _D, _PD, _PD, _D, _D, _D, _PD]),
        "vibe_mi_ic_balance_pi": (_I, [_SZ, _PD, _PD, _PD, _PD, _PD]),
        "vibe_mi_vibebin_header": (_I, [ctypes.c_char_p] + [ctypes.POINTER(_I)] * 4 + [_PD] +
                                   [ctypes.POINTER(_I)] * 2),
        "vibe_mi_vibebin_name": (_I, [ctypes.c_char_p, _I, ctypes.c_char_p]),
        "vibe_mi_vibebin_write": (_I, [ctypes.c_char_p, _I, _I, _I, _D, _I,
                                       ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_PD)]),
        "vibe_mi_vibebin_read": (_I, [ctypes.c_char_p, ctypes.c_char_p, _PD, _SZ]),
        "vibe_mi_restart_write": (_I, [ctypes.c_char_p, _I, _I, _I, _D, _I] + [_PD] * 9),
        "vibe_mi_restart_read": (_I, [ctypes.c_char_p, _I, _I, _I, _PD, ctypes.POINTER(_I)] + [_PD] * 9),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _pd(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected contiguous float64"
    return a.ctypes.data_as(_PD)


def _pf(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "expected contiguous float32"
    return a.ctypes.data_as(_PF)


def _check(rc: int, what: str):
    if rc == VIBE_MI_ERR_NONFINITE:
        raise NumericalError(rc, what)
    if rc in (VIBE_MI_ERR_IO, VIBE_MI_ERR_FORMAT, VIBE_MI_ERR_NOT_FOUND):
        raise IoError(rc, what)
    if rc != 0:
        raise VibeError(rc, what)


def _check_last(what: str):
    rc = lib().vibe_mi_last_error()
    if rc != 0:
        raise VibeError(rc, what)


# --------------------------------------------------------------------------
# C11 kernel ABI (c_kernels.h:54-113)
# --------------------------------------------------------------------------

def tridiagonal_solve(a, b, c, d, single=False):
    """x with a[k-1]x[k-1] + b[k]x[k] + c[k]x[k+1] = d[k]; returns (rc, x)."""
    dt = np.float32 if single else np.float64
    b = np.ascontiguousarray(b, dtype=dt)
    d = np.ascontiguousarray(d, dtype=dt)
    n = b.size
    a = None if a is None else np.ascontiguousarray(a, dtype=dt)
    c = None if c is None else np.ascontiguousarray(c, dtype=dt)
    x = np.zeros(max(n, 1), dtype=dt)
    ws = np.zeros(max(n, 1), dtype=dt)
    p = _pf if single else _pd
    fn = lib().vibe_tridiagonal_solve_f32 if single else lib().vibe_tridiagonal_solve_f64
    rc = fn(p(a), p(b), p(c), p(d), n, p(x), p(ws))
    if rc >= 100:
        raise VibeError(rc, "vibe_tridiagonal_solve")
    return rc, x[:n]


def weno5_reconstruct(qm2, qm1, q0, q1, q2, single=False):
    """(q_left, q_right) of c_kernels.h:64-65."""
    if single:
        l, r = ctypes.c_float(), ctypes.c_float()
        lib().vibe_weno5_reconstruct_f32(qm2, qm1, q0, q1, q2, ctypes.byref(l), ctypes.byref(r))
    else:
        l, r = ctypes.c_double(), ctypes.c_double()
        lib().vibe_weno5_reconstruct_f64(qm2, qm1, q0, q1, q2, ctypes.byref(l), ctypes.byref(r))
    _check_last("vibe_weno5_reconstruct")
    return l.value, r.value


def bilinear(f, halo, nsx, nsy, nz, xi, yj, k):
    f = np.ascontiguousarray(f, dtype=np.float64)
    r = lib().vibe_bilinear_f64(_pd(f), halo, nsx, nsy, nz, xi, yj, k)
    _check_last("vibe_bilinear")
    return r


def cfl_scan(u, v, w, dzeta, zs, nx, ny, nz, dx, dy, z_top, dt):
    """(cfl, max_wind, min_dz) of c_kernels.h:86-89."""
    mw, md = ctypes.c_double(), ctypes.c_double()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, w, dzeta)]
    zsa = None if zs is None else np.ascontiguousarray(zs, dtype=np.float64)
    r = lib().vibe_cfl_scan_f64(*[_pd(x) for x in arrs], _pd(zsa), nx, ny, nz, dx, dy, z_top, dt,
                                ctypes.byref(mw), ctypes.byref(md))
    _check_last("vibe_cfl_scan")
    return r, mw.value, md.value


def neumaier_sum(x, single=False):
    if single:
        x = np.ascontiguousarray(x, dtype=np.float32)
        r = lib().vibe_neumaier_sum_f32(_pf(x), x.size)
    else:
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = lib().vibe_neumaier_sum_f64(_pd(x), x.size)
    _check_last("vibe_neumaier_sum")
    return r


def advect_scalar_flux(q, rho, u, v, w, dzeta, zs, nx, ny, nz, dx, dy, z_top):
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (q, rho, u, v, w, dzeta)]
    zsa = None if zs is None else np.ascontiguousarray(zs, dtype=np.float64)
    out = np.zeros(nx * ny * nz)
    lib().vibe_advect_scalar_flux_f64(*[_pd(x) for x in arrs], _pd(zsa), nx, ny, nz, dx, dy, z_top, _pd(out))
    _check_last("vibe_advect_scalar_flux")
    return out


# --------------------------------------------------------------------------
# .vibebin files (include/vibe_mi/vibebin.h; src/io/binary_writer.cpp roles)
# --------------------------------------------------------------------------

def vibebin_shape(name: str, nx: int, ny: int, nz: int):
    """(ni, nj, nk) of a record, from its name (binary_writer.cpp:117-138)."""
    if name == "__restart_info__":
        return (2, 1, 1)
    if name in ("u", "u_facex"):
        return (nx + 1, ny, nz)
    if name in ("v", "v_facey"):
        return (nx, ny + 1, nz)
    if name in ("w", "zeta", "w_facez"):
        return (nx, ny, nz + 1)
    return (nx, ny, nz)


def vibebin_header(path) -> dict:
    """Header of a .vibebin file (read_binary_header role)."""
    v = [_I() for _ in range(4)]
    t = _D()
    nv, nl = _I(), _I()
    _check(lib().vibe_mi_vibebin_header(os.fsencode(path), *[ctypes.byref(x) for x in v], ctypes.byref(t),
                                        ctypes.byref(nv), ctypes.byref(nl)), "vibe_mi_vibebin_header")
    return {"nx": v[0].value, "ny": v[1].value, "nz": v[2].value, "real_kind": v[3].value,
            "time": t.value, "nvars": nv.value, "nlevels": nl.value}


def vibebin_variables(path) -> list:
    """Record names in file order (BinaryReader::variables role)."""
    n = vibebin_header(path)["nvars"]
    out = []
    for i in range(n):
        buf = ctypes.create_string_buffer(32)
        _check(lib().vibe_mi_vibebin_name(os.fsencode(path), i, buf), "vibe_mi_vibebin_name")
        out.append(buf.value.decode("ascii", errors="replace"))
    return out


def vibebin_write(path, nx: int, ny: int, nz: int, records: dict, time: float = 0.0):
    """Writes records {name: array} in insertion order (BinaryWriter role);
    each array is flattened in C order and must hold the name's shape."""
    names, arrs = [], []
    for name, a in records.items():
        shp = vibebin_shape(name[:31], nx, ny, nz)
        arr = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if arr.size != shp[0] * shp[1] * shp[2]:
            raise ValueError(f"record {name!r} holds {arr.size} values, its name implies {shp}")
        names.append(name.encode("ascii"))
        arrs.append(arr)
    cn = (ctypes.c_char_p * max(1, len(names)))(*names)
    cd = (_PD * max(1, len(arrs)))(*[_pd(a) for a in arrs])
    _check(lib().vibe_mi_vibebin_write(os.fsencode(path), nx, ny, nz, float(time), len(names), cn, cd),
           "vibe_mi_vibebin_write")


def vibebin_read(path, name: str) -> np.ndarray:
    """One record as an (nk, nj, ni) array (BinaryReader::read role)."""
    h = vibebin_header(path)
    ni, nj, nk = vibebin_shape(name, h["nx"], h["ny"], h["nz"])
    out = np.empty(ni * nj * nk)
    _check(lib().vibe_mi_vibebin_read(os.fsencode(path), name.encode("ascii"), _pd(out), out.size),
           "vibe_mi_vibebin_read")
    return out.reshape(nk, nj, ni)


def restart_write(path, case: "Case", state: dict, time: float = 0.0, step: int = 0):
    """Restart file of a state in the layout of dycore.h (write_restart role,
    field_io.cpp:207-227): totals of the reference's 12 species plus the exact
    perturbations."""
    st = [np.ascontiguousarray(state[k], dtype=np.float64).reshape(-1) for k in STATE_FIELDS]
    ref = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (case.rho0, case.theta0, case.pi0)]
    _check(lib().vibe_mi_restart_write(os.fsencode(path), case.nx, case.ny, case.nz, float(time), int(step),
                                       *[_pd(a) for a in st + ref]), "vibe_mi_restart_write")


def restart_read(path, case: "Case"):
    """(state, time, step) from a restart file (read_restart role,
    field_io.cpp:229-250); a file without the exact perturbations (one the
    reference wrote) gives totals minus the case's reference state."""
    cell = case.nx * case.ny * case.nz
    st = {k: np.empty(cell) for k in STATE_FIELDS}
    st["w"] = np.empty(case.nx * case.ny * (case.nz + 1))
    ref = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (case.rho0, case.theta0, case.pi0)]
    t, s = _D(), _I()
    _check(lib().vibe_mi_restart_read(os.fsencode(path), case.nx, case.ny, case.nz, ctypes.byref(t),
                                      ctypes.byref(s), *[_pd(st[k]) for k in STATE_FIELDS],
                                      *[_pd(a) for a in ref]), "vibe_mi_restart_read")
    return st, t.value, s.value


# --------------------------------------------------------------------------
# dynamical core
# --------------------------------------------------------------------------

STATE_FIELDS = ("u", "v", "w", "rho", "theta", "pi")


@dataclass
class Case:
    """Geometry, reference state and initial perturbation state of one domain
    (build with vibe_mi.configs)."""

    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    ztop: float
    zeta: np.ndarray
    zs: np.ndarray
    rho0: np.ndarray
    theta0: np.ndarray
    pi0: np.ndarray
    sponge: np.ndarray
    state: dict
    dt: float
    nsub: int
    steps: int
    eps_s: float = 0.1
    meta: dict = field(default_factory=dict)

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz

    def copy_state(self):
        return {k: v.copy() for k, v in self.state.items()}


def _desc_of(case: Case, precision: int):
    """DomainDesc of a case plus the arrays it points to (keep them alive)."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (case.zeta, case.zs, case.rho0, case.theta0, case.pi0, case.sponge)]
    z, zs, r0, t0, p0, sp = keep
    desc = DomainDesc(case.nx, case.ny, case.nz, case.dx, case.dy, case.ztop,
                      _pd(z), _pd(zs), _pd(r0), _pd(t0), _pd(p0), _pd(sp), case.eps_s, precision)
    return desc, keep


class Domain:
    """One domain resident on the GPU (RungeKutta3Integrator role)."""

    def __init__(self, case: Case, precision: int = FP64):
        L = lib()
        self.case = case
        self.precision = precision
        self.owner = None
        self.desc, self._keep = _desc_of(case, precision)
        h = _VP()
        _check(L.vibe_mi_domain_create(ctypes.byref(self.desc), ctypes.byref(h)), "vibe_mi_domain_create")
        self.h = h
        self.set_state(case.state)

    @classmethod
    def _member(cls, case: Case, precision: int, handle, owner):
        """Non-owning view of a batch member."""
        self = cls.__new__(cls)
        self.case = case
        self.precision = precision
        self.owner = owner
        self.h = handle
        return self

    def close(self):
        if getattr(self, "h", None):
            if self.owner is None:
                lib().vibe_mi_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shapes(self):
        c = self.case
        cell = c.nx * c.ny * c.nz
        return {"u": cell, "v": cell, "w": c.nx * c.ny * (c.nz + 1), "rho": cell, "theta": cell, "pi": cell}

    def set_state(self, st: dict):
        arrs = [np.ascontiguousarray(st[k], dtype=np.float64).reshape(-1) for k in STATE_FIELDS]
        for k, a in zip(STATE_FIELDS, arrs):
            assert a.size == self._shapes()[k], f"field {k} has {a.size} values"
        _check(lib().vibe_mi_domain_set_state(self.h, *[_pd(a) for a in arrs]), "vibe_mi_domain_set_state")

    def get_state(self) -> dict:
        out = {k: np.empty(n) for k, n in self._shapes().items()}
        _check(lib().vibe_mi_domain_get_state(self.h, *[_pd(out[k]) for k in STATE_FIELDS]),
               "vibe_mi_domain_get_state")
        return out

    def step(self, dt: float, substeps: int, nsteps: int = 1):
        _check(lib().vibe_mi_domain_step(self.h, dt, substeps, nsteps), "vibe_mi_domain_step")

    def step_timed(self, dt: float, substeps: int, nsteps: int = 1) -> float:
        ms = ctypes.c_double()
        _check(lib().vibe_mi_domain_step_timed(self.h, dt, substeps, nsteps, ctypes.byref(ms)),
               "vibe_mi_domain_step_timed")
        return ms.value

    def kernel_times(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        na, nb = _I(), _I()
        _check(lib().vibe_mi_domain_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(na),
                                                 ctypes.byref(nb)), "vibe_mi_domain_kernel_times")
        return {"ms_stage": a.value, "ms_substep": b.value, "n_stage": na.value, "n_substep": nb.value}

    def slow_tendency(self) -> dict:
        out = {k: np.empty(n) for k, n in self._shapes().items()}
        _check(lib().vibe_mi_domain_slow_tendency(self.h, *[_pd(out[k]) for k in STATE_FIELDS]),
               "vibe_mi_domain_slow_tendency")
        return out

    def acoustic_substep(self, dtau: float):
        _check(lib().vibe_mi_domain_substep(self.h, dtau), "vibe_mi_domain_substep")

    # ---- physics coupling (runge_kutta.cpp:46-63; include/vibe_mi/dycore.h)

    def set_physics_tendency(self, theta=None, u=None, v=None):
        """Holds theta' (K/s), u and v (m/s^2) rates, added to every stage's
        slow tendency; all None switches the coupling off."""
        c = self.case
        cell = c.nx * c.ny * c.nz
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (theta, u, v)]
        for a in arrs:
            if a is not None and a.size != cell:
                raise ValueError(f"physics rates hold {a.size} values, the domain has {cell} cells")
        self._phys_keep = arrs
        _check(lib().vibe_mi_domain_set_physics_tendency(self.h, *[_pd(a) for a in arrs]),
               "vibe_mi_domain_set_physics_tendency")

    def physics_buffers(self):
        """Device addresses (ints) of the theta', u, v rate buffers."""
        p = [_VP() for _ in range(3)]
        _check(lib().vibe_mi_domain_physics_buffers(self.h, *[ctypes.byref(x) for x in p]),
               "vibe_mi_domain_physics_buffers")
        return tuple(x.value for x in p)

    def set_physics_callback(self, fn, host: bool = True):
        """Per-stage physics: before each stage, fn(stage, dt, state, rates).
        host=True: state is a dict of the stage's FP64 host arrays (u, v, w,
        rho, theta, pi) and rates a dict of zeroed host arrays (theta, u, v)
        that fn fills; host=False: both hold device addresses (ints) and fn
        enqueues its work on rates["stream"].  fn returns None or 0 on
        success; fn=None removes the callback."""
        if fn is None:
            self._phys_cb = None
            _check(lib().vibe_mi_domain_set_physics_callback(self.h, PHYSICS_FN(), None, 0),
                   "vibe_mi_domain_set_physics_callback")
            return
        c = self.case
        cell = c.nx * c.ny * c.nz
        nwf = c.nx * c.ny * (c.nz + 1)

        def arr(p, n):
            return np.ctypeslib.as_array(ctypes.cast(p, _PD), shape=(n,))

        def tramp(_user, stage, dt, stream, u, v, w, rho, th, pi, rt, ru, rv):
            try:
                if host:
                    st = {"u": arr(u, cell), "v": arr(v, cell), "w": arr(w, nwf), "rho": arr(rho, cell),
                          "theta": arr(th, cell), "pi": arr(pi, cell)}
                    rates = {"theta": arr(rt, cell), "u": arr(ru, cell), "v": arr(rv, cell)}
                else:
                    st = {"u": u, "v": v, "w": w, "rho": rho, "theta": th, "pi": pi}
                    rates = {"theta": rt, "u": ru, "v": rv, "stream": stream}
                r = fn(stage, dt, st, rates)
                return int(r or 0)
            except Exception:  # an exception must not cross the C frame
                return 1

        self._phys_cb = PHYSICS_FN(tramp)
        _check(lib().vibe_mi_domain_set_physics_callback(self.h, self._phys_cb, None, 1 if host else 0),
               "vibe_mi_domain_set_physics_callback")

    def write_restart(self, path, time: float = 0.0, step: int = 0):
        """Downloads the state and writes it as a restart file
        (write_restart role, src/io/field_io.cpp:207-227)."""
        restart_write(path, self.case, self.get_state(), time, step)

    def read_restart(self, path):
        """Loads a restart file into the device state; returns (time, step)
        (read_restart role, src/io/field_io.cpp:229-250).  FP64 domains
        continue bit for bit; FP32-mixed domains recompute their background
        wind from the loaded state, as every set_state does."""
        st, t, s = restart_read(path, self.case)
        self.set_state(st)
        return t, s

    def energy_budget(self) -> np.ndarray:
        out = np.zeros(6)
        _check(lib().vibe_mi_domain_energy(self.h, _pd(out)), "vibe_mi_domain_energy")
        return out

    def stability(self, dt: float) -> dict:
        """TimeStepController::inspect role (timestep_control.cpp:23-56)."""
        out = np.zeros(6)
        _check(lib().vibe_mi_domain_stability(self.h, dt, _pd(out)), "vibe_mi_domain_stability")
        return {"cfl_advection": out[0], "cfl_acoustic": out[1], "cfl_vertical": out[2], "max_wind": out[3],
                "min_dz": out[4], "stable": bool(out[5])}

    def max_stable_dt(self) -> float:
        """RungeKutta3Integrator::max_stable_dt (runge_kutta.cpp:121-141) of
        the device state: min(0.8 / CFL(dt = 1), 2 / (cs_max max(1/dx + 1/dy
        + 1/dz)))."""
        out = ctypes.c_double()
        _check(lib().vibe_mi_domain_max_stable_dt(self.h, ctypes.byref(out)), "vibe_mi_domain_max_stable_dt")
        return out.value

    def physics_enable(self, on: bool) -> None:
        """StepContext::physics_on: with False the registered physics (rates
        or callback) is skipped until enabled again."""
        _check(lib().vibe_mi_domain_physics_enable(self.h, 1 if on else 0), "vibe_mi_domain_physics_enable")

    def cfl(self, dt: float):
        c, w = ctypes.c_double(), ctypes.c_double()
        _check(lib().vibe_mi_domain_cfl(self.h, dt, ctypes.byref(c), ctypes.byref(w)), "vibe_mi_domain_cfl")
        return c.value, w.value

    def device_bytes(self) -> int:
        return lib().vibe_mi_domain_device_bytes(self.h)


class Batch:
    """Many domains of uneven size advanced together: one persistent launch
    per kernel phase over a device tile work list, the step's launch
    sequence replayed from a hipGraph (vibe_mi/batch.h)."""

    def __init__(self, cases, precision: int = FP64, use_graph: bool = True):
        L = lib()
        self.cases = list(cases)
        self.precision = precision
        n = len(self.cases)
        descs = (DomainDesc * n)()
        self._keep = []
        for m, c in enumerate(self.cases):
            d, keep = _desc_of(c, precision)
            descs[m] = d
            self._keep.append(keep)
        h = _VP()
        _check(L.vibe_mi_batch_create(descs, n, ctypes.byref(h)), "vibe_mi_batch_create")
        self.h = h
        self._keep = None  # the members hold their own copies of the inputs
        self.members = []
        for m, c in enumerate(self.cases):
            mh = _VP()
            _check(L.vibe_mi_batch_member(h, m, ctypes.byref(mh)), "vibe_mi_batch_member")
            dom = Domain._member(c, precision, mh, self)
            dom.set_state(c.state)
            self.members.append(dom)
        self.set_graph(use_graph)

    def __len__(self):
        return len(self.members)

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            lib().vibe_mi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _args(self, dt, active):
        n = len(self.members)
        if dt is None:
            dt = [c.dt for c in self.cases]
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)))
        act = None
        if active is not None:
            act = np.ascontiguousarray(np.asarray(active, dtype=np.uint8).reshape(n))
        return dt, act

    def step(self, dt=None, nsub: int = 6, nsteps: int = 1, active=None):
        dt, act = self._args(dt, active)
        _check(lib().vibe_mi_batch_step(self.h, _pd(dt), nsub, nsteps,
                                        None if act is None else act.ctypes.data_as(_VP)), "vibe_mi_batch_step")

    def step_timed(self, dt=None, nsub: int = 6, nsteps: int = 1, active=None) -> float:
        dt, act = self._args(dt, active)
        ms = ctypes.c_double()
        _check(lib().vibe_mi_batch_step_timed(self.h, _pd(dt), nsub, nsteps,
                                              None if act is None else act.ctypes.data_as(_VP), ctypes.byref(ms)),
               "vibe_mi_batch_step_timed")
        return ms.value

    def set_graph(self, use_graph: bool):
        _check(lib().vibe_mi_batch_set_graph(self.h, 1 if use_graph else 0), "vibe_mi_batch_set_graph")

    def kernel_times(self):
        """Per-group times of the last timed call made without a graph."""
        a, b = ctypes.c_double(), ctypes.c_double()
        na, nb = _I(), _I()
        _check(lib().vibe_mi_batch_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(na),
                                                ctypes.byref(nb)), "vibe_mi_batch_kernel_times")
        return {"ms_stage": a.value, "ms_substep": b.value, "n_stage": na.value, "n_substep": nb.value}

    def tuning(self):
        a, b = _I(), _I()
        _check(lib().vibe_mi_batch_tuning(self.h, ctypes.byref(a), ctypes.byref(b)), "vibe_mi_batch_tuning")
        return {"substep_waves": a.value, "stage_rows": b.value}

    def device_bytes(self) -> int:
        return lib().vibe_mi_batch_device_bytes(self.h)

    @property
    def ncell(self):
        return sum(c.ncell for c in self.cases)


# --------------------------------------------------------------------------
# nest lateral boundary (vibe_mi/nest.h)
# --------------------------------------------------------------------------

def davies_profile(zone: int, alpha_max: float = 1.0) -> np.ndarray:
    out = np.zeros(zone)
    _check(lib().vibe_mi_davies_profile(zone, alpha_max, _pd(out)), "vibe_mi_davies_profile")
    return out


def nest_interpolate(child: "Domain", parent: "Domain", desc: NestDesc):
    """Child state <- bilinear interpolation of the parent's current state."""
    _check(lib().vibe_mi_nest_interpolate(child.h, parent.h, ctypes.byref(desc)), "vibe_mi_nest_interpolate")


def nest_relax(child: "Domain", parent: "Domain", desc: NestDesc, frac: float = 1.0):
    """Davies relaxation of the child's lateral zone toward the parent state
    interpolated in space and (with frac) in time between the parent's last two
    states."""
    _check(lib().vibe_mi_nest_relax(child.h, parent.h, ctypes.byref(desc), frac), "vibe_mi_nest_relax")


def batch_nest_relax(batch: "Batch", child: int, parent: int, desc: NestDesc, frac: float = 1.0):
    """nest_relax between members of one batch, enqueued on the batch stream
    behind its last step without a host synchronisation."""
    _check(lib().vibe_mi_batch_nest_relax(batch.h, child, parent, ctypes.byref(desc), frac),
           "vibe_mi_batch_nest_relax")


def nest_restrict(child: "Domain", parent: "Domain", desc: NestDesc):
    """Two-way feedback: the parent cells covered by the child take the mean
    of the child state over them (vibe_mi/nest.h)."""
    _check(lib().vibe_mi_nest_restrict(child.h, parent.h, ctypes.byref(desc)), "vibe_mi_nest_restrict")


def batch_nest_restrict(batch: "Batch", child: int, parent: int, desc: NestDesc):
    """nest_restrict between members of one batch, on the batch stream."""
    _check(lib().vibe_mi_batch_nest_restrict(batch.h, child, parent, ctypes.byref(desc)),
           "vibe_mi_batch_nest_restrict")


def batch_nested_step(batch: "Batch", links, two_way: bool, dt0: float, nsub: int, nsteps: int = 1,
                      timed: bool = False) -> float:
    """nsteps parent steps of a nest hierarchy in one call (vibe_mi/nest.h:
    vibe_mi_batch_nested_step); links = [(child, parent, NestDesc)].  Returns
    the elapsed ms on the batch stream when timed."""
    n = len(links)
    ch = (_I * n)(*[c for c, _, _ in links])
    pa = (_I * n)(*[p for _, p, _ in links])
    ds = (NestDesc * n)(*[d for _, _, d in links])
    if timed:
        ms = _D(0.0)
        _check(lib().vibe_mi_batch_nested_step_timed(batch.h, n, ch, pa, ds, int(two_way), dt0, nsub, nsteps,
                                                     ctypes.byref(ms)), "vibe_mi_batch_nested_step_timed")
        return ms.value
    _check(lib().vibe_mi_batch_nested_step(batch.h, n, ch, pa, ds, int(two_way), dt0, nsub, nsteps),
           "vibe_mi_batch_nested_step")
    return 0.0


HELM_BICGSTAB = 0
HELM_REFINE = 1
HELM_MULTIGRID = 2


def multigrid_restrict(fine, coarse):
    """multigrid_restrict (helmholtz.cpp:131-145) of the device tensor fine
    [nz, ny, nx] into the device tensor coarse (float64, contiguous, on the
    current device; enqueued on the null stream)."""
    _check_dev3(fine, coarse)
    fz, fy, fx = fine.shape
    cz, cy, cx = coarse.shape
    _check(lib().vibe_mi_multigrid_restrict_f64(fine.data_ptr(), fx, fy, fz, coarse.data_ptr(), cx, cy, cz, None),
           "vibe_mi_multigrid_restrict_f64")


def multigrid_prolong(coarse, fine):
    """multigrid_prolong (helmholtz.cpp:147-164) of the device tensor coarse
    into the device tensor fine (as multigrid_restrict)."""
    _check_dev3(fine, coarse)
    fz, fy, fx = fine.shape
    cz, cy, cx = coarse.shape
    _check(lib().vibe_mi_multigrid_prolong_f64(coarse.data_ptr(), cx, cy, cz, fine.data_ptr(), fx, fy, fz, None),
           "vibe_mi_multigrid_prolong_f64")


def _check_dev3(*ts):
    import torch
    for t in ts:
        if t.dtype != torch.float64 or t.dim() != 3 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("expected contiguous float64 device tensors of shape [nz, ny, nx]")


class Helmholtz:
    """3-D Helmholtz operator -div(a grad x) + b x and its solvers on the GPU
    (HelmholtzSolver role, include/vibe/time/helmholtz.hpp:63-86)."""

    def __init__(self, nx, ny, nz, dx, dy, ztop, zeta, zs=None):
        self.shape = (nz, ny, nx)
        self._zeta = np.ascontiguousarray(zeta, dtype=np.float64)
        self._zs = np.ascontiguousarray(np.zeros(nx * ny) if zs is None else zs, dtype=np.float64).reshape(-1)
        if self._zeta.size != nz + 1 or self._zs.size != nx * ny:
            raise ValueError("zeta needs nz+1 values and zs nx*ny")
        self.desc = HelmholtzDesc(nx, ny, nz, dx, dy, ztop, _pd(self._zeta), _pd(self._zs))
        h = _VP()
        _check(lib().vibe_mi_helmholtz_create(ctypes.byref(self.desc), ctypes.byref(h)), "vibe_mi_helmholtz_create")
        self.h = h
        self.iterations = 0
        self.residual_rel = float("nan")

    def _arr(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if a.size != self.shape[0] * self.shape[1] * self.shape[2]:
            raise ValueError("expected an nz*ny*nx array")
        return a

    def set_coefficients(self, a, b):
        _check(lib().vibe_mi_helmholtz_set_coefficients(self.h, _pd(self._arr(a)), _pd(self._arr(b))),
               "vibe_mi_helmholtz_set_coefficients")

    def apply(self, x):
        y = np.empty(self.shape)
        _check(lib().vibe_mi_helmholtz_apply(self.h, _pd(self._arr(x)), _pd(y)), "vibe_mi_helmholtz_apply")
        return y

    def residual(self, f, x) -> float:
        r = _D()
        _check(lib().vibe_mi_helmholtz_residual(self.h, _pd(self._arr(f)), _pd(self._arr(x)), ctypes.byref(r)),
               "vibe_mi_helmholtz_residual")
        return r.value

    def solve(self, f, kind: int = HELM_BICGSTAB, tol: float = 1e-10, max_iter: int = 300):
        """x with A x = f from x = 0; sets .iterations and .residual_rel."""
        x = np.empty(self.shape)
        its, rel = _I(), _D()
        _check(lib().vibe_mi_helmholtz_solve(self.h, _pd(self._arr(f)), _pd(x), kind, tol, max_iter,
                                             ctypes.byref(its), ctypes.byref(rel)), "vibe_mi_helmholtz_solve")
        self.iterations, self.residual_rel = its.value, rel.value
        return x

    def close(self):
        if getattr(self, "h", None):
            lib().vibe_mi_helmholtz_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SPECIES = ("u", "v", "w", "rho", "theta", "pi", "qv", "qc", "qr", "qi", "qs", "qg")


class SemiImplicit:
    """The reference's semi-implicit integrator on the GPU
    (SemiImplicitIntegrator, src/time/semi_implicit.cpp:42-233; C ABI
    vibe_mi/semi_implicit.h) over its literal equations: FULL prognostic
    values of the 12 species in the reference's State order, bounded x and
    y.  The device copy of the state is authoritative between calls.

    grid: nx, ny, nz, dx, dy, z_top, zeta (nz+1), zs (ny*nx or None);
    reference state rho0, theta0, pi0 (cells); numerics: advection order
    (2/4/6), coriolis (0/1/2), latitude, div_coeff, div_order, moist;
    sponge (nz rates or None); solver (HELM_*), tol, max_iter;
    time_filter (set_time_filter alpha)."""

    def __init__(self, nx, ny, nz, dx, dy, z_top, zeta, rho0, theta0, pi0, zs=None, order=2, coriolis=0,
                 latitude=45.0, div_coeff=0.0, div_order=2, moist=0, sponge=None, solver=HELM_BICGSTAB,
                 tol=1e-8, max_iter=200, time_filter=0.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        nc = nx * ny * nz
        self._keep = [np.ascontiguousarray(zeta, dtype=np.float64)]
        if self._keep[0].size != nz + 1:
            raise ValueError("zeta needs nz+1 values")
        ref = []
        for a in (rho0, theta0, pi0):
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.size != nc:
                raise ValueError("reference state arrays need nx*ny*nz values")
            ref.append(a)
        self._keep += ref
        zsa = None
        if zs is not None:
            zsa = np.ascontiguousarray(zs, dtype=np.float64).reshape(-1)
            if zsa.size != nx * ny:
                raise ValueError("zs needs nx*ny values")
            self._keep.append(zsa)
        sp = None
        if sponge is not None:
            sp = np.ascontiguousarray(sponge, dtype=np.float64).reshape(-1)
            if sp.size != nz:
                raise ValueError("sponge needs nz rates")
            self._keep.append(sp)
        eq = LiteralDesc(nx, ny, nz, dx, dy, z_top, _pd(self._keep[0]), _pd(zsa), _pd(ref[0]), _pd(ref[1]),
                         _pd(ref[2]), order, coriolis, latitude, div_coeff, div_order, moist)
        self.desc = SiDesc(eq, _pd(sp), solver, tol, max_iter, time_filter)
        h = _VP()
        _check(lib().vibe_mi_si_create(ctypes.byref(self.desc), ctypes.byref(h)), "vibe_mi_si_create")
        self.h = h

    def sizes(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        c = nx * ny * nz
        return [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)] + [c] * 9

    def _ptrs(self, arrs):
        return (_PD * 12)(*[_pd(a) for a in arrs])

    def set_state(self, s):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in s]
        if [a.size for a in arrs] != self.sizes():
            raise ValueError("state sizes do not match the grid")
        _check(lib().vibe_mi_si_set_state(self.h, self._ptrs(arrs)), "vibe_mi_si_set_state")

    def get_state(self):
        out = [np.empty(n) for n in self.sizes()]
        _check(lib().vibe_mi_si_get_state(self.h, self._ptrs(out)), "vibe_mi_si_get_state")
        return out

    def set_physics_tendency(self, theta=None, qv=None, u=None, v=None):
        self._phys = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                      for a in (theta, qv, u, v)]
        _check(lib().vibe_mi_si_set_physics_tendency(self.h, *[_pd(a) for a in self._phys]),
               "vibe_mi_si_set_physics_tendency")

    def step(self, dt, nsteps=1):
        _check(lib().vibe_mi_si_step(self.h, dt, nsteps), "vibe_mi_si_step")

    def set_time_filter(self, alpha):
        _check(lib().vibe_mi_si_set_time_filter(self.h, alpha), "vibe_mi_si_set_time_filter")

    def stats(self) -> dict:
        st = SiStats()
        _check(lib().vibe_mi_si_get_stats(self.h, ctypes.byref(st)), "vibe_mi_si_get_stats")
        return {f: getattr(st, f) for f, _ in SiStats._fields_}

    def reset_stats(self):
        _check(lib().vibe_mi_si_reset_stats(self.h), "vibe_mi_si_reset_stats")

    def max_stable_dt(self) -> float:
        r = _D()
        _check(lib().vibe_mi_si_max_stable_dt(self.h, ctypes.byref(r)), "vibe_mi_si_max_stable_dt")
        return r.value

    def rhs(self, dt):
        """(a, rhs) of assemble_rhs for the device state."""
        c = self.nx * self.ny * self.nz
        a, r = np.empty(c), np.empty(c)
        _check(lib().vibe_mi_si_rhs(self.h, dt, _pd(a), _pd(r)), "vibe_mi_si_rhs")
        return a, r

    def pressure(self):
        """pi_new of the last step's Helmholtz solve."""
        x = np.empty(self.nx * self.ny * self.nz)
        _check(lib().vibe_mi_si_pressure(self.h, _pd(x)), "vibe_mi_si_pressure")
        return x

    def close(self):
        if getattr(self, "h", None):
            lib().vibe_mi_si_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
