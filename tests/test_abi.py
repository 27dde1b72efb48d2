"""The C-ABI library loads on the CPU host and exports every entry point the
public headers declare (no compute calls: no GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDRS = sorted(os.path.join(ROOT, "include", "vibe_mi", h) for h in os.listdir(os.path.join(ROOT, "include", "vibe_mi"))
              if h.endswith(".h"))


def declared_functions():
    names = []
    for h in HDRS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"\b(vibe_\w+)\s*\(", src)
    return sorted(set(names))


def test_headers_declare_reference_entry_points():
    names = declared_functions()
    # the reference's C ABI (include/vibe/common/c_kernels.h:54-113)
    for ref in ("vibe_tridiagonal_solve_f64", "vibe_weno5_reconstruct_f64", "vibe_bilinear_f64",
                "vibe_cfl_scan_f64", "vibe_neumaier_sum_f64", "vibe_advect_scalar_flux_f64"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    import vibe_mi
    L = vibe_mi.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", vibe_mi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    assert set(declared_functions()) <= exported


def test_library_is_gfx950_code_object():
    import vibe_mi
    blob = open(vibe_mi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_domain_create_rejects_fields_of_4_gib():
    """Kernels address a field with 32-bit byte offsets: a grid whose FP64 w
    array (nx*ny*(nz+1)*8 B) reaches 4 GiB is refused with VIBE_C_ERR_BAD_ARG
    before anything is allocated (no GPU is touched)."""
    import ctypes
    import numpy as np
    import vibe_mi
    L = vibe_mi.lib()
    nz = 96
    one = np.zeros(nz + 1)
    nx = ny = 2400  # 2400*2400*97*8 B = 4.47e9 B
    desc = vibe_mi.DomainDesc(nx, ny, nz, 1000.0, 1000.0, 20000.0, vibe_mi._pd(one), vibe_mi._pd(one),
                              vibe_mi._pd(one), vibe_mi._pd(one), vibe_mi._pd(one), vibe_mi._pd(one), 0.1,
                              vibe_mi.FP64)
    h = ctypes.c_void_p()
    assert L.vibe_mi_domain_create(ctypes.byref(desc), ctypes.byref(h)) == 2
    assert not h.value


@pytest.mark.parametrize("nx,ny,nz,prec", [(3, 8, 10, 0), (8, 3, 10, 0), (8, 8, 2, 0), (8, 8, 104, 0),
                                           (8, 8, 129, 1), (8, 8, 129, 2), (8, 8, 10, 3), (8, 8, 10, -1)])
def test_domain_create_rejects_sizes_outside_the_kernels_range(nx, ny, nz, prec):
    """nx, ny >= 4 (3-cell periodic stencils) and 3 <= nz <= 103 (FP64) or
    128 (FP32-mixed and FP32-compensated), the tallest column one substep
    workgroup takes; precision one of VIBE_MI_FP64 (0), VIBE_MI_MIXED (1),
    VIBE_MI_MIXED_COMP (2); refused before any allocation."""
    import ctypes
    import numpy as np
    import vibe_mi
    L = vibe_mi.lib()
    big = np.zeros(max(nx * ny * (nz + 1), nz + 1))
    desc = vibe_mi.DomainDesc(nx, ny, nz, 1000.0, 1000.0, 20000.0, *[vibe_mi._pd(big)] * 6, 0.1, prec)
    h = ctypes.c_void_p()
    assert L.vibe_mi_domain_create(ctypes.byref(desc), ctypes.byref(h)) == 2
    assert not h.value
