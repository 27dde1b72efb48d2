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
