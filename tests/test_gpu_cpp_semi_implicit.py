"""The C++ caller of the semi-implicit step
(vibe-model_amd/examples/run_semi_implicit.cpp, built by the library's
Makefile against include/vibe_mi/semi_implicit.hpp) on the GPU, against the
oracle's restatement of the reference's step (orc_si_step, oracle/literal.c;
SemiImplicitIntegrator, src/time/semi_implicit.cpp:167-233).

The caller builds its grid with the library's set-up helpers (stretched
levels, Agnesi ridge, the reference's isothermal state, the sponge profile:
all pinned bit for bit against the reference's own code in
test_setup_golden.py) and a hashed initial state; this test rebuilds the same
inputs, runs the oracle for the same 4 steps (moisture, fourth-order
advection, Coriolis, divergence damping, sponge, time filter, a theta
physics rate) and requires the dumped 12 species to equal it bit for bit.
The statistics (4 steps, Helmholtz iterations > 0, residual below the
solver tolerance), the boundary callback (once per step) and max_stable_dt
(0.6 / advective CFL of the final state, computed here in numpy) are checked
from the caller's output.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from literal_cases import LiteralDesc, hash_r, oracle_si_step
from oracle_bind import _PD

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "vibe-model_amd", "bin", "run_semi_implicit")
NX, NY, NZ, STEPS, DT = 40, 24, 16, 4, 3.0
DX, DY, ZTOP = 1000.0, 1200.0, 15000.0


def _shape(q):
    return (NZ + (q == 2), NY + (q == 1), NX + (q == 0))


def test_cpp_semi_implicit_caller_matches_oracle(vmi, tmp_path):
    assert os.path.exists(EXE), "build the library first (__graft_entry__.build())"
    dump = tmp_path / "si_state.bin"
    print("[cpp-si] running", EXE, flush=True)
    out = subprocess.run([EXE, "--dump", str(dump)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    got = np.fromfile(dump, dtype=np.float64)

    L, pd = vmi.lib(), vmi._pd
    nc = NX * NY * NZ
    zeta, zs = np.empty(NZ + 1), np.empty(NX * NY)
    r0, t0, p0, sponge = np.empty(nc), np.empty(nc), np.empty(nc), np.empty(NZ)
    assert L.vibe_mi_zeta_stretched(NZ, 1.0, 1.05, pd(zeta)) == 0
    assert L.vibe_mi_terrain_agnesi_bell(NX, NY, DX, DY, 600.0, 8000.0, 20000.0, 12000.0, pd(zs)) == 0
    assert L.vibe_mi_reference_state(0, 300.0, 0.0, 100000.0, NX, NY, NZ, ZTOP, pd(zeta), pd(zs), pd(r0), pd(t0),
                                     pd(p0)) == 0
    assert L.vibe_mi_sponge(NZ, ZTOP, pd(zeta), 0.0, 0.7, 0.02, pd(sponge)) == 0
    assert np.any(sponge > 0)
    s = []
    for q in range(12):
        r = hash_r(q, _shape(q)).ravel()
        if q == 0:
            a = 8.0 + 4.0 * r
        elif q == 1:
            a = -2.0 + 3.0 * r
        elif q == 2:
            a = 0.2 * r
        elif q == 3:
            a = r0 * (1.0 + 1e-3 * r)
        elif q == 4:
            a = t0 + r
        elif q == 5:
            a = p0 * (1.0 + 1e-4 * r)
        elif q == 6:
            a = 5e-3 * (r + 0.5)
        else:
            a = 1e-4 * (r + 0.5)
        s.append(np.ascontiguousarray(a))
    rate = np.ascontiguousarray(1e-4 * hash_r(12, (NZ, NY, NX)).ravel())
    d = LiteralDesc(NX, NY, NZ, DX, DY, ZTOP, zeta.ctypes.data_as(_PD), zs.ctypes.data_as(_PD),
                    r0.ctypes.data_as(_PD), t0.ctypes.data_as(_PD), p0.ctypes.data_as(_PD), 4, 1, 40.0, 1e5, 2, 1)
    assert oracle_si_step(d, s, DT, sponge=sponge, time_filter=0.05, first=True, phys=[rate, None, None, None],
                          nsteps=STEPS) == 0
    want = np.concatenate(s)
    assert got.size == want.size
    off = 0
    for q in range(12):
        n = s[q].size
        assert np.array_equal(got[off:off + n], s[q]), q
        off += n

    m = re.search(r"(\d+) steps, (\d+) Helmholtz iterations, last residual ([0-9.e+-]+), boundary calls (\d+), "
                  r"time ([0-9.]+) s", out.stdout)
    assert m, out.stdout
    assert int(m.group(1)) == STEPS and int(m.group(4)) == STEPS
    assert int(m.group(2)) >= STEPS and float(m.group(3)) < 1e-9
    assert float(m.group(5)) == STEPS * DT
    msd = float(re.search(r"max_stable_dt ([0-9.e+-]+)", out.stdout).group(1))
    u, v, w = (s[q].reshape(_shape(q)) for q in range(3))
    dzeta = np.diff(zeta)
    dz = np.maximum(dzeta[:, None, None] * (ZTOP - zs.reshape(NY, NX))[None], 1.0)
    cfl = np.max(1.0 * (np.abs(0.5 * (u[:, :, :-1] + u[:, :, 1:])) / DX + np.abs(0.5 * (v[:, :-1] + v[:, 1:])) / DY +
                        np.abs(0.5 * (w[:-1] + w[1:])) / dz))
    assert msd == 0.6 / cfl
