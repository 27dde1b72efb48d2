"""3-D Helmholtz solver on the GPU (include/vibe_mi/helmholtz.h) against the
numpy restatement (oracle/helmholtz.py) and its direct solution.

Bars: the FP64 operator and the multigrid transfers equal the oracle's bit
for bit; residual norms agree to 1e-13 relative; the three solvers (FP64
BiCGSTAB, FP32 + FP64 iterative refinement, the two-grid start) reach the
FP64 relative residual 1e-10 and the direct solution to 1e-8 relative; the
reference's own test cases hold (tests/unit/test_time.cpp:101-171)."""
import os
import sys
import time

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import helmholtz as orc  # noqa: E402

pytestmark = pytest.mark.gpu


def _stretched(nz, ratio=1.08):
    d = ratio ** np.arange(nz)
    return np.r_[0.0, np.cumsum(d)] / np.sum(d)


def _problem(nx, ny, nz, seed, terrain=True):
    rng = np.random.default_rng(seed)
    zeta = _stretched(nz)
    zs = rng.uniform(0, 800, (ny, nx)) if terrain else np.zeros((ny, nx))
    sh = (nz, ny, nx)
    # a ~ (dt/2)^2 (Rd/cv) pi0 cp theta0 for dt = 6 s, b = 1 (semi_implicit.cpp:66-72)
    a = 9.0 * 0.4 * 0.9 * 1004.64 * rng.uniform(290, 330, sh)
    b = np.ones(sh)
    return zeta, zs, a, b, rng


def test_apply_and_residual_match_oracle(vmi):
    nx, ny, nz = 37, 23, 19
    zeta, zs, a, b, rng = _problem(nx, ny, nz, 1)
    x = rng.normal(size=(nz, ny, nx))
    f = rng.normal(size=(nz, ny, nx))
    h = vmi.Helmholtz(nx, ny, nz, 1500.0, 1200.0, 18000.0, zeta, zs)
    h.set_coefficients(a, b)
    dz = orc.dz_of(zeta, zs, 18000.0)
    y = h.apply(x)
    assert np.array_equal(y, orc.apply(a, b, x, 1500.0, 1200.0, dz))
    r = h.residual(f, x)
    ro = orc.residual_l2(a, b, f, x, 1500.0, 1200.0, dz)
    assert abs(r - ro) <= 1e-13 * ro
    h.close()


@pytest.mark.parametrize("kind", ["bicgstab", "refine", "multigrid"])
@pytest.mark.parametrize("terrain", [False, True])
def test_solvers_reach_the_direct_solution(vmi, kind, terrain):
    nx, ny, nz = 20, 16, 24
    zeta, zs, a, b, rng = _problem(nx, ny, nz, 7, terrain)
    f = rng.normal(size=(nz, ny, nx))
    dx = dy = 1000.0
    h = vmi.Helmholtz(nx, ny, nz, dx, dy, 20000.0, zeta, zs)
    h.set_coefficients(a, b)
    k = {"bicgstab": vmi.HELM_BICGSTAB, "refine": vmi.HELM_REFINE, "multigrid": vmi.HELM_MULTIGRID}[kind]
    x = h.solve(f, kind=k, tol=1e-10, max_iter=400)
    print(kind, terrain, "iterations", h.iterations, "rel residual", h.residual_rel)
    assert h.residual_rel < 1e-10
    dz = orc.dz_of(zeta, zs, 20000.0)
    xd = orc.solve(a, b, f, dx, dy, dz)
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    assert err < 1e-8, err
    assert orc.residual_l2(a, b, f, x, dx, dy, dz) <= 2e-10 * np.linalg.norm(f)
    h.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_reference_test_cases(vmi, kind):
    # vertical manufactured solution (test_time.cpp:116-152), grid 4x4x16
    nz, ztop = 16, 20000.0
    zeta = _stretched(nz)
    sh = (nz, 4, 4)
    xt = np.broadcast_to(np.sin(np.pi * (np.arange(nz) + 1) / (nz + 1))[:, None, None], sh).copy()
    f = np.empty(sh)
    for k in range(nz):
        d = (zeta[k + 1] - zeta[k]) * ztop
        f[k] = (2.0 / (d * d) + 0.01) * xt[k] - (xt[max(k - 1, 0)] + xt[min(k + 1, nz - 1)]) / (d * d)
    h = vmi.Helmholtz(4, 4, nz, 2000.0, 2000.0, ztop, zeta)
    h.set_coefficients(np.ones(sh), np.full(sh, 0.01))
    x = h.solve(f, kind=kind, tol=1e-10, max_iter=300)
    assert np.max(np.abs(x - xt)) < 1e-8
    assert h.iterations >= 1
    h.close()
    # smooth right-hand side (test_time.cpp:154-171), grid 8x8x8, a = b = 1
    sh = (8, 8, 8)
    j, i = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    f = np.broadcast_to(np.sin(0.4 * i) * np.cos(0.3 * j), sh).copy()
    h = vmi.Helmholtz(8, 8, 8, 2000.0, 2000.0, ztop, _stretched(8))
    h.set_coefficients(np.ones(sh), np.ones(sh))
    x = h.solve(f, kind=kind, tol=1e-10, max_iter=300)
    assert np.all(np.isfinite(x)) and h.residual_rel < 1e-10
    h.close()


def test_bad_arguments(vmi):
    h = vmi.Helmholtz(4, 4, 4, 1.0, 1.0, 10.0, np.linspace(0, 1, 5))
    with pytest.raises(vmi.VibeError) as e:
        h.solve(np.ones((4, 4, 4)))
    assert e.value.code == vmi.VIBE_C_ERR_BAD_ARG  # no coefficients yet
    with pytest.raises(vmi.VibeError) as e:
        h.set_coefficients(np.ones((4, 4, 4)), np.zeros((4, 4, 4)))
    assert e.value.code == vmi.VIBE_C_ERR_BAD_ARG  # b must be positive
    h.close()


def test_large_solve_iterations_and_time(vmi):
    """256x256x64 with terrain: both solvers converge; prints iterations and
    wall time (a measurement, not a bar)."""
    nx, ny, nz = 256, 256, 64
    zeta, zs, a, b, rng = _problem(nx, ny, nz, 3)
    f = rng.normal(size=(nz, ny, nx))
    h = vmi.Helmholtz(nx, ny, nz, 1000.0, 1000.0, 20000.0, zeta, zs)
    h.set_coefficients(a, b)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    smooth = np.sin(2 * np.pi * i / nx) * np.cos(4 * np.pi * j / ny) * np.sin(np.pi * (k + 0.5) / nz)
    for name, rhs in (("noise", f), ("smooth", smooth)):
        for kind in (vmi.HELM_BICGSTAB, vmi.HELM_REFINE, vmi.HELM_MULTIGRID):
            h.solve(rhs, kind=kind, tol=1e-10, max_iter=500)
            t0 = time.perf_counter()
            h.solve(rhs, kind=kind, tol=1e-10, max_iter=500)
            ms = (time.perf_counter() - t0) * 1e3
            print(f"helmholtz 256x256x64 {name} right-hand side, kind {kind}: {h.iterations} iterations, "
                  f"rel residual {h.residual_rel:.2e}, {ms:.1f} ms (host copies included)", flush=True)
            assert h.residual_rel < 1e-10
    h.close()


@pytest.mark.parametrize("shape", [((3, 16, 20), (3, 8, 10)), ((2, 9, 13), (2, 5, 7)), ((4, 8, 8), (4, 8, 8)),
                                   ((1, 1, 5), (2, 3, 3))])
def test_multigrid_transfers_match_oracle_bit_for_bit(vmi, shape):
    """vibe_mi_multigrid_restrict_f64 / _prolong_f64 against the oracle on
    even, odd, same-size (the reference's own use) and degenerate extents."""
    import torch
    fs, cs = shape
    rng = np.random.default_rng(sum(fs) + sum(cs))
    fine = rng.normal(size=fs)
    coarse = rng.normal(size=cs)
    dfine = torch.tensor(fine, device="cuda")
    dco = torch.empty(cs, dtype=torch.float64, device="cuda")
    vmi.multigrid_restrict(dfine, dco)
    torch.cuda.synchronize()
    assert np.array_equal(dco.cpu().numpy(), orc.restrict(fine, cs))
    dco.copy_(torch.tensor(coarse))
    dback = torch.empty(fs, dtype=torch.float64, device="cuda")
    vmi.multigrid_prolong(dco, dback)
    torch.cuda.synchronize()
    assert np.array_equal(dback.cpu().numpy(), orc.prolong(coarse, fs))


def test_multigrid_transfers_reference_norm_test(vmi):
    # tests/unit/test_time.cpp:101-114 on the device
    import torch
    i = np.arange(8.0)
    fine = np.broadcast_to(np.sin(0.7 * i)[None, None, :] + np.cos(0.5 * i)[None, :, None], (4, 8, 8)).copy()
    df = torch.tensor(fine, device="cuda")
    dc = torch.empty((4, 4, 4), dtype=torch.float64, device="cuda")
    db = torch.empty((4, 8, 8), dtype=torch.float64, device="cuda")
    vmi.multigrid_restrict(df, dc)
    vmi.multigrid_prolong(dc, db)
    torch.cuda.synchronize()
    assert torch.isfinite(dc).all()
    assert float(torch.linalg.norm(db)) <= 1.5 * float(np.linalg.norm(fine))
