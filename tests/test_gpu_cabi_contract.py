"""The six C-ABI drop-ins under the reference's contract (c_kernels.h:13-19:
no allocation, caller-provided workspace, no global state, re-entrant and
safe from concurrent threads) on the GPU.

* In place: device arrays and hipHostMalloc'd (pinned) host arrays are read
  and written where they are; the results equal those of pageable host
  arrays (copied through the library's static staging array) bit for bit.
* Chunked staging: CFL scans, flux advection and Neumaier sums larger than
  one staging chunk (several chunks and level slabs; zs in point order when
  the terrain plane is too large to stage whole) equal the oracle and the
  in-place results bit for bit.
* No allocation: after a first call of each entry point (which loads the
  code object and the runtime's per-thread stream), a sweep of calls of every
  entry point at sizes from 1 to 3e7 values leaves the GPU's free memory
  unchanged (hipMemGetInfo).
* Concurrency: 8 host threads calling the entry points at once get the
  single-thread results bit for bit.
* Latency: per-call time of each entry point (pageable host arrays, the
  reference's own test sizes) and of the reference's compiled C on the same
  inputs, printed; each GPU call stays under 2 ms.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from oracle_bind import _pd, oracle_lib, ref_lib

pytestmark = pytest.mark.gpu

PD = ctypes.POINTER(ctypes.c_double)


def say(*a):
    print("[cabi]", *a, flush=True)


def _dev(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), PD)


def _torch_copies(arr):
    import torch
    h = torch.from_numpy(np.ascontiguousarray(arr))
    return h.to("cuda"), h.pin_memory()


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert _hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_in_place_equals_staged(vmi):
    import torch
    L = vmi.lib()
    rng = np.random.default_rng(11)
    # Thomas
    n = 96
    a, c = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - 1)
    b, d = 3 + rng.uniform(0, 1, n), rng.normal(size=n)
    rc, x_host = vmi.tridiagonal_solve(a, b, c, d)
    assert rc == 0
    for kind in (0, 1):  # device, pinned
        ins = [_torch_copies(q)[kind] for q in (a, b, c, d)]
        x = _torch_copies(np.zeros(n))[kind]
        ws = _torch_copies(np.zeros(n))[kind]
        rc = L.vibe_tridiagonal_solve_f64(*[_dev(t) for t in ins], n, _dev(x), _dev(ws))
        torch.cuda.synchronize()
        assert rc == 0 and np.array_equal(x.cpu().numpy(), x_host), kind
    # a singular system in place reports it
    z = torch.zeros(2, dtype=torch.float64, device="cuda")
    one = torch.ones(2, dtype=torch.float64, device="cuda")
    xx, ww = torch.zeros_like(z), torch.zeros_like(z)
    assert L.vibe_tridiagonal_solve_f64(None, _dev(z), None, _dev(one), 2, _dev(xx), _dev(ww)) == 1
    # CFL, flux, Neumaier, bilinear on a small grid
    nx, ny, nz = 13, 9, 7
    m = nx * ny * nz
    u, v, w = rng.normal(size=m) * 10, rng.normal(size=m) * 10, rng.normal(size=m)
    q, rho = rng.uniform(1, 2, m), rng.uniform(0.5, 1.2, m)
    dz, zs = rng.uniform(0.05, 0.2, nz), rng.uniform(0, 500, nx * ny)
    want_cfl = vmi.cfl_scan(u, v, w, dz, zs, nx, ny, nz, 1000., 900., 20000., 3.)
    want_flux = vmi.advect_scalar_flux(q, rho, u, v, w, dz, zs, nx, ny, nz, 1000., 900., 20000.)
    want_sum = vmi.neumaier_sum(u)
    halo, bnx, bny, bnz = 2, 5, 4, 2
    f = rng.normal(size=(bnx + 2 * halo) * (bny + 2 * halo) * (bnz + 2 * halo))
    want_bil = vmi.bilinear(f, halo, bnx + 2 * halo, bny + 2 * halo, bnz, 3.7, 2.2, 1)
    for kind in (0, 1):
        tu, tv, tw, tq, tr, tz, ts, tf = (_torch_copies(x)[kind] for x in (u, v, w, q, rho, dz, zs, f))
        mw, md = ctypes.c_double(), ctypes.c_double()
        got = L.vibe_cfl_scan_f64(_dev(tu), _dev(tv), _dev(tw), _dev(tz), _dev(ts), nx, ny, nz, 1000., 900.,
                                  20000., 3., ctypes.byref(mw), ctypes.byref(md))
        assert (got, mw.value, md.value) == want_cfl, kind
        out = _torch_copies(np.zeros(m))[kind]
        L.vibe_advect_scalar_flux_f64(_dev(tq), _dev(tr), _dev(tu), _dev(tv), _dev(tw), _dev(tz), _dev(ts), nx, ny,
                                      nz, 1000., 900., 20000., _dev(out))
        torch.cuda.synchronize()
        assert L.vibe_mi_last_error() == 0
        assert np.array_equal(out.cpu().numpy(), want_flux), kind
        assert L.vibe_neumaier_sum_f64(_dev(tu), m) == want_sum, kind
        assert L.vibe_bilinear_f64(_dev(tf), halo, bnx + 2 * halo, bny + 2 * halo, bnz, 3.7, 2.2, 1) == want_bil
    # an interior pointer of a pinned array is used in place as well
    pinned = torch.from_numpy(np.concatenate([np.zeros(5), u])).pin_memory()
    sub = ctypes.cast(ctypes.c_void_p(pinned.data_ptr() + 5 * 8), PD)
    assert L.vibe_neumaier_sum_f64(sub, m) == want_sum


def test_chunked_staging_matches_oracle_and_in_place(vmi):
    import torch
    O = oracle_lib()
    L = vmi.lib()
    rng = np.random.default_rng(12)
    # CFL: 4096 x 2304 x 2 = 1.9e7 points; zs plane 75 MB (> a quarter of
    # the staging body: staged in point order) and several chunks
    nx, ny, nz = 4096, 2304, 2
    m = nx * ny * nz
    say("cfl", nx, ny, nz)
    u = rng.normal(size=m) * 10
    v = rng.normal(size=m) * 10
    w = rng.normal(size=m)
    dz = rng.uniform(0.05, 0.2, nz)
    zs = rng.uniform(0, 500, nx * ny)
    mw, md = ctypes.c_double(), ctypes.c_double()
    o = O.orc_cfl_scan(_pd(u), _pd(v), _pd(w), _pd(dz), _pd(zs), nx, ny, nz, 1000., 900., 20000., 3.,
                       ctypes.byref(mw), ctypes.byref(md))
    g = vmi.cfl_scan(u, v, w, dz, zs, nx, ny, nz, 1000., 900., 20000., 3.)
    assert g == (o, mw.value, md.value)
    tu, tv, tw, tz, ts = (torch.from_numpy(x).to("cuda") for x in (u, v, w, dz, zs))
    got = L.vibe_cfl_scan_f64(_dev(tu), _dev(tv), _dev(tw), _dev(tz), _dev(ts), nx, ny, nz, 1000., 900., 20000.,
                              3., ctypes.byref(mw), ctypes.byref(md))
    assert (got, mw.value, md.value) == g
    del tu, tv, tw, tz, ts
    # Neumaier over 2.4e7 values: three chunks of 2^23
    say("neumaier", u.size + v.size // 4)
    x = np.concatenate([u, v[: v.size // 4]]) * 10.0 ** rng.integers(-6, 6, size=u.size + v.size // 4)
    s_host = vmi.neumaier_sum(x)
    tx = torch.from_numpy(x).to("cuda")
    assert L.vibe_neumaier_sum_f64(_dev(tx), x.size) == s_host
    out = ctypes.c_double()
    assert L.vibe_mi_neumaier_sum_device_f64(_dev(tx), x.size, ctypes.byref(out), None) == 0
    assert out.value == s_host
    exact = float(np.sum(x.astype(np.longdouble)))  # 64-bit-mantissa reference
    assert abs(s_host - exact) <= 1e-12 * np.sum(np.abs(x))
    del tx
    # flux advection: 512 x 512 x 40, level slabs of 19 levels (3 slabs)
    nx, ny, nz = 512, 512, 40
    m = nx * ny * nz
    say("flux", nx, ny, nz)
    q, rho = rng.uniform(1, 2, m), rng.uniform(0.5, 1.2, m)
    u, v, w = rng.normal(size=m) * 10, rng.normal(size=m) * 10, rng.normal(size=m)
    dz, zs = rng.uniform(0.01, 0.05, nz), rng.uniform(0, 500, nx * ny)
    want = np.zeros(m)
    O.orc_advect_scalar_flux(_pd(q), _pd(rho), _pd(u), _pd(v), _pd(w), _pd(dz), _pd(zs), nx, ny, nz, 1000., 900.,
                             20000., _pd(want))
    gf = vmi.advect_scalar_flux(q, rho, u, v, w, dz, zs, nx, ny, nz, 1000., 900., 20000.)
    assert np.array_equal(gf, want)
    ts = [torch.from_numpy(a).to("cuda") for a in (q, rho, u, v, w, dz, zs)]
    tout = torch.zeros(m, dtype=torch.float64, device="cuda")
    L.vibe_advect_scalar_flux_f64(*[_dev(t) for t in ts], nx, ny, nz, 1000., 900., 20000., _dev(tout))
    torch.cuda.synchronize()
    assert np.array_equal(tout.cpu().numpy(), want)


def test_no_allocation_across_calls(vmi):
    rng = np.random.default_rng(13)
    L = vmi.lib()

    def sweep(sizes):
        for n in sizes:
            b = 3 + rng.uniform(0, 1, n)
            vmi.tridiagonal_solve(rng.uniform(-1, 1, n - 1), b, rng.uniform(-1, 1, n - 1), rng.normal(size=n))
            vmi.weno5_reconstruct(*rng.normal(size=5))
            x = rng.normal(size=n * 37)
            vmi.neumaier_sum(x)
            nz = max(1, n // 64)
            nx = ny = 8
            m = nx * ny * nz
            u = rng.normal(size=m)
            vmi.cfl_scan(u, u, u, np.full(nz, 1.0 / nz), None, nx, ny, nz, 1e3, 1e3, 1e4, 1.0)
            vmi.advect_scalar_flux(1 + u * 0, 1 + u * 0, u, u, u, np.full(nz, 1.0 / nz), None, nx, ny, nz, 1e3, 1e3,
                                   1e4)
            f = rng.normal(size=9 * 8 * 6)
            vmi.bilinear(f, 2, 9, 8, 2, 1.5, 1.2, 0)
        assert L.vibe_mi_last_error() == 0

    sweep([2])  # first use: code object, per-thread stream
    before = _free_bytes()
    sweep([1, 7, 96, 4096, 100000, 800000])
    after = _free_bytes()
    say("free before", before, "after", after)
    assert after == before


def test_concurrent_threads(vmi):
    rng = np.random.default_rng(14)
    cases = []
    for t in range(8):
        n = 50 + 13 * t
        cases.append((rng.uniform(-1, 1, n - 1), 3 + rng.uniform(0, 1, n), rng.uniform(-1, 1, n - 1),
                      rng.normal(size=n), rng.normal(size=20000 + t)))
    want = [(vmi.tridiagonal_solve(a, b, c, d)[1], vmi.neumaier_sum(x)) for a, b, c, d, x in cases]
    got = [None] * len(cases)
    errors = []

    def work(i):
        try:
            a, b, c, d, x = cases[i]
            for _ in range(20):
                r = (vmi.tridiagonal_solve(a, b, c, d)[1], vmi.neumaier_sum(x))
                if not (np.array_equal(r[0], want[i][0]) and r[1] == want[i][1]):
                    errors.append(i)
            got[i] = True
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=60)
    assert not errors and all(got)


def _time(fn, reps):
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return (time.perf_counter() - t0) / reps


def test_per_call_latency_against_reference_c(vmi):
    R = ref_lib()
    L = vmi.lib()
    rng = np.random.default_rng(15)
    n = 40
    a, c = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - 1)
    b, d = 3 + rng.uniform(0, 1, n), rng.normal(size=n)
    x, ws = np.zeros(n), np.zeros(n)
    q5 = [float(t) for t in rng.normal(size=5)]
    l, r = ctypes.c_double(), ctypes.c_double()
    nx, ny, nz = 8, 6, 4  # the reference's test sizes (test_common.cpp:227-266)
    m = nx * ny * nz
    u = rng.normal(size=m)
    dz = np.full(nz, 0.25)
    out = np.zeros(m)
    one = np.ones(m)
    f = rng.normal(size=9 * 8 * 6)
    s = rng.normal(size=10000)
    mw, md = ctypes.c_double(), ctypes.c_double()
    calls = {
        "tridiagonal_solve n=40": lambda lib: lib.vibe_tridiagonal_solve_f64(_pd(a), _pd(b), _pd(c), _pd(d), n,
                                                                            _pd(x), _pd(ws)),
        "weno5_reconstruct": lambda lib: lib.vibe_weno5_reconstruct_f64(*q5, ctypes.byref(l), ctypes.byref(r)),
        "bilinear": lambda lib: lib.vibe_bilinear_f64(_pd(f), 2, 9, 8, 2, 1.5, 1.2, 0),
        "cfl_scan 8x6x4": lambda lib: lib.vibe_cfl_scan_f64(_pd(u), _pd(u), _pd(u), _pd(dz), None, nx, ny, nz, 1e3,
                                                            1e3, 1e4, 1.0, ctypes.byref(mw), ctypes.byref(md)),
        "neumaier_sum n=1e4": lambda lib: lib.vibe_neumaier_sum_f64(_pd(s), s.size),
        "advect_scalar_flux 8x6x4": lambda lib: lib.vibe_advect_scalar_flux_f64(
            _pd(one), _pd(one), _pd(u), _pd(u), _pd(u), _pd(dz), None, nx, ny, nz, 1e3, 1e3, 1e4, _pd(out)),
    }
    rows = []
    for name, fn in calls.items():
        g = _time(lambda: fn(L), 200)
        rr = _time(lambda: fn(R), 2000) if R is not None else float("nan")
        rows.append((name, g, rr))
        say(f"{name:28s} gpu {g * 1e6:9.1f} us/call   reference C {rr * 1e6:9.3f} us/call")
        assert g < 2e-3, name
    assert L.vibe_mi_last_error() == 0
