"""The no-allocation batched Thomas entry point
(vibe_mi_tridiagonal_solve_batched_noalloc_f64, static device scratch)
against the allocating one it varies (vibe_mi_tridiagonal_solve_batched_f64,
caller workspace): bit-identical solutions and return codes on two calls in
a row, the second smaller than the first (the scratch keeps the first call's
contents), a singular system reported in the same way, and an input beyond
the scratch capacity rejected before the GPU is touched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Dev:
    """hipMalloc'd copies of numpy arrays (freed on close)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(arr.nbytes, 8))) == 0
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0
        self.bufs.append(p)
        return p

    def get(self, p, shape):
        out = np.zeros(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _system(rng, n, nsys, singular_at=None):
    a = rng.uniform(-1, 1, (n - 1, nsys))
    c = rng.uniform(-1, 1, (n - 1, nsys))
    b = 3.0 + rng.uniform(0, 1, (n, nsys))
    d = rng.normal(size=(n, nsys))
    if singular_at is not None:
        b[0, singular_at] = 0.0
    return a, b, c, d


def _both(vmi, dev, n, nsys, a, b, c, d):
    L = vmi.lib()
    da, db, dc, dd = (dev.put(q) for q in (a, b, c, d))
    x1, ws = dev.put(np.zeros((n, nsys))), dev.put(np.zeros((n, nsys)))
    x2 = dev.put(np.full((n, nsys), np.nan))
    rc1 = L.vibe_mi_tridiagonal_solve_batched_f64(da, db, dc, dd, n, nsys, x1, ws, None)
    rc2 = L.vibe_mi_tridiagonal_solve_batched_noalloc_f64(da, db, dc, dd, n, nsys, x2, None)
    return rc1, rc2, dev.get(x1, (n, nsys)), dev.get(x2, (n, nsys))


def test_noalloc_thomas_equals_allocating_entry_two_calls(vmi):
    rng = np.random.default_rng(21)
    dev = _Dev()
    try:
        # first call large (96 levels x 40000 columns), second smaller
        for n, nsys in ((96, 40000), (40, 1000)):
            rc1, rc2, x1, x2 = _both(vmi, dev, n, nsys, *_system(rng, n, nsys))
            assert rc1 == 0 and rc2 == 0, (n, nsys, rc1, rc2)
            assert np.array_equal(x1, x2), (n, nsys)
            assert np.all(np.isfinite(x2))
        # singular pivot: both report it, the other systems agree
        n, nsys = 24, 300
        rc1, rc2, x1, x2 = _both(vmi, dev, n, nsys, *_system(rng, n, nsys, singular_at=7))
        assert rc1 == vmi.VIBE_C_ERR_SINGULAR and rc2 == vmi.VIBE_C_ERR_SINGULAR
        keep = np.arange(nsys) != 7
        assert np.array_equal(x1[:, keep], x2[:, keep])
        # and a regular call right after the singular one is clean again
        rc1, rc2, x1, x2 = _both(vmi, dev, n, nsys, *_system(rng, n, nsys))
        assert rc1 == 0 and rc2 == 0 and np.array_equal(x1, x2)
    finally:
        dev.close()


def test_noalloc_thomas_rejects_inputs_beyond_scratch(vmi):
    L = vmi.lib()
    cap = L.vibe_mi_noalloc_scratch_capacity()
    assert cap == (32 * 2 ** 20 - 256) // 8
    dev = _Dev()
    try:
        p = dev.put(np.zeros(8))
        n = 97
        nsys = cap // n + 1  # n * nsys > cap; no kernel is launched for it
        rc = L.vibe_mi_tridiagonal_solve_batched_noalloc_f64(p, p, p, p, n, nsys, p, None)
        assert rc == vmi.VIBE_MI_ERR_TOO_LARGE
        assert L.vibe_mi_tridiagonal_solve_batched_noalloc_f64(p, p, p, p, 0, 4, p, None) == vmi.VIBE_C_ERR_BAD_ARG
    finally:
        dev.close()
