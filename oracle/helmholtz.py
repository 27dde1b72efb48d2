"""CPU restatement (numpy) of the reference's 3-D Helmholtz operator.

Test infrastructure only (tests/ import it as the checker of
libvibe_mi.so's vibe_mi_helmholtz_* entry points).

Follows:
  apply_operator          src/time/helmholtz.cpp:173-201: y = -div(a grad x) + b x,
                          face a = mean of the two cells, neighbours clamped at
                          every boundary, dz = max(dzeta_k (ztop - zs), 1)
  helmholtz_residual_l2   src/time/helmholtz.cpp:211-225
  multigrid_restrict      src/time/helmholtz.cpp:131-145: 2:1 full weighting
  multigrid_prolong       src/time/helmholtz.cpp:147-164: bilinear, i -> i/2
and, for a reference solution, the operator assembled as a sparse matrix
and solved directly (scipy.sparse.linalg.spsolve, a library routine).

Pinned (tests/test_oracle_helmholtz.py) by the reference's unit-test
values (tests/unit/test_time.cpp:116-185: residual sqrt(64) of x = 0,
f = 1 on 4x4x4; the vertical manufactured solution) and by closed forms
(constant x gives b x; second differences of quadratics).  The transfers
are pinned by the reference's own test (restrict then prolong does not grow
the norm by more than 1.5, tests/unit/test_time.cpp:101-114) and by what
the mathematics fixes: both reproduce constants (the weights sum to one),
restriction of a field linear in (i, j) is its value at (2i, 2j) away from
the clamped edges, prolongation reproduces a field bilinear in the coarse
indices exactly inside the coarse extent.  The reference's helmholtz.cpp
does not compile from its sources here ('Precision' has not been declared,
helmholtz.cpp:431), so no compiled golden exists.
"""
from __future__ import annotations

import numpy as np


def dz_of(zeta, zs, ztop):
    """dz[k, j, i] = max(dzeta_k (ztop - zs_ij), 1)."""
    dzeta = np.diff(np.asarray(zeta, dtype=np.float64))
    zs = np.asarray(zs, dtype=np.float64)
    return np.maximum(dzeta[:, None, None] * (ztop - zs)[None, :, :], 1.0)


def _sh(f, axis, d):
    """f at index +d along axis with the index clamped (clamp_at)."""
    n = f.shape[axis]
    idx = np.clip(np.arange(n) + d, 0, n - 1)
    return np.take(f, idx, axis=axis)


def apply(a, b, x, dx, dy, dz):
    """y = -div(a grad x) + b x on [nz, ny, nx] arrays (same operation order
    as the reference's loop body)."""
    a = np.asarray(a, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    ae = 0.5 * (a + _sh(a, 2, 1))
    aw = 0.5 * (a + _sh(a, 2, -1))
    an = 0.5 * (a + _sh(a, 1, 1))
    as_ = 0.5 * (a + _sh(a, 1, -1))
    at = 0.5 * (a + _sh(a, 0, 1))
    ab = 0.5 * (a + _sh(a, 0, -1))
    fx = (ae * (_sh(x, 2, 1) - x) - aw * (x - _sh(x, 2, -1))) / (dx * dx)
    fy = (an * (_sh(x, 1, 1) - x) - as_ * (x - _sh(x, 1, -1))) / (dy * dy)
    fz = (at * (_sh(x, 0, 1) - x) - ab * (x - _sh(x, 0, -1))) / (dz * dz)
    return -fx - fy - fz + b * x


def residual_l2(a, b, f, x, dx, dy, dz):
    """||A x - f||_2 (helmholtz_residual_l2)."""
    r = apply(a, b, x, dx, dy, dz) - f
    return float(np.sqrt(np.sum(r * r)))


def matrix(a, b, dx, dy, dz):
    """The operator as a scipy sparse matrix (column n = A e_n)."""
    import scipy.sparse as sp
    nz, ny, nx = np.shape(a)
    n = nz * ny * nx
    rows, cols, vals = [], [], []
    # A is a 7-point stencil: apply it to 7-coloured probe vectors would be
    # cheaper; small test grids afford one column per unit vector
    e = np.zeros(n)
    for q in range(n):
        e[q] = 1.0
        col = apply(a, b, e.reshape(nz, ny, nx), dx, dy, dz).reshape(-1)
        nzr = np.nonzero(col)[0]
        rows.extend(nzr.tolist())
        cols.extend([q] * len(nzr))
        vals.extend(col[nzr].tolist())
        e[q] = 0.0
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def solve(a, b, f, dx, dy, dz):
    """Direct solution of A x = f (scipy spsolve)."""
    import scipy.sparse.linalg as sla
    A = matrix(a, b, dx, dy, dz)
    return sla.spsolve(A.tocsc(), np.asarray(f, dtype=np.float64).reshape(-1)).reshape(np.shape(a))


def _clamped(f, k, j, i):
    """Field::clamp_at of a [nz, ny, nx] array at integer index arrays."""
    nz, ny, nx = f.shape
    return f[np.clip(k, 0, nz - 1), np.clip(j, 0, ny - 1), np.clip(i, 0, nx - 1)]


def restrict(fine, coarse_shape):
    """multigrid_restrict into a field of coarse_shape = (nz, ny, nx); the
    reference's expression, summed in its order."""
    fine = np.asarray(fine, dtype=np.float64)
    k, j, i = np.meshgrid(*[np.arange(n) for n in coarse_shape], indexing="ij")
    fi, fj = 2 * i, 2 * j
    at = lambda di, dj: _clamped(fine, k, fj + dj, fi + di)  # noqa: E731
    acc = 0.25 * at(0, 0)
    acc = acc + 0.125 * (((at(-1, 0) + at(1, 0)) + at(0, -1)) + at(0, 1))
    acc = acc + 0.0625 * (((at(-1, -1) + at(1, -1)) + at(-1, 1)) + at(1, 1))
    return acc


def prolong(coarse, fine_shape):
    """multigrid_prolong into a field of fine_shape = (nz, ny, nx)."""
    coarse = np.asarray(coarse, dtype=np.float64)
    k, j, i = np.meshgrid(*[np.arange(n) for n in fine_shape], indexing="ij")
    xs, ys = 0.5 * i, 0.5 * j
    i0, j0 = np.floor(xs).astype(np.int64), np.floor(ys).astype(np.int64)
    wx, wy = xs - i0, ys - j0
    lerp = lambda a, b, w: a + w * (b - a)  # noqa: E731
    c = lambda di, dj: _clamped(coarse, k, j0 + dj, i0 + di)  # noqa: E731
    return lerp(lerp(c(0, 0), c(1, 0), wx), lerp(c(0, 1), c(1, 1), wx), wy)
