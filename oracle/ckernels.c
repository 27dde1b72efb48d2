/* oracle/ckernels.c -- CPU restatement of the reference's C11 hot kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Each function restates one routine of the reference's
 * src/common/c_kernels.c (cited per function), written fresh.  The Thomas
 * solver follows the HEADER CONTRACT (include/vibe/common/c_kernels.h:38-55)
 * with the textbook recurrence of src/time/helmholtz.cpp:57-86, not the
 * reference's C implementation, which deviates from its own contract.
 *
 * Pinning (tests/test_oracle_ckernels.py): golden vectors written from the
 * reference's own c_kernels.c (tests/golden/), the known-answer tests of
 * tests/unit/test_common.cpp:163-266, closed forms (sine-solution Thomas,
 * linear-stencil reduction of WENO5), numpy/scipy banded solves and exact
 * rational sums.  Built with -ffp-contract=off so that a*b+c is never fused.
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>

/* Thomas algorithm per the header contract c_kernels.h:38-55:
 *   a[k-1] x[k-1] + b[k] x[k] + c[k] x[k+1] = d[k];  a, c of length n-1.
 * Recurrence as helmholtz.cpp:57-86 (c' kept in ws, d' kept in x).
 * Return codes as c_kernels.c:22-29 / :34 / :40 (exact-zero pivot). */
int orc_tridiagonal_solve(const double *a, const double *b, const double *c,
                          const double *d, int n, double *x, double *ws) {
  if (n <= 0 || b == NULL || d == NULL || x == NULL || ws == NULL) return 2;
  if (n == 1) {
    if (b[0] == 0.0) return 1;
    x[0] = d[0] / b[0];
    return 0;
  }
  double denom = b[0];
  if (denom == 0.0) return 1;
  ws[0] = (c != NULL ? c[0] : 0.0) / denom;
  x[0] = d[0] / denom;
  for (int k = 1; k < n; ++k) {
    const double am = (a != NULL) ? a[k - 1] : 0.0;
    denom = b[k] - am * ws[k - 1];
    if (denom == 0.0) return 1;
    if (k < n - 1) ws[k] = (c != NULL ? c[k] : 0.0) / denom;
    x[k] = (d[k] - am * x[k - 1]) / denom;
  }
  for (int k = n - 2; k >= 0; --k) x[k] = x[k] - ws[k] * x[k + 1];
  return 0;
}

/* WENO5 reconstruction, c_kernels.c:64-104: Jiang-Shu weights
 * d = (0.1, 0.6, 0.3), eps = 1e-6.  q_left is the left-biased value at the
 * face i+1/2; q_right is the mirror-image reconstruction over the same five
 * cells, i.e. the right-biased value at face i-1/2. */
void orc_weno5(double qm2, double qm1, double q0, double q1, double q2,
               double *q_left, double *q_right) {
  const double eps = 1e-6;
  if (q_left == NULL || q_right == NULL) return;
  {
    const double p0 = (2.0 * qm2 - 7.0 * qm1 + 11.0 * q0) / 6.0;
    const double p1 = (-qm1 + 5.0 * q0 + 2.0 * q1) / 6.0;
    const double p2 = (2.0 * q0 + 5.0 * q1 - q2) / 6.0;
    const double s0a = qm2 - 2.0 * qm1 + q0, s0b = qm2 - 4.0 * qm1 + 3.0 * q0;
    const double s1a = qm1 - 2.0 * q0 + q1, s1b = qm1 - q1;
    const double s2a = q0 - 2.0 * q1 + q2, s2b = 3.0 * q0 - 4.0 * q1 + q2;
    const double be0 = (13.0 / 12.0) * s0a * s0a + 0.25 * s0b * s0b;
    const double be1 = (13.0 / 12.0) * s1a * s1a + 0.25 * s1b * s1b;
    const double be2 = (13.0 / 12.0) * s2a * s2a + 0.25 * s2b * s2b;
    const double w0 = 0.1 / ((eps + be0) * (eps + be0));
    const double w1 = 0.6 / ((eps + be1) * (eps + be1));
    const double w2 = 0.3 / ((eps + be2) * (eps + be2));
    *q_left = (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
  }
  {
    const double p0 = (2.0 * q2 - 7.0 * q1 + 11.0 * q0) / 6.0;
    const double p1 = (-q1 + 5.0 * q0 + 2.0 * qm1) / 6.0;
    const double p2 = (2.0 * q0 + 5.0 * qm1 - qm2) / 6.0;
    const double s0a = q2 - 2.0 * q1 + q0, s0b = q2 - 4.0 * q1 + 3.0 * q0;
    const double s1a = q1 - 2.0 * q0 + qm1, s1b = q1 - qm1;
    const double s2a = q0 - 2.0 * qm1 + qm2, s2b = 3.0 * q0 - 4.0 * qm1 + qm2;
    const double be0 = (13.0 / 12.0) * s0a * s0a + 0.25 * s0b * s0b;
    const double be1 = (13.0 / 12.0) * s1a * s1a + 0.25 * s1b * s1b;
    const double be2 = (13.0 / 12.0) * s2a * s2a + 0.25 * s2b * s2b;
    const double w0 = 0.1 / ((eps + be0) * (eps + be0));
    const double w1 = 0.6 / ((eps + be1) * (eps + be1));
    const double w2 = 0.3 / ((eps + be2) * (eps + be2));
    *q_right = (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
  }
}

/* Bilinear interpolation on a halo'd field, c_kernels.c:108-144: indices
 * clamped to the interior, weights from the fractional cell-centre index. */
double orc_bilinear(const double *f, int halo, int nsx, int nsy, int nz,
                    double xi, double yj, int k) {
  if (f == NULL || nz <= 0) return 0.0;
  const int nxi = nsx - 2 * halo, nyi = nsy - 2 * halo;
  if (k < 0) k = 0;
  if (k >= nz) k = nz - 1;
  const int i0 = (int)floor(xi), j0 = (int)floor(yj);
  const double wx = xi - (double)i0, wy = yj - (double)j0;
  int ia = i0, ib = i0 + 1, ja = j0, jb = j0 + 1;
  ia = ia < 0 ? 0 : (ia > nxi - 1 ? nxi - 1 : ia);
  ib = ib < 0 ? 0 : (ib > nxi - 1 ? nxi - 1 : ib);
  ja = ja < 0 ? 0 : (ja > nyi - 1 ? nyi - 1 : ja);
  jb = jb < 0 ? 0 : (jb > nyi - 1 ? nyi - 1 : jb);
#define ORC_AT(I, J)                                                       \
  f[(((size_t)(k + halo) * (size_t)nsy + (size_t)((J) + halo)) * (size_t)nsx) + \
    (size_t)((I) + halo)]
  const double q00 = ORC_AT(ia, ja), q10 = ORC_AT(ib, ja);
  const double q01 = ORC_AT(ia, jb), q11 = ORC_AT(ib, jb);
#undef ORC_AT
  const double lo = q00 + wx * (q10 - q00);
  const double hi = q01 + wx * (q11 - q01);
  return lo + wy * (hi - lo);
}

/* CFL scan, c_kernels.c:148-187. */
double orc_cfl_scan(const double *u, const double *v, const double *w,
                    const double *dzeta, const double *zs, int nx, int ny, int nz,
                    double dx, double dy, double z_top, double dt,
                    double *max_wind_out, double *min_dz_out) {
  double worst = 0.0, max_wind = 0.0, min_dz = 1e30;
  if (u == NULL || v == NULL || w == NULL || dzeta == NULL) return 0.0;
  const size_t nxy = (size_t)nx * (size_t)ny;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t col = (size_t)j * (size_t)nx + (size_t)i;
        const size_t idx = (size_t)k * nxy + col;
        const double scale = zs != NULL ? z_top - zs[col] : z_top;
        double dz = dzeta[k] * scale;
        if (dz < 1.0) dz = 1.0;
        if (dz < min_dz) min_dz = dz;
        const double uu = u[idx], vv = v[idx], ww = w[idx];
        const double speed = sqrt(uu * uu + vv * vv + ww * ww);
        if (speed > max_wind) max_wind = speed;
        const double cfl = dt * (fabs(uu) / dx + fabs(vv) / dy + fabs(ww) / dz);
        if (cfl > worst) worst = cfl;
      }
  if (max_wind_out != NULL) *max_wind_out = max_wind;
  if (min_dz_out != NULL) *min_dz_out = (min_dz > 1e29) ? 0.0 : min_dz;
  return worst;
}

/* Sequential Neumaier sum, c_kernels.c:190-206. */
double orc_neumaier_sum(const double *x, size_t n) {
  double s = 0.0, comp = 0.0;
  if (x == NULL) return 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double t = s + x[i];
    if (fabs(s) >= fabs(x[i]))
      comp += (s - t) + x[i];
    else
      comp += (x[i] - t) + s;
    s = t;
  }
  return s + comp;
}

/* Second-order flux-form scalar advection, c_kernels.c:209-286: periodic
 * modulo neighbours in x and y, vertical fluxes only for 0 < k < nz-1,
 * dz >= 1, rho floor 1e-12. */
void orc_advect_scalar_flux(const double *q, const double *rho, const double *u,
                            const double *v, const double *w, const double *dzeta,
                            const double *zs, int nx, int ny, int nz, double dx,
                            double dy, double z_top, double *dqdt) {
  if (q == NULL || rho == NULL || u == NULL || v == NULL || w == NULL ||
      dzeta == NULL || dqdt == NULL)
    return;
  const size_t nxy = (size_t)nx * (size_t)ny;
  const double rdx = 1.0 / dx, rdy = 1.0 / dy;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t col = (size_t)j * (size_t)nx + (size_t)i;
        const size_t c = (size_t)k * nxy + col;
        const double scale = zs != NULL ? z_top - zs[col] : z_top;
        double dz = dzeta[k] * scale;
        if (dz < 1.0) dz = 1.0;
        const size_t e = (size_t)k * nxy + (size_t)j * nx + (size_t)((i + 1) % nx);
        const size_t wv = (size_t)k * nxy + (size_t)j * nx + (size_t)((i + nx - 1) % nx);
        const size_t nn = (size_t)k * nxy + (size_t)((j + 1) % ny) * nx + (size_t)i;
        const size_t ss = (size_t)k * nxy + (size_t)((j + ny - 1) % ny) * nx + (size_t)i;
        const double fxe = u[e] * (0.5 * (rho[c] + rho[e])) * (0.5 * (q[c] + q[e]));
        const double fxw = u[c] * (0.5 * (rho[wv] + rho[c])) * (0.5 * (q[wv] + q[c]));
        const double fyn = v[nn] * (0.5 * (rho[c] + rho[nn])) * (0.5 * (q[c] + q[nn]));
        const double fys = v[c] * (0.5 * (rho[ss] + rho[c])) * (0.5 * (q[ss] + q[c]));
        double fzt = 0.0, fzb = 0.0;
        if (k > 0 && k + 1 < nz) {
          const size_t t = c + nxy, b = c - nxy;
          fzt = w[t] * (0.5 * (rho[c] + rho[t])) * (0.5 * (q[c] + q[t]));
          fzb = w[c] * (0.5 * (rho[b] + rho[c])) * (0.5 * (q[b] + q[c]));
        }
        const double div = (fxe - fxw) * rdx + (fyn - fys) * rdy + (fzt - fzb) / dz;
        const double rr = rho[c] > 1e-12 ? rho[c] : 1e-12;
        dqdt[c] = -div / rr;
      }
}
