/* oracle/nest.c -- CPU restatement of the nest lateral boundary (test
 * infrastructure only).
 *
 * Follows src/grid/nest.cpp of the reference:
 *   davies_profile            nest.cpp:34-43
 *   interpolate_from_parent   nest.cpp:52-83 (child centre at parent index
 *                             (i0 + i + 0.5)/r - 0.5, bilinear lerp form of
 *                             c_kernels.c:108-144)
 *   apply_lateral_boundary    nest.cpp:110-128 (d = min distance to the
 *                             edge, c = (1 - alpha(d)) c + alpha(d) p)
 *   restrict_to_parent        nest.cpp:85-108 (r x r mean of the child cells
 *                             over each covered parent cell, summed jj-outer
 *                             and ii-inner, times 1/r^2)
 * with two choices of the library: parent indices wrap (periodic parent)
 * instead of clamping, and u / v are sampled at their own staggered
 * positions (u: xp = (i0 + i)/r, v: yp = (j0 + j)/r).  The parent value is
 * interpolated in time between its previous and current state,
 * p = p_old + frac (p_new - p_old) (frac = 1: p_new alone).
 */
#include <math.h>
#include <stddef.h>

#include "oracle.h"

void orc_davies_profile(int zone, double alpha_max, double *alpha) {
  const double pi = 3.14159265358979323846;
  for (int n = 0; n < zone; ++n)
    alpha[n] = alpha_max * 0.5 * (1.0 - cos(pi * (double)(zone - n) / (double)zone));
}

static int wrap_mod(int i, int n) {
  const int m = i % n;
  return m < 0 ? m + n : m;
}

static double bilinear(const double *f, int nx, int ny, double xp, double yp) {
  const double fx = floor(xp), fy = floor(yp);
  const int ia = (int)fx, ja = (int)fy;
  const double wx = xp - fx, wy = yp - fy;
  const int i0 = wrap_mod(ia, nx), i1 = wrap_mod(ia + 1, nx);
  const size_t r0 = (size_t)wrap_mod(ja, ny) * nx, r1 = (size_t)wrap_mod(ja + 1, ny) * nx;
  const double q00 = f[r0 + i0], q10 = f[r0 + i1], q01 = f[r1 + i0], q11 = f[r1 + i1];
  return (q00 + wx * (q10 - q00)) + wy * ((q01 + wx * (q11 - q01)) - (q00 + wx * (q10 - q00)));
}

void orc_nest_apply(const orc_state *pnew, const orc_state *pold, int pnx, int pny, orc_state *child, int cnx,
                    int cny, int nz, int r, int i0, int j0, int zone, double alpha_max, double frac, int full) {
  double alpha[64];
  if (zone > 64) zone = 64;
  orc_davies_profile(zone, alpha_max, alpha);
  const double *pn[6] = {pnew->u, pnew->v, pnew->w, pnew->rho, pnew->theta, pnew->pi};
  const double *po[6] = {pold ? pold->u : 0, pold ? pold->v : 0, pold ? pold->w : 0,
                         pold ? pold->rho : 0, pold ? pold->theta : 0, pold ? pold->pi : 0};
  double *c[6] = {child->u, child->v, child->w, child->rho, child->theta, child->pi};
  const size_t pxy = (size_t)pnx * pny, cxy = (size_t)cnx * cny;
  const double rinv = 1.0 / (double)r;
  for (int j = 0; j < cny; ++j)
    for (int i = 0; i < cnx; ++i) {
      int d = i;
      if (cnx - 1 - i < d) d = cnx - 1 - i;
      if (j < d) d = j;
      if (cny - 1 - j < d) d = cny - 1 - j;
      if (!full && d >= zone) continue;
      const double al = full ? 1.0 : alpha[d];
      const double xc = ((double)(i0 + i) + 0.5) * rinv - 0.5;
      const double yc = ((double)(j0 + j) + 0.5) * rinv - 0.5;
      const double xu = (double)(i0 + i) * rinv;
      const double yv = (double)(j0 + j) * rinv;
      for (int f = 0; f < 6; ++f) {
        const double xp = f == 0 ? xu : xc;
        const double yp = f == 1 ? yv : yc;
        const int nk = f == 2 ? nz + 1 : nz;
        for (int k = 0; k < nk; ++k) {
          const double a = bilinear(pn[f] + k * pxy, pnx, pny, xp, yp);
          double p = a;
          if (!full && frac != 1.0) {
            const double b = bilinear(po[f] + k * pxy, pnx, pny, xp, yp);
            p = b + frac * (a - b);
          }
          double *cf = c[f] + k * cxy + (size_t)j * cnx + i;
          *cf = full ? p : (1.0 - al) * *cf + al * p;
        }
      }
    }
}

/* Two-way feedback (nest.cpp:85-108): parent cell (i0/r + i, j0/r + j),
 * i < cnx/r, j < cny/r, takes the r x r mean of the child cells over it for
 * rho', theta', pi' and w (w sits at cell centres horizontally); u (west
 * face) takes the mean of the r child faces on the parent face,
 * u(i r, j r + jj), and v the mean of v(i r + ii, j r).  The reference
 * averages every field as a cell field and writes to parent index
 * i0 + i with i0 read as parent cells; here i0, j0 are child cells, as in
 * the interpolation, and must be multiples of r.  Parent indices wrap. */
void orc_nest_restrict(const orc_state *child, int cnx, int cny, orc_state *parent, int pnx, int pny, int nz,
                       int r, int i0, int j0) {
  const double *c[6] = {child->u, child->v, child->w, child->rho, child->theta, child->pi};
  double *p[6] = {parent->u, parent->v, parent->w, parent->rho, parent->theta, parent->pi};
  const size_t pxy = (size_t)pnx * pny, cxy = (size_t)cnx * cny;
  const int iext = cnx / r, jext = cny / r, I0 = i0 / r, J0 = j0 / r;
  const double w2 = 1.0 / (double)(r * r), w1 = 1.0 / (double)r;
  for (int f = 0; f < 6; ++f) {
    const int nk = f == 2 ? nz + 1 : nz;
    for (int j = 0; j < jext; ++j)
      for (int i = 0; i < iext; ++i)
        for (int k = 0; k < nk; ++k) {
          const double *src = c[f] + (size_t)k * cxy;
          double acc = 0.0, wt = w2;
          if (f == 0) {
            for (int jj = 0; jj < r; ++jj) acc += src[(size_t)(j * r + jj) * cnx + i * r];
            wt = w1;
          } else if (f == 1) {
            for (int ii = 0; ii < r; ++ii) acc += src[(size_t)(j * r) * cnx + i * r + ii];
            wt = w1;
          } else {
            for (int jj = 0; jj < r; ++jj)
              for (int ii = 0; ii < r; ++ii) acc += src[(size_t)(j * r + jj) * cnx + i * r + ii];
          }
          const size_t pc = (size_t)k * pxy + (size_t)wrap_mod(J0 + j, pny) * pnx + wrap_mod(I0 + i, pnx);
          p[f][pc] = acc * wt;
        }
  }
}
