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
 * with two choices of this build: parent indices wrap (periodic parent)
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
