/* oracle/dycore.c -- FP64 CPU oracle of the nonhydrostatic dynamical-core
 * step (WS-RK3 large step + forward-backward HEVI acoustic substeps).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library (see oracle.h).
 *
 * Structure follows the reference's time step:
 *   RungeKutta3Integrator::step      src/time/runge_kutta.cpp:77-118
 *     stages dt/3, dt/2, dt; s = s0 + dt_k R(s); n_s substeps per stage with
 *     dtau = dt_k / n_s; non-finite check per stage; sponge once per step.
 *   acoustic_substep                 src/time/runge_kutta.cpp:65-75 and
 *   AcousticSubstepper (HEVI)        src/time/acoustic_substep.cpp:38-149
 *   Equations::tendencies            src/dyn/equations.cpp:60-150
 *   Equations::acoustic_tendencies   src/dyn/equations.cpp:152-209
 *   Advection::scalar (WENO5 + LLF)  src/dyn/advection.cpp:149-187
 *   Damping::apply_sponge            src/dyn/damping.cpp:76-138
 *   energy_budget                    src/dyn/diagnostics.cpp:255-277
 * with the balanced ("spec") equations of docs/design/02_dynamical_core.md
 * and 04_time_integration.md: pressure gradient on pi' = pi - pi0, theta'
 * buoyancy, w/pi' vertically implicit (off-centred Crank-Nicolson) coupling,
 * rigid lid and kinematic lower boundary, terrain-following (Gal-Chen)
 * metric terms, WENO5 flux-form theta' and 5th-order upwind momentum
 * fluxes, periodic lateral boundaries.  The exact discrete equations are
 * written out in the comments below; the GPU kernels implement the same
 * equations independently.
 *
 * Compiled with -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants: include/vibe/common/constants.hpp:16-26 */
#define G_ 9.80665
#define RD_ 287.05
#define CP_ 1004.64
#define CV_ 717.63

struct orc_dycore {
  orc_domain d;
  int nx, ny, nz;
  size_t nxy, ncell, nwf;
  double *dzeta;  /* nz   : zeta[k+1]-zeta[k] */
  double *zc;     /* nz   : cell-centre zeta */
  double *dzw;    /* nz+1 : zc[k]-zc[k-1] for 1..nz-1 */
  double *Dc;     /* nxy  : ztop - zs */
  double *Du;     /* nxy  : ztop - zs at u-faces */
  double *Dv;     /* nxy  */
  double *sx;     /* nxy  : d zs/dx at u-faces */
  double *sy;     /* nxy  : d zs/dy at v-faces */
  double *kap;    /* ncell: (Rd/cv) pi0 / rho0 */
  double *lam;    /* ncell: (Rd/cv) pi0 / theta0 */
  /* scratch */
  double *om;     /* nwf  : Omega */
  double *mc;     /* ncell: terrain-slope flux at cell centres */
  orc_state s0, R;
  double *buf;
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

static inline int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}
static inline int clampi(int k, int lo, int hi) { return k < lo ? lo : (k > hi ? hi : k); }

#define IX(i, j, k) (((size_t)(k) * (size_t)ny + (size_t)wrap((j), ny)) * (size_t)nx + (size_t)wrap((i), nx))
#define HX(i, j) ((size_t)wrap((j), ny) * (size_t)nx + (size_t)wrap((i), nx))

orc_dycore *orc_dycore_create(const orc_domain *d) {
  orc_dycore *dc = (orc_dycore *)calloc(1, sizeof(orc_dycore));
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  dc->d = *d;
  dc->nx = nx;
  dc->ny = ny;
  dc->nz = nz;
  dc->nxy = (size_t)nx * ny;
  dc->ncell = dc->nxy * nz;
  dc->nwf = dc->nxy * (nz + 1);
  dc->dzeta = (double *)calloc(nz, sizeof(double));
  dc->zc = (double *)calloc(nz, sizeof(double));
  dc->dzw = (double *)calloc(nz + 1, sizeof(double));
  for (int k = 0; k < nz; ++k) {
    dc->dzeta[k] = d->zeta[k + 1] - d->zeta[k];
    dc->zc[k] = 0.5 * (d->zeta[k] + d->zeta[k + 1]);
  }
  for (int k = 1; k < nz; ++k) dc->dzw[k] = dc->zc[k] - dc->zc[k - 1];
  dc->Dc = (double *)calloc(dc->nxy, sizeof(double));
  dc->Du = (double *)calloc(dc->nxy, sizeof(double));
  dc->Dv = (double *)calloc(dc->nxy, sizeof(double));
  dc->sx = (double *)calloc(dc->nxy, sizeof(double));
  dc->sy = (double *)calloc(dc->nxy, sizeof(double));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double z = d->zs[HX(i, j)], zw = d->zs[HX(i - 1, j)], zsth = d->zs[HX(i, j - 1)];
      dc->Dc[HX(i, j)] = d->ztop - z;
      dc->Du[HX(i, j)] = d->ztop - 0.5 * (zw + z);
      dc->Dv[HX(i, j)] = d->ztop - 0.5 * (zsth + z);
      dc->sx[HX(i, j)] = (z - zw) / d->dx;
      dc->sy[HX(i, j)] = (z - zsth) / d->dy;
    }
  dc->kap = (double *)calloc(dc->ncell, sizeof(double));
  dc->lam = (double *)calloc(dc->ncell, sizeof(double));
  for (size_t n = 0; n < dc->ncell; ++n) {
    dc->kap[n] = (RD_ / CV_) * d->pi0[n] / d->rho0[n];
    dc->lam[n] = (RD_ / CV_) * d->pi0[n] / d->theta0[n];
  }
  dc->om = (double *)calloc(dc->nwf, sizeof(double));
  dc->mc = (double *)calloc(dc->ncell, sizeof(double));
  /* s0 and R: 5 cell-sized + 1 w-sized each */
  const size_t per = 5 * dc->ncell + dc->nwf;
  dc->buf = (double *)calloc(2 * per, sizeof(double));
  double *p = dc->buf;
  orc_state *ss[2] = {&dc->s0, &dc->R};
  for (int q = 0; q < 2; ++q) {
    ss[q]->u = p; p += dc->ncell;
    ss[q]->v = p; p += dc->ncell;
    ss[q]->w = p; p += dc->nwf;
    ss[q]->rho = p; p += dc->ncell;
    ss[q]->theta = p; p += dc->ncell;
    ss[q]->pi = p; p += dc->ncell;
  }
  return dc;
}

void orc_dycore_destroy(orc_dycore *dc) {
  if (!dc) return;
  free(dc->dzeta); free(dc->zc); free(dc->dzw);
  free(dc->Dc); free(dc->Du); free(dc->Dv); free(dc->sx); free(dc->sy);
  free(dc->kap); free(dc->lam); free(dc->om); free(dc->mc); free(dc->buf);
  free(dc);
}

/* ---- stencil primitives ------------------------------------------------ */

/* 5th-order upwind face value (Wicker & Skamarock 2002), face between q0 and
 * q1: flux = U*c6 - |U|*d5. */
static inline double c6(double qm2, double qm1, double q0, double q1, double q2, double q3) {
  return (37.0 * (q0 + q1) - 8.0 * (qm1 + q2) + (qm2 + q3)) / 60.0;
}
static inline double d5(double qm2, double qm1, double q0, double q1, double q2, double q3) {
  return (10.0 * (q1 - q0) - 5.0 * (q2 - qm1) + (q3 - qm2)) / 60.0;
}
static inline double up5(double U, double qm2, double qm1, double q0, double q1, double q2,
                         double q3) {
  return U * c6(qm2, qm1, q0, q1, q2, q3) - fabs(U) * d5(qm2, qm1, q0, q1, q2, q3);
}

/* WENO5 (c_kernels.c:64-104) split into its two halves. */
static inline double weno_left(double qm2, double qm1, double q0, double q1, double q2) {
  double l, r;
  orc_weno5(qm2, qm1, q0, q1, q2, &l, &r);
  return l;
}
static inline double weno_right(double qm2, double qm1, double q0, double q1, double q2) {
  double l, r;
  orc_weno5(qm2, qm1, q0, q1, q2, &l, &r);
  return r;
}
/* Local Lax-Friedrichs face flux with alpha = max(|U|, 1e-3), as
 * advection.cpp:151-156; qm/qp are the left/right states of this face. */
static inline double llf(double U, double qm, double qp) {
  const double a = fmax(fabs(U), 1e-3);
  return 0.5 * (U * (qm + qp) - a * (qp - qm));
}

/* ---- kinematics -------------------------------------------------------- */

/* Mc = horizontal velocity . grad(zs) at cell centres;
 * Omega(k) = w(k) - (1-zeta_k) * (Mc(k-1)+Mc(k))/2 at interior interfaces,
 * Omega = 0 at k = 0 and k = nz (no flux through ground or lid). */
static void compute_omega(orc_dycore *dc, const double *u, const double *v, const double *w,
                          double *om, double *mc) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  const double *sx = dc->sx, *sy = dc->sy, *zeta = dc->d.zeta;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        mc[IX(i, j, k)] = 0.5 * (u[IX(i, j, k)] * sx[HX(i, j)] + u[IX(i + 1, j, k)] * sx[HX(i + 1, j)]) +
                          0.5 * (v[IX(i, j, k)] * sy[HX(i, j)] + v[IX(i, j + 1, k)] * sy[HX(i, j + 1)]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double o = 0.0;
        if (k > 0 && k < nz)
          o = w[IX(i, j, k)] - (1.0 - zeta[k]) * (0.5 * (mc[IX(i, j, k - 1)] + mc[IX(i, j, k)]));
        om[IX(i, j, k)] = o;
      }
}

/* vertical pi' gradient per unit zeta at interface k (0 at ground and lid) */
static inline double gz(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  if (k <= 0 || k >= nz) return 0.0;
  return (pi[IX(i, j, k)] - pi[IX(i, j, k - 1)]) / dc->dzw[k];
}
static inline double gzc(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  return 0.5 * (gz(dc, pi, i, j, k) + gz(dc, pi, i, j, k + 1));
}
/* horizontal pressure-gradient operators along constant z (Gal-Chen):
 *   PX = d_x pi' |zeta - (1-zeta_c) * sx / Du * (d pi'/d zeta)  at u-faces */
static inline double px_op(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny;
  const double dpx = (pi[IX(i, j, k)] - pi[IX(i - 1, j, k)]) / dc->d.dx;
  const double met = (1.0 - dc->zc[k]) * dc->sx[HX(i, j)] / dc->Du[HX(i, j)];
  return dpx - met * (0.5 * (gzc(dc, pi, i - 1, j, k) + gzc(dc, pi, i, j, k)));
}
static inline double py_op(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny;
  const double dpy = (pi[IX(i, j, k)] - pi[IX(i, j - 1, k)]) / dc->d.dy;
  const double met = (1.0 - dc->zc[k]) * dc->sy[HX(i, j)] / dc->Dv[HX(i, j)];
  return dpy - met * (0.5 * (gzc(dc, pi, i, j - 1, k) + gzc(dc, pi, i, j, k)));
}

/* d theta0 / dz at interface k; interfaces 0 and nz copy their neighbour */
static inline double gam(const orc_dycore *dc, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  int kk = clampi(k, 1, nz - 1);
  const double *t0 = dc->d.theta0;
  return (t0[IX(i, j, kk)] - t0[IX(i, j, kk - 1)]) / (dc->Dc[HX(i, j)] * dc->dzw[kk]);
}

/* ---- slow tendencies --------------------------------------------------- */

void orc_dycore_slow_tendency(orc_dycore *dc, const orc_state *s, orc_state *R) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  const double dx = dc->d.dx, dy = dc->d.dy;
  const double *u = s->u, *v = s->v, *w = s->w, *rho = s->rho, *th = s->theta, *pi = s->pi;
  const double *Du = dc->Du, *Dv = dc->Dv, *Dc = dc->Dc, *dzeta = dc->dzeta, *dzw = dc->dzw;
  const double *om = dc->om;
  compute_omega(dc, u, v, w, dc->om, dc->mc);

  /* ---- u: A_u - cp theta'_u PX(pi') ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
#define U_(ii, jj, kk) u[IX(ii, jj, clampi(kk, 0, nz - 1))]
        double fxc[2], uxc[2], fyc[2], vyc[2], fzc[2], ozc[2];
        for (int q = 0; q < 2; ++q) {
          const int c = i - 1 + q; /* cell centre c: between faces c and c+1 */
          const double U = 0.5 * (Du[HX(c, j)] * u[IX(c, j, k)] + Du[HX(c + 1, j)] * u[IX(c + 1, j, k)]);
          uxc[q] = U;
          fxc[q] = up5(U, U_(c - 2, j, k), U_(c - 1, j, k), U_(c, j, k), U_(c + 1, j, k),
                       U_(c + 2, j, k), U_(c + 3, j, k));
          const int jf = j + q; /* corner between rows jf-1 and jf */
          const double V = 0.5 * (Dv[HX(i - 1, jf)] * v[IX(i - 1, jf, k)] + Dv[HX(i, jf)] * v[IX(i, jf, k)]);
          vyc[q] = V;
          fyc[q] = up5(V, U_(i, jf - 3, k), U_(i, jf - 2, k), U_(i, jf - 1, k), U_(i, jf, k),
                       U_(i, jf + 1, k), U_(i, jf + 2, k));
          const int kf = k + q; /* interface kf between levels kf-1 and kf */
          double O = 0.0, F = 0.0;
          if (kf > 0 && kf < nz) {
            O = 0.5 * (om[IX(i - 1, j, kf)] + om[IX(i, j, kf)]);
            F = up5(O, U_(i, j, kf - 3), U_(i, j, kf - 2), U_(i, j, kf - 1), U_(i, j, kf),
                    U_(i, j, kf + 1), U_(i, j, kf + 2));
          }
          ozc[q] = O;
          fzc[q] = F;
        }
        const double uc = u[IX(i, j, k)];
        const double adv = -(fxc[1] - fxc[0]) / dx - (fyc[1] - fyc[0]) / dy -
                           (fzc[1] - fzc[0]) / dzeta[k] +
                           uc * ((uxc[1] - uxc[0]) / dx + (vyc[1] - vyc[0]) / dy +
                                 (ozc[1] - ozc[0]) / dzeta[k]);
        const double thu = 0.5 * (th[IX(i - 1, j, k)] + th[IX(i, j, k)]);
        R->u[IX(i, j, k)] = adv / Du[HX(i, j)] - CP_ * thu * px_op(dc, pi, i, j, k);
#undef U_
      }

  /* ---- v ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
#define V_(ii, jj, kk) v[IX(ii, jj, clampi(kk, 0, nz - 1))]
        double fxc[2], uxc[2], fyc[2], vyc[2], fzc[2], ozc[2];
        for (int q = 0; q < 2; ++q) {
          const int ifc = i + q; /* corner between columns ifc-1 and ifc */
          const double U = 0.5 * (Du[HX(ifc, j - 1)] * u[IX(ifc, j - 1, k)] + Du[HX(ifc, j)] * u[IX(ifc, j, k)]);
          uxc[q] = U;
          fxc[q] = up5(U, V_(ifc - 3, j, k), V_(ifc - 2, j, k), V_(ifc - 1, j, k), V_(ifc, j, k),
                       V_(ifc + 1, j, k), V_(ifc + 2, j, k));
          const int c = j - 1 + q; /* cell centre row c */
          const double V = 0.5 * (Dv[HX(i, c)] * v[IX(i, c, k)] + Dv[HX(i, c + 1)] * v[IX(i, c + 1, k)]);
          vyc[q] = V;
          fyc[q] = up5(V, V_(i, c - 2, k), V_(i, c - 1, k), V_(i, c, k), V_(i, c + 1, k),
                       V_(i, c + 2, k), V_(i, c + 3, k));
          const int kf = k + q;
          double O = 0.0, F = 0.0;
          if (kf > 0 && kf < nz) {
            O = 0.5 * (om[IX(i, j - 1, kf)] + om[IX(i, j, kf)]);
            F = up5(O, V_(i, j, kf - 3), V_(i, j, kf - 2), V_(i, j, kf - 1), V_(i, j, kf),
                    V_(i, j, kf + 1), V_(i, j, kf + 2));
          }
          ozc[q] = O;
          fzc[q] = F;
        }
        const double vc = v[IX(i, j, k)];
        const double adv = -(fxc[1] - fxc[0]) / dx - (fyc[1] - fyc[0]) / dy -
                           (fzc[1] - fzc[0]) / dzeta[k] +
                           vc * ((uxc[1] - uxc[0]) / dx + (vyc[1] - vyc[0]) / dy +
                                 (ozc[1] - ozc[0]) / dzeta[k]);
        const double thv = 0.5 * (th[IX(i, j - 1, k)] + th[IX(i, j, k)]);
        R->v[IX(i, j, k)] = adv / Dv[HX(i, j)] - CP_ * thv * py_op(dc, pi, i, j, k);
#undef V_
      }

  /* ---- w (interior interfaces 1..nz-1; R_w = 0 at 0 and nz) ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (k == 0 || k == nz) {
          R->w[IX(i, j, k)] = 0.0;
          continue;
        }
#define W_(ii, jj, kk) w[IX(ii, jj, clampi(kk, 0, nz))]
        double fxc[2], uxc[2], fyc[2], vyc[2], fzc[2], ozc[2];
        for (int q = 0; q < 2; ++q) {
          const int ifc = i + q; /* u-face column ifc */
          const double U = 0.5 * (Du[HX(ifc, j)] * u[IX(ifc, j, k - 1)] + Du[HX(ifc, j)] * u[IX(ifc, j, k)]);
          uxc[q] = U;
          fxc[q] = up5(U, W_(ifc - 3, j, k), W_(ifc - 2, j, k), W_(ifc - 1, j, k), W_(ifc, j, k),
                       W_(ifc + 1, j, k), W_(ifc + 2, j, k));
          const int jf = j + q;
          const double V = 0.5 * (Dv[HX(i, jf)] * v[IX(i, jf, k - 1)] + Dv[HX(i, jf)] * v[IX(i, jf, k)]);
          vyc[q] = V;
          fyc[q] = up5(V, W_(i, jf - 3, k), W_(i, jf - 2, k), W_(i, jf - 1, k), W_(i, jf, k),
                       W_(i, jf + 1, k), W_(i, jf + 2, k));
          const int c = k - 1 + q; /* cell centre level c: between interfaces c and c+1 */
          const double O = 0.5 * (om[IX(i, j, c)] + om[IX(i, j, c + 1)]);
          ozc[q] = O;
          fzc[q] = up5(O, W_(i, j, c - 2), W_(i, j, c - 1), W_(i, j, c), W_(i, j, c + 1),
                       W_(i, j, c + 2), W_(i, j, c + 3));
        }
        const double wc = w[IX(i, j, k)];
        const double adv = -(fxc[1] - fxc[0]) / dx - (fyc[1] - fyc[0]) / dy -
                           (fzc[1] - fzc[0]) / dzw[k] +
                           wc * ((uxc[1] - uxc[0]) / dx + (vyc[1] - vyc[0]) / dy +
                                 (ozc[1] - ozc[0]) / dzw[k]);
        const double D = Dc[HX(i, j)];
        const double thw = 0.5 * (th[IX(i, j, k - 1)] + th[IX(i, j, k)]);
        const double th0w = 0.5 * (dc->d.theta0[IX(i, j, k - 1)] + dc->d.theta0[IX(i, j, k)]);
        const double pz = ((pi[IX(i, j, k)] - pi[IX(i, j, k - 1)]) / dzw[k]) / D;
        R->w[IX(i, j, k)] = adv / D - CP_ * thw * pz + G_ * thw / th0w;
#undef W_
      }

  /* ---- cell fields: rho', theta', pi' ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
#define T_(ii, jj, kk) th[IX(ii, jj, clampi(kk, 0, nz - 1))]
        const double D = Dc[HX(i, j)];
        /* mass: -(1/D) div(rho' V), second-order face averages */
        double frx[2], fry[2], frz[2], ftx[2], fty[2], ftz[2], mfx[2], mfy[2], mfz[2];
        for (int q = 0; q < 2; ++q) {
          const int f = i + q;
          const double U = u[IX(f, j, k)];
          mfx[q] = Du[HX(f, j)] * U;
          frx[q] = Du[HX(f, j)] * (0.5 * (rho[IX(f - 1, j, k)] + rho[IX(f, j, k)])) * U;
          ftx[q] = Du[HX(f, j)] * llf(U, weno_left(T_(f - 3, j, k), T_(f - 2, j, k), T_(f - 1, j, k), T_(f, j, k), T_(f + 1, j, k)),
                                      weno_right(T_(f - 2, j, k), T_(f - 1, j, k), T_(f, j, k), T_(f + 1, j, k), T_(f + 2, j, k)));
          const int g = j + q;
          const double V = v[IX(i, g, k)];
          mfy[q] = Dv[HX(i, g)] * V;
          fry[q] = Dv[HX(i, g)] * (0.5 * (rho[IX(i, g - 1, k)] + rho[IX(i, g, k)])) * V;
          fty[q] = Dv[HX(i, g)] * llf(V, weno_left(T_(i, g - 3, k), T_(i, g - 2, k), T_(i, g - 1, k), T_(i, g, k), T_(i, g + 1, k)),
                                      weno_right(T_(i, g - 2, k), T_(i, g - 1, k), T_(i, g, k), T_(i, g + 1, k), T_(i, g + 2, k)));
          const int h = k + q;
          double O = 0.0, FR = 0.0, FT = 0.0;
          if (h > 0 && h < nz) {
            O = om[IX(i, j, h)];
            FR = (0.5 * (rho[IX(i, j, h - 1)] + rho[IX(i, j, h)])) * O;
            FT = llf(O, weno_left(T_(i, j, h - 3), T_(i, j, h - 2), T_(i, j, h - 1), T_(i, j, h), T_(i, j, h + 1)),
                     weno_right(T_(i, j, h - 2), T_(i, j, h - 1), T_(i, j, h), T_(i, j, h + 1), T_(i, j, h + 2)));
          }
          mfz[q] = O;
          frz[q] = FR;
          ftz[q] = FT;
        }
        const size_t c = IX(i, j, k);
        const double rr = -((frx[1] - frx[0]) / dx + (fry[1] - fry[0]) / dy + (frz[1] - frz[0]) / dzeta[k]) / D;
        const double div = (mfx[1] - mfx[0]) / dx + (mfy[1] - mfy[0]) / dy + (mfz[1] - mfz[0]) / dzeta[k];
        const double fd = (ftx[1] - ftx[0]) / dx + (fty[1] - fty[0]) / dy + (ftz[1] - ftz[0]) / dzeta[k];
        const double rt = -(fd - th[c] * div) / D -
                          0.5 * (w[IX(i, j, k)] * gam(dc, i, j, k) + w[IX(i, j, k + 1)] * gam(dc, i, j, k + 1));
        R->rho[c] = rr;
        R->theta[c] = rt;
        R->pi[c] = dc->kap[c] * rr + dc->lam[c] * rt;
#undef T_
      }
}

/* ---- acoustic substep -------------------------------------------------- */

void orc_dycore_substep(orc_dycore *dc, orc_state *s, double dtau) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  const double dx = dc->d.dx, dy = dc->d.dy;
  const double *r0 = dc->d.rho0, *t0 = dc->d.theta0, *zeta = dc->d.zeta;
  const double an = 0.5 * (1.0 + dc->d.eps_s), ao = 0.5 * (1.0 - dc->d.eps_s);
  const double an2 = an * an;
  double *u = s->u, *v = s->v, *w = s->w, *rho = s->rho, *pi = s->pi;

  /* 1) horizontal momentum, forward with the old pi' (explicit) */
  double *un = dc->R.u, *vn = dc->R.v; /* scratch */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double th0u = 0.5 * (t0[IX(i - 1, j, k)] + t0[IX(i, j, k)]);
        const double th0v = 0.5 * (t0[IX(i, j - 1, k)] + t0[IX(i, j, k)]);
        un[IX(i, j, k)] = u[IX(i, j, k)] - dtau * (CP_ * th0u * px_op(dc, pi, i, j, k));
        vn[IX(i, j, k)] = v[IX(i, j, k)] - dtau * (CP_ * th0v * py_op(dc, pi, i, j, k));
      }
  memcpy(u, un, dc->ncell * sizeof(double));
  memcpy(v, vn, dc->ncell * sizeof(double));

  /* 2) per column: w / pi' / rho' with the vertical terms off-centred
   *    implicit, solved by Thomas on the interior interfaces 1..nz-1. */
#pragma omp parallel
  {
    double *mcol = (double *)malloc(sizeof(double) * (size_t)(6 * (nz + 1) + 4 * nz));
    double *Mw = mcol, *r0w = Mw + (nz + 1), *gk = r0w + (nz + 1), *wn = gk + (nz + 1);
    double *Fh = wn + (nz + 1), *F = Fh + (nz + 1);
    double *HD = F + (nz + 1), *ck = HD + nz, *Ph = ck + nz, *mcc = Ph + nz;
    double ta[512], tb[512], tc[512], td[512], tx[512], tw[512];
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double D = dc->Dc[HX(i, j)];
        for (int k = 0; k < nz; ++k) {
          mcc[k] = 0.5 * (u[IX(i, j, k)] * dc->sx[HX(i, j)] + u[IX(i + 1, j, k)] * dc->sx[HX(i + 1, j)]) +
                   0.5 * (v[IX(i, j, k)] * dc->sy[HX(i, j)] + v[IX(i, j + 1, k)] * dc->sy[HX(i, j + 1)]);
          const double r0e = 0.5 * (r0[IX(i, j, k)] + r0[IX(i + 1, j, k)]);
          const double r0wst = 0.5 * (r0[IX(i - 1, j, k)] + r0[IX(i, j, k)]);
          const double r0n = 0.5 * (r0[IX(i, j, k)] + r0[IX(i, j + 1, k)]);
          const double r0s = 0.5 * (r0[IX(i, j - 1, k)] + r0[IX(i, j, k)]);
          HD[k] = (dc->Du[HX(i + 1, j)] * r0e * u[IX(i + 1, j, k)] - dc->Du[HX(i, j)] * r0wst * u[IX(i, j, k)]) / dx +
                  (dc->Dv[HX(i, j + 1)] * r0n * v[IX(i, j + 1, k)] - dc->Dv[HX(i, j)] * r0s * v[IX(i, j, k)]) / dy;
          ck[k] = dtau * dc->kap[IX(i, j, k)] / (D * dc->dzeta[k]);
        }
        for (int k = 0; k <= nz; ++k) {
          Mw[k] = r0w[k] = gk[k] = Fh[k] = 0.0;
          if (k > 0 && k < nz) {
            Mw[k] = (1.0 - zeta[k]) * (0.5 * (mcc[k - 1] + mcc[k]));
            r0w[k] = 0.5 * (r0[IX(i, j, k - 1)] + r0[IX(i, j, k)]);
            const double th0w = 0.5 * (t0[IX(i, j, k - 1)] + t0[IX(i, j, k)]);
            gk[k] = dtau * CP_ * th0w / (D * dc->dzw[k]);
            Fh[k] = r0w[k] * (ao * w[IX(i, j, k)] - Mw[k]);
          }
        }
        for (int k = 0; k < nz; ++k)
          Ph[k] = pi[IX(i, j, k)] - dc->kap[IX(i, j, k)] * (dtau / D) * HD[k] - ck[k] * (Fh[k + 1] - Fh[k]);
        const int n = nz - 1;
        for (int m = 0; m < n; ++m) {
          const int k = m + 1;
          const double ga = gk[k] * an2;
          if (m > 0) ta[m - 1] = -ga * (ck[k - 1] * r0w[k - 1]);
          tb[m] = 1.0 + ga * ((ck[k] + ck[k - 1]) * r0w[k]);
          if (m < n - 1) tc[m] = -ga * (ck[k] * r0w[k + 1]);
          td[m] = w[IX(i, j, k)] - gk[k] * ao * (pi[IX(i, j, k)] - pi[IX(i, j, k - 1)]) -
                  gk[k] * an * (Ph[k] - Ph[k - 1]);
        }
        if (n > 0) orc_tridiagonal_solve(ta, tb, tc, td, n, tx, tw);
        wn[0] = mcc[0];
        wn[nz] = w[IX(i, j, nz)];
        for (int m = 0; m < n; ++m) wn[m + 1] = tx[m];
        for (int k = 0; k <= nz; ++k) {
          F[k] = 0.0;
          if (k > 0 && k < nz) F[k] = r0w[k] * (an * wn[k] + ao * w[IX(i, j, k)] - Mw[k]);
        }
        for (int k = 0; k < nz; ++k) {
          const double drho = -(dtau / D) * (HD[k] + (F[k + 1] - F[k]) / dc->dzeta[k]);
          rho[IX(i, j, k)] += drho;
          pi[IX(i, j, k)] += dc->kap[IX(i, j, k)] * drho;
        }
        for (int k = 0; k <= nz; ++k) w[IX(i, j, k)] = wn[k];
      }
    free(mcol);
  }
}

/* ---- sponge (damping.cpp:76-138, perturbation targets are 0) ----------- */

void orc_dycore_sponge(orc_dycore *dc, orc_state *s, double dt) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  const double *sp = dc->d.sponge;
  for (int k = 0; k <= nz; ++k) {
    const double a = sp[clampi(k, 0, nz - 1)];
    if (a <= 0.0) continue;
    const double f = 1.0 / (1.0 + dt * a);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        s->w[IX(i, j, k)] = f * s->w[IX(i, j, k)];
        if (k < nz) {
          const size_t c = IX(i, j, k);
          s->u[c] = f * s->u[c];
          s->v[c] = f * s->v[c];
          s->rho[c] = f * s->rho[c];
          s->theta[c] = f * s->theta[c];
          s->pi[c] = f * s->pi[c];
        }
      }
  }
}

/* ---- RK3 step ---------------------------------------------------------- */

static int nonfinite(const orc_dycore *dc, const orc_state *s) {
  int bad = 0;
  const size_t nc = dc->ncell, nw = dc->nwf;
  for (size_t n = 0; n < nc && !bad; ++n)
    bad = !isfinite(s->u[n]) || !isfinite(s->v[n]) || !isfinite(s->rho[n]) ||
          !isfinite(s->theta[n]) || !isfinite(s->pi[n]);
  for (size_t n = 0; n < nw && !bad; ++n) bad = !isfinite(s->w[n]);
  return bad;
}

static void copy_state(const orc_dycore *dc, orc_state *dst, const orc_state *src) {
  const size_t nc = dc->ncell * sizeof(double);
  memcpy(dst->u, src->u, nc);
  memcpy(dst->v, src->v, nc);
  memcpy(dst->w, src->w, dc->nwf * sizeof(double));
  memcpy(dst->rho, src->rho, nc);
  memcpy(dst->theta, src->theta, nc);
  memcpy(dst->pi, src->pi, nc);
}

int orc_dycore_step(orc_dycore *dc, orc_state *s, double dt, int nsub) {
#ifdef _OPENMP
  if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
  const int nz = dc->nz;
  const size_t nxy = dc->nxy;
  const int ns = nsub < 1 ? 1 : nsub;
  const double dtk[3] = {dt / 3.0, dt / 2.0, dt};
  copy_state(dc, &dc->s0, s);
  for (int st = 0; st < 3; ++st) {
    orc_dycore_slow_tendency(dc, s, &dc->R);
    const double h = dtk[st];
    const orc_state *s0 = &dc->s0, *R = &dc->R;
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < dc->ncell; ++n) {
      s->u[n] = s0->u[n] + h * R->u[n];
      s->v[n] = s0->v[n] + h * R->v[n];
      s->rho[n] = s0->rho[n] + h * R->rho[n];
      s->theta[n] = s0->theta[n] + h * R->theta[n];
      s->pi[n] = s0->pi[n] + h * R->pi[n];
    }
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < dc->nwf; ++n) {
      const size_t k = n / nxy;
      s->w[n] = (k == 0 || k == (size_t)nz) ? s0->w[n] : s0->w[n] + h * R->w[n];
    }
    const double dtau = h / (double)ns;
    for (int q = 0; q < ns; ++q) orc_dycore_substep(dc, s, dtau);
    if (nonfinite(dc, s)) return 102;
  }
  orc_dycore_sponge(dc, s, dt);
  return 0;
}

/* ---- compensated totals (diagnostics.cpp:255-277, Neumaier-summed) ----- */

typedef struct {
  double s, c;
} nsum;
static inline void nadd(nsum *a, double x) {
  const double t = a->s + x;
  if (fabs(a->s) >= fabs(x))
    a->c += (a->s - t) + x;
  else
    a->c += (x - t) + a->s;
  a->s = t;
}

void orc_dycore_energy(orc_dycore *dc, const orc_state *s, double out[6]) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  const double dx = dc->d.dx, dy = dc->d.dy, H = dc->d.ztop;
  nsum m = {0, 0}, mp = {0, 0}, ke = {0, 0}, ie = {0, 0}, pe = {0, 0};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t c = IX(i, j, k);
        const double zs = dc->d.zs[HX(i, j)];
        const double dv = dx * dy * ((H - zs) / H) * dc->dzeta[k] * H;
        const double r = dc->d.rho0[c] + s->rho[c];
        const double uc = 0.5 * (s->u[c] + s->u[IX(i + 1, j, k)]);
        const double vc = 0.5 * (s->v[c] + s->v[IX(i, j + 1, k)]);
        const double wc = 0.5 * (s->w[c] + s->w[IX(i, j, k + 1)]);
        const double T = (dc->d.theta0[c] + s->theta[c]) * (dc->d.pi0[c] + s->pi[c]);
        const double z = zs + dc->zc[k] * (H - zs);
        nadd(&m, r * dv);
        nadd(&mp, s->rho[c] * dv);
        nadd(&ke, 0.5 * r * (uc * uc + vc * vc + wc * wc) * dv);
        nadd(&ie, r * CV_ * T * dv);
        nadd(&pe, r * G_ * z * dv);
      }
  out[0] = m.s + m.c;
  out[1] = mp.s + mp.c;
  out[2] = ke.s + ke.c;
  out[3] = ie.s + ie.c;
  out[4] = pe.s + pe.c;
  out[5] = out[2] + out[3] + out[4];
}
