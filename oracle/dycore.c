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
  double *rDc, *rDu, *rDv; /* nxy: reciprocal column depths */
  double *metx, *mety;     /* nxy: sx / Du, sy / Dv */
  double *rdzeta, *rdzw;   /* 1/dzeta (nz), 1/dzw (nz+1, 1..nz-1) */
  double rdx, rdy;
  /* physics coupling (runge_kutta.cpp:46-63): rates added to R each stage */
  double *ph_th, *ph_u, *ph_v;
  int ph_on;
  orc_physics_fn ph_fn;
  void *ph_user;
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

/* planted-mistake factors (oracle.h): exactly 1 (and pl_zeta 0) unless a
 * test plants a slip */
static double pl_buoy = 1.0, pl_gam = 1.0, pl_met = 1.0, pl_om = 1.0;
static int pl_zeta = 0;
void orc_plant(int which) {
  pl_buoy = which == 1 ? -1.0 : 1.0;
  pl_gam = which == 2 ? 0.5 : 1.0;
  pl_met = which == 3 ? -1.0 : 1.0;
  pl_om = which == 4 ? -1.0 : 1.0;
  pl_zeta = which == 5;
}

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
  dc->rdzeta = (double *)calloc(nz, sizeof(double));
  dc->rdzw = (double *)calloc(nz + 1, sizeof(double));
  for (int k = 0; k < nz; ++k) dc->rdzeta[k] = 1.0 / dc->dzeta[k];
  for (int k = 1; k < nz; ++k) dc->rdzw[k] = 1.0 / dc->dzw[k];
  dc->rdx = 1.0 / d->dx;
  dc->rdy = 1.0 / d->dy;
  dc->Dc = (double *)calloc(dc->nxy, sizeof(double));
  dc->Du = (double *)calloc(dc->nxy, sizeof(double));
  dc->Dv = (double *)calloc(dc->nxy, sizeof(double));
  dc->sx = (double *)calloc(dc->nxy, sizeof(double));
  dc->sy = (double *)calloc(dc->nxy, sizeof(double));
  dc->rDc = (double *)calloc(dc->nxy, sizeof(double));
  dc->rDu = (double *)calloc(dc->nxy, sizeof(double));
  dc->rDv = (double *)calloc(dc->nxy, sizeof(double));
  dc->metx = (double *)calloc(dc->nxy, sizeof(double));
  dc->mety = (double *)calloc(dc->nxy, sizeof(double));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double z = d->zs[HX(i, j)], zw = d->zs[HX(i - 1, j)], zsth = d->zs[HX(i, j - 1)];
      dc->Dc[HX(i, j)] = d->ztop - z;
      dc->Du[HX(i, j)] = d->ztop - 0.5 * (zw + z);
      dc->Dv[HX(i, j)] = d->ztop - 0.5 * (zsth + z);
      dc->sx[HX(i, j)] = (z - zw) / d->dx;
      dc->sy[HX(i, j)] = (z - zsth) / d->dy;
      dc->rDc[HX(i, j)] = 1.0 / dc->Dc[HX(i, j)];
      dc->rDu[HX(i, j)] = 1.0 / dc->Du[HX(i, j)];
      dc->rDv[HX(i, j)] = 1.0 / dc->Dv[HX(i, j)];
      dc->metx[HX(i, j)] = dc->sx[HX(i, j)] * dc->rDu[HX(i, j)];
      dc->mety[HX(i, j)] = dc->sy[HX(i, j)] * dc->rDv[HX(i, j)];
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
  free(dc->rDc); free(dc->rDu); free(dc->rDv); free(dc->metx); free(dc->mety);
  free(dc->rdzeta); free(dc->rdzw);
  free(dc->kap); free(dc->lam); free(dc->om); free(dc->mc); free(dc->buf);
  free(dc->ph_th); free(dc->ph_u); free(dc->ph_v);
  free(dc);
}

static void ensure_physics(orc_dycore *dc) {
  if (!dc->ph_th) {
    dc->ph_th = (double *)calloc(dc->ncell, sizeof(double));
    dc->ph_u = (double *)calloc(dc->ncell, sizeof(double));
    dc->ph_v = (double *)calloc(dc->ncell, sizeof(double));
  }
}

void orc_dycore_set_physics(orc_dycore *dc, const double *theta_rate, const double *u_rate, const double *v_rate) {
  if (!theta_rate && !u_rate && !v_rate) {
    dc->ph_on = 0;
    return;
  }
  ensure_physics(dc);
  const double *src[3] = {theta_rate, u_rate, v_rate};
  double *dst[3] = {dc->ph_th, dc->ph_u, dc->ph_v};
  for (int q = 0; q < 3; ++q) {
    if (src[q])
      memcpy(dst[q], src[q], dc->ncell * sizeof(double));
    else
      memset(dst[q], 0, dc->ncell * sizeof(double));
  }
  dc->ph_on = 1;
}

void orc_dycore_set_physics_callback(orc_dycore *dc, orc_physics_fn fn, void *user) {
  dc->ph_fn = fn;
  dc->ph_user = user;
  dc->ph_on = fn != NULL;
  if (fn) ensure_physics(dc);
}

/* ---- stencil primitives ------------------------------------------------ */

/* 5th-order upwind face value (Wicker & Skamarock 2002), face between q0 and
 * q1: flux = U*c6 - |U|*d5. */
static inline double c6(double qm2, double qm1, double q0, double q1, double q2, double q3) {
  return (37.0 * (q0 + q1) - 8.0 * (qm1 + q2) + (qm2 + q3)) * (1.0 / 60.0);
}
static inline double d5(double qm2, double qm1, double q0, double q1, double q2, double q3) {
  return (10.0 * (q1 - q0) - 5.0 * (q2 - qm1) + (q3 - qm2)) * (1.0 / 60.0);
}
static inline double up5(double U, double qm2, double qm1, double q0, double q1, double q2,
                         double q3) {
  return U * c6(qm2, qm1, q0, q1, q2, q3) - fabs(U) * d5(qm2, qm1, q0, q1, q2, q3);
}

/* WENO5 value at the face right of c from the stencil (a, b, c, d, e) with
 * the weights of c_kernels.c:64-104 (d = 0.1, 0.6, 0.3; eps = 1e-6), written
 * with a single division: multiplying the weights alpha_k = d_k/(eps+beta_k)^2
 * by prod_j (eps+beta_j)^2 leaves w_k = d_k prod_{j!=k} (eps+beta_j)^2. */
static inline double weno_face(double a, double b, double c, double d, double e) {
  const double eps = 1e-6;
  const double p0 = 2.0 * a - 7.0 * b + 11.0 * c;
  const double p1 = -b + 5.0 * c + 2.0 * d;
  const double p2 = 2.0 * c + 5.0 * d - e;
  const double s0a = a - 2.0 * b + c, s0b = a - 4.0 * b + 3.0 * c;
  const double s1a = b - 2.0 * c + d, s1b = b - d;
  const double s2a = c - 2.0 * d + e, s2b = 3.0 * c - 4.0 * d + e;
  const double e0 = eps + ((13.0 / 12.0) * s0a * s0a + 0.25 * s0b * s0b);
  const double e1 = eps + ((13.0 / 12.0) * s1a * s1a + 0.25 * s1b * s1b);
  const double e2 = eps + ((13.0 / 12.0) * s2a * s2a + 0.25 * s2b * s2b);
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double w0 = 0.1 * (q1 * q2), w1 = 0.6 * (q0 * q2), w2 = 0.3 * (q0 * q1);
  return (w0 * p0 + w1 * p1 + w2 * p2) / (6.0 * (w0 + w1 + w2));
}
/* Upwind WENO5 flux through the face between cells f-1 and f, given the six
 * cells q[f-3..f+2]: U >= 0 takes the left-biased value (stencil centred on
 * f-1), U < 0 the right-biased one (stencil centred on f, mirrored). */
static inline double upwind_weno(double U, double qm3, double qm2, double qm1, double q0, double q1,
                                 double q2) {
  const double qf = (U >= 0.0) ? weno_face(qm3, qm2, qm1, q0, q1) : weno_face(q2, q1, q0, qm1, qm2);
  return U * qf;
}

/* Exported views of the two face primitives for the oracle's own tests. */
double orc_weno_face(double a, double b, double c, double d, double e) { return weno_face(a, b, c, d, e); }
double orc_upwind_weno(double U, double qm3, double qm2, double qm1, double q0, double q1, double q2) {
  return upwind_weno(U, qm3, qm2, qm1, q0, q1, q2);
}
double orc_up5(double U, double qm2, double qm1, double q0, double q1, double q2, double q3) {
  return up5(U, qm2, qm1, q0, q1, q2, q3);
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
          o = w[IX(i, j, k)] -
              ((pl_zeta ? zeta[k] : 1.0 - zeta[k]) * pl_om) * (0.5 * (mc[IX(i, j, k - 1)] + mc[IX(i, j, k)]));
        om[IX(i, j, k)] = o;
      }
}

/* vertical pi' gradient per unit zeta at interface k (0 at ground and lid) */
static inline double gz(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  if (k <= 0 || k >= nz) return 0.0;
  return (pi[IX(i, j, k)] - pi[IX(i, j, k - 1)]) * dc->rdzw[k];
}
static inline double gzc(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  return 0.5 * (gz(dc, pi, i, j, k) + gz(dc, pi, i, j, k + 1));
}
/* horizontal pressure-gradient operators along constant z (Gal-Chen):
 *   PX = d_x pi' |zeta - (1-zeta_c) * sx / Du * (d pi'/d zeta)  at u-faces */
static inline double px_op(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny;
  const double dpx = (pi[IX(i, j, k)] - pi[IX(i - 1, j, k)]) * dc->rdx;
  const double met = ((1.0 - dc->zc[k]) * pl_met) * dc->metx[HX(i, j)];
  return dpx - met * (0.5 * (gzc(dc, pi, i - 1, j, k) + gzc(dc, pi, i, j, k)));
}
static inline double py_op(const orc_dycore *dc, const double *pi, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny;
  const double dpy = (pi[IX(i, j, k)] - pi[IX(i, j - 1, k)]) * dc->rdy;
  const double met = ((1.0 - dc->zc[k]) * pl_met) * dc->mety[HX(i, j)];
  return dpy - met * (0.5 * (gzc(dc, pi, i, j - 1, k) + gzc(dc, pi, i, j, k)));
}

/* d theta0 / dz at interface k; interfaces 0 and nz copy their neighbour */
static inline double gam(const orc_dycore *dc, int i, int j, int k) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  int kk = clampi(k, 1, nz - 1);
  const double *t0 = dc->d.theta0;
  return (t0[IX(i, j, kk)] - t0[IX(i, j, kk - 1)]) * dc->rDc[HX(i, j)] * dc->rdzw[kk] * pl_gam;
}

/* ---- slow tendencies --------------------------------------------------- */

void orc_dycore_slow_tendency(orc_dycore *dc, const orc_state *s, orc_state *R) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
  const double rdx = dc->rdx, rdy = dc->rdy;
  const double *rdzeta = dc->rdzeta, *rdzw = dc->rdzw, *rDc = dc->rDc, *rDu = dc->rDu, *rDv = dc->rDv;
  const double *u = s->u, *v = s->v, *w = s->w, *rho = s->rho, *th = s->theta, *pi = s->pi;
  const double *Du = dc->Du, *Dv = dc->Dv;
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
        const double adv = -(fxc[1] - fxc[0]) * rdx - (fyc[1] - fyc[0]) * rdy -
                           (fzc[1] - fzc[0]) * rdzeta[k] +
                           uc * ((uxc[1] - uxc[0]) * rdx + (vyc[1] - vyc[0]) * rdy +
                                 (ozc[1] - ozc[0]) * rdzeta[k]);
        const double thu = 0.5 * (th[IX(i - 1, j, k)] + th[IX(i, j, k)]);
        double ru = adv * rDu[HX(i, j)] - CP_ * thu * px_op(dc, pi, i, j, k);
        if (dc->ph_on) ru += dc->ph_u[IX(i, j, k)]; /* physics rate (runge_kutta.cpp:60) */
        R->u[IX(i, j, k)] = ru;
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
        const double adv = -(fxc[1] - fxc[0]) * rdx - (fyc[1] - fyc[0]) * rdy -
                           (fzc[1] - fzc[0]) * rdzeta[k] +
                           vc * ((uxc[1] - uxc[0]) * rdx + (vyc[1] - vyc[0]) * rdy +
                                 (ozc[1] - ozc[0]) * rdzeta[k]);
        const double thv = 0.5 * (th[IX(i, j - 1, k)] + th[IX(i, j, k)]);
        double rv = adv * rDv[HX(i, j)] - CP_ * thv * py_op(dc, pi, i, j, k);
        if (dc->ph_on) rv += dc->ph_v[IX(i, j, k)]; /* runge_kutta.cpp:61 */
        R->v[IX(i, j, k)] = rv;
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
        const double adv = -(fxc[1] - fxc[0]) * rdx - (fyc[1] - fyc[0]) * rdy -
                           (fzc[1] - fzc[0]) * rdzw[k] +
                           wc * ((uxc[1] - uxc[0]) * rdx + (vyc[1] - vyc[0]) * rdy +
                                 (ozc[1] - ozc[0]) * rdzw[k]);
        const double rD = rDc[HX(i, j)];
        const double thw = 0.5 * (th[IX(i, j, k - 1)] + th[IX(i, j, k)]);
        const double th0w = 0.5 * (dc->d.theta0[IX(i, j, k - 1)] + dc->d.theta0[IX(i, j, k)]);
        const double pz = (pi[IX(i, j, k)] - pi[IX(i, j, k - 1)]) * rdzw[k] * rD;
        R->w[IX(i, j, k)] = adv * rD - CP_ * thw * pz + (pl_buoy * G_) * thw / th0w;
#undef W_
      }

  /* ---- cell fields: rho', theta', pi' ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
#define T_(ii, jj, kk) th[IX(ii, jj, clampi(kk, 0, nz - 1))]
        const double rD = rDc[HX(i, j)];
        /* mass: -(1/D) div(rho' V), second-order face averages */
        double frx[2], fry[2], frz[2], ftx[2], fty[2], ftz[2], mfx[2], mfy[2], mfz[2];
        for (int q = 0; q < 2; ++q) {
          const int f = i + q;
          const double U = u[IX(f, j, k)];
          mfx[q] = Du[HX(f, j)] * U;
          frx[q] = Du[HX(f, j)] * (0.5 * (rho[IX(f - 1, j, k)] + rho[IX(f, j, k)])) * U;
          ftx[q] = Du[HX(f, j)] * upwind_weno(U, T_(f - 3, j, k), T_(f - 2, j, k), T_(f - 1, j, k), T_(f, j, k),
                                              T_(f + 1, j, k), T_(f + 2, j, k));
          const int g = j + q;
          const double V = v[IX(i, g, k)];
          mfy[q] = Dv[HX(i, g)] * V;
          fry[q] = Dv[HX(i, g)] * (0.5 * (rho[IX(i, g - 1, k)] + rho[IX(i, g, k)])) * V;
          fty[q] = Dv[HX(i, g)] * upwind_weno(V, T_(i, g - 3, k), T_(i, g - 2, k), T_(i, g - 1, k), T_(i, g, k),
                                              T_(i, g + 1, k), T_(i, g + 2, k));
          const int h = k + q;
          double O = 0.0, FR = 0.0, FT = 0.0;
          if (h > 0 && h < nz) {
            O = om[IX(i, j, h)];
            FR = (0.5 * (rho[IX(i, j, h - 1)] + rho[IX(i, j, h)])) * O;
            FT = upwind_weno(O, T_(i, j, h - 3), T_(i, j, h - 2), T_(i, j, h - 1), T_(i, j, h), T_(i, j, h + 1),
                             T_(i, j, h + 2));
          }
          mfz[q] = O;
          frz[q] = FR;
          ftz[q] = FT;
        }
        const size_t c = IX(i, j, k);
        const double rr = -((frx[1] - frx[0]) * rdx + (fry[1] - fry[0]) * rdy + (frz[1] - frz[0]) * rdzeta[k]) * rD;
        const double div = (mfx[1] - mfx[0]) * rdx + (mfy[1] - mfy[0]) * rdy + (mfz[1] - mfz[0]) * rdzeta[k];
        const double fd = (ftx[1] - ftx[0]) * rdx + (fty[1] - fty[0]) * rdy + (ftz[1] - ftz[0]) * rdzeta[k];
        double rt = -(fd - th[c] * div) * rD -
                    0.5 * (w[IX(i, j, k)] * gam(dc, i, j, k) + w[IX(i, j, k + 1)] * gam(dc, i, j, k + 1));
        /* heating (runge_kutta.cpp:53) enters before R_pi' is formed */
        if (dc->ph_on) rt += dc->ph_th[c];
        R->rho[c] = rr;
        R->theta[c] = rt;
        R->pi[c] = dc->kap[c] * rr + dc->lam[c] * rt;
#undef T_
      }
}

/* ---- acoustic substep -------------------------------------------------- */

void orc_dycore_substep(orc_dycore *dc, orc_state *s, double dtau) {
  const int nx = dc->nx, ny = dc->ny, nz = dc->nz;
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
   *    implicit.  Explicit parts E (rho'), Q (pi'), Fh (vertical flux with
   *    the old w), then a tridiagonal system for the new w on interfaces
   *    1..nz-1 solved by a factored Thomas sweep (reciprocal pivots), then
   *    the vertical flux divergence closes rho' and pi'. */
#pragma omp parallel
  {
    double *col = (double *)malloc(sizeof(double) * (size_t)(12 * (nz + 1)));
    double *mcc = col, *HD = mcc + (nz + 1), *E = HD + (nz + 1), *Q = E + (nz + 1);
    double *ck = Q + (nz + 1), *Mw = ck + (nz + 1), *r0w = Mw + (nz + 1), *Fh = r0w + (nz + 1);
    double *Ph = Fh + (nz + 1), *gk = Ph + (nz + 1), *wn = gk + (nz + 1), *F = wn + (nz + 1);
    double ta[512], tb[512], tc[512], td[512], tinv[512], tcp[512];
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double dtD = dtau * dc->rDc[HX(i, j)];
        for (int k = 0; k < nz; ++k) {
          mcc[k] = 0.5 * (u[IX(i, j, k)] * dc->sx[HX(i, j)] + u[IX(i + 1, j, k)] * dc->sx[HX(i + 1, j)]) +
                   0.5 * (v[IX(i, j, k)] * dc->sy[HX(i, j)] + v[IX(i, j + 1, k)] * dc->sy[HX(i, j + 1)]);
          const double r0e = 0.5 * (r0[IX(i, j, k)] + r0[IX(i + 1, j, k)]);
          const double r0wst = 0.5 * (r0[IX(i - 1, j, k)] + r0[IX(i, j, k)]);
          const double r0n = 0.5 * (r0[IX(i, j, k)] + r0[IX(i, j + 1, k)]);
          const double r0s = 0.5 * (r0[IX(i, j - 1, k)] + r0[IX(i, j, k)]);
          HD[k] = (dc->Du[HX(i + 1, j)] * r0e * u[IX(i + 1, j, k)] - dc->Du[HX(i, j)] * r0wst * u[IX(i, j, k)]) * dc->rdx +
                  (dc->Dv[HX(i, j + 1)] * r0n * v[IX(i, j + 1, k)] - dc->Dv[HX(i, j)] * r0s * v[IX(i, j, k)]) * dc->rdy;
          const double cq = dc->kap[IX(i, j, k)] * dtD;
          ck[k] = cq * dc->rdzeta[k];
          E[k] = rho[IX(i, j, k)] - dtD * HD[k];
          Q[k] = pi[IX(i, j, k)] - cq * HD[k];
        }
        for (int k = 0; k <= nz; ++k) {
          Mw[k] = r0w[k] = gk[k] = Fh[k] = 0.0;
          if (k > 0 && k < nz) {
            Mw[k] = ((pl_zeta ? zeta[k] : 1.0 - zeta[k]) * pl_om) * (0.5 * (mcc[k - 1] + mcc[k]));
            r0w[k] = 0.5 * (r0[IX(i, j, k - 1)] + r0[IX(i, j, k)]);
            const double th0w = 0.5 * (t0[IX(i, j, k - 1)] + t0[IX(i, j, k)]);
            gk[k] = dtD * CP_ * th0w * dc->rdzw[k];
            Fh[k] = r0w[k] * (ao * w[IX(i, j, k)] - Mw[k]);
          }
        }
        for (int k = 0; k < nz; ++k) Ph[k] = Q[k] - ck[k] * (Fh[k + 1] - Fh[k]);
        const int n = nz - 1;
        for (int m = 0; m < n; ++m) {
          const int k = m + 1;
          const double ga = gk[k] * an2;
          ta[m] = (m > 0) ? -ga * (ck[k - 1] * r0w[k - 1]) : 0.0;
          tb[m] = 1.0 + ga * ((ck[k] + ck[k - 1]) * r0w[k]);
          tc[m] = (m < n - 1) ? -ga * (ck[k] * r0w[k + 1]) : 0.0;
          td[m] = w[IX(i, j, k)] - gk[k] * ao * (pi[IX(i, j, k)] - pi[IX(i, j, k - 1)]) -
                  gk[k] * an * (Ph[k] - Ph[k - 1]);
        }
        /* factored Thomas (helmholtz.cpp:88-125 form) with reciprocal
         * pivots inv_m, scaled sub-diagonal beta_m = a_m inv_m and c'_m:
         *   d'_m = td_m inv_m - beta_m d'_{m-1},  x_m = d'_m - c'_m x_{m+1} */
        for (int m = 0; m < n; ++m) {
          const double den = (m == 0) ? tb[0] : tb[m] - ta[m] * tcp[m - 1];
          tinv[m] = 1.0 / den;
          tcp[m] = tc[m] * tinv[m];
          ta[m] = ta[m] * tinv[m];
        }
        for (int m = 0; m < n; ++m) td[m] = (m == 0) ? td[0] * tinv[0] : td[m] * tinv[m] - ta[m] * td[m - 1];
        for (int m = n - 2; m >= 0; --m) td[m] = td[m] - tcp[m] * td[m + 1];
        wn[0] = mcc[0];
        wn[nz] = w[IX(i, j, nz)];
        for (int m = 0; m < n; ++m) wn[m + 1] = td[m];
        for (int k = 0; k <= nz; ++k) F[k] = (k > 0 && k < nz) ? Fh[k] + (an * r0w[k]) * wn[k] : 0.0;
        for (int k = 0; k < nz; ++k) {
          const double vd = -dtD * ((F[k + 1] - F[k]) * dc->rdzeta[k]);
          rho[IX(i, j, k)] = E[k] + vd;
          pi[IX(i, j, k)] = Q[k] + dc->kap[IX(i, j, k)] * vd;
        }
        for (int k = 0; k <= nz; ++k) w[IX(i, j, k)] = wn[k];
      }
    free(col);
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
    if (dc->ph_fn) { /* physics on the stage state, fresh zero rates (runge_kutta.cpp:48-51) */
      memset(dc->ph_th, 0, dc->ncell * sizeof(double));
      memset(dc->ph_u, 0, dc->ncell * sizeof(double));
      memset(dc->ph_v, 0, dc->ncell * sizeof(double));
      if (dc->ph_fn(dc->ph_user, st, dt, s, dc->ph_th, dc->ph_u, dc->ph_v)) return 103;
    }
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
