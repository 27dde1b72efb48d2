/* oracle/oracle.h -- FP64 CPU oracle of the dynamical-core step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Nothing in the
 * product (vibe-model_amd/) includes, links or calls this code.
 *
 * Field layout (all arrays dense, x fastest):
 *   cell / u-face / v-face fields: [nz][ny][nx]
 *   w-face fields:                 [nz+1][ny][nx]
 * u(i,j,k) is the x-face on the west side of cell i, v(i,j,k) the y-face on
 * the south side of cell j, w(i,j,k) the interface below cell k.  x and y are
 * periodic.  The prognostic state is (u, v, w, rho', theta', pi'), the
 * primed fields being perturbations from the reference state (rho0, theta0,
 * pi0) stored at cell centres.
 */
#ifndef VIBE_ORACLE_H
#define VIBE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- C11 kernel restatements (ckernels.c) ------------------------------ */
int orc_tridiagonal_solve(const double *a, const double *b, const double *c,
                          const double *d, int n, double *x, double *ws);
void orc_weno5(double qm2, double qm1, double q0, double q1, double q2,
               double *q_left, double *q_right);
double orc_bilinear(const double *f, int halo, int nsx, int nsy, int nz,
                    double xi, double yj, int k);
double orc_cfl_scan(const double *u, const double *v, const double *w,
                    const double *dzeta, const double *zs, int nx, int ny, int nz,
                    double dx, double dy, double z_top, double dt,
                    double *max_wind_out, double *min_dz_out);
double orc_neumaier_sum(const double *x, size_t n);
void orc_advect_scalar_flux(const double *q, const double *rho, const double *u,
                            const double *v, const double *w, const double *dzeta,
                            const double *zs, int nx, int ny, int nz, double dx,
                            double dy, double z_top, double *dqdt);

/* ---- the reference's literal equations (literal.c) --------------------
 * State and tendency arrays s[12], t[12] in the species order of the
 * reference's State (u, v, w, rho, theta, pi, qv, qc, qr, qi, qs, qg), FULL
 * values, dense, i fastest: u [nz][ny][nx+1] (faces 0..nx), v [nz][ny+1][nx]
 * (faces 0..ny), w [nz+1][ny][nx] (faces 0..nz), the rest [nz][ny][nx]. */
typedef struct {
  int nx, ny, nz;
  double dx, dy, z_top;
  const double *zeta;   /* nz+1 interfaces */
  const double *zs;     /* ny*nx terrain, NULL: flat */
  const double *rho0, *theta0, *pi0; /* reference state, cells */
  int order;            /* advection: 2, 4 or 6 (central2/4/6) */
  int coriolis;         /* 0 none, 1 f-plane, 2 beta plane */
  double latitude;      /* degrees */
  double div_coeff;     /* divergence damping coefficient (0: off) */
  int div_order;        /* 2: Del^2, 4: Del^4 */
  int moist;            /* advect qc..qg (microphysics != "none") */
} orc_literal_desc;

int orc_literal_tendencies(const orc_literal_desc *d, const double *const s[12], double *const t[12]);
int orc_literal_acoustic(const orc_literal_desc *d, const double *const s[12], double *const t[12]);
double orc_literal_max_sound_speed(const orc_literal_desc *d, const double *const s[12]);
void orc_literal_diagnose(const orc_literal_desc *d, const double *const s[12], double *pi_out);
/* one literal RK3 step (runge_kutta.cpp:77-118) with explicit
 * forward-backward substeps and the sponge rates sponge[nz] (NULL: none);
 * 0, 3 if non-finite, 2 bad arguments */
int orc_literal_step(const orc_literal_desc *d, double *const s[12], double dt, int n_sub, const double *sponge);
/* SemiImplicitIntegrator (semi_implicit.cpp): assemble_rhs of state s and
 * tendency n into a[ncell], rhs[ncell]; the state update of one step
 * (tendencies + physics rates {theta, qv, u, v}, forward Euler of the
 * scalars, time filter, sponge); 0, 3 if non-finite, 2 bad arguments */
void orc_si_assemble(const orc_literal_desc *d, const double *const s[12], const double *const n[12], double dt,
                     double *a, double *rhs);
int orc_si_step(const orc_literal_desc *d, double *const s[12], double dt, const double *sponge, double filter,
                int first, const double *const phys[4]);

/* ---- dynamical core (dycore.c) ---------------------------------------- */
typedef struct {
  int nx, ny, nz;
  double dx, dy, ztop;
  const double *zeta;   /* nz+1 normalised interfaces, zeta[0]=0, zeta[nz]=1 */
  const double *zs;     /* ny*nx terrain height (m) */
  const double *rho0;   /* reference state at cell centres, nz*ny*nx */
  const double *theta0;
  const double *pi0;
  const double *sponge; /* nz Rayleigh rates (1/s), 0 below the sponge */
  double eps_s;         /* HEVI off-centering, weights (1+eps)/2, (1-eps)/2 */
} orc_domain;

typedef struct {
  double *u, *v, *w, *rho, *theta, *pi;
} orc_state;

typedef struct orc_dycore orc_dycore;

orc_dycore *orc_dycore_create(const orc_domain *d);
void orc_dycore_destroy(orc_dycore *dc);
/* One WS-RK3 step with nsub acoustic substeps per stage and the sponge.
 * Returns 0, or 102 if the state became non-finite. */
int orc_dycore_step(orc_dycore *dc, orc_state *s, double dt, int nsub);
/* Slow (large-step) tendencies of state s into R (same layout). */
void orc_dycore_slow_tendency(orc_dycore *dc, const orc_state *s, orc_state *R);
/* One forward-backward HEVI substep in place on (u, v, w, rho', pi'). */
void orc_dycore_substep(orc_dycore *dc, orc_state *s, double dtau);
/* Backward-Euler Rayleigh sponge over dt, in place. */
void orc_dycore_sponge(orc_dycore *dc, orc_state *s, double dt);
/* Compensated domain totals: out = {mass, perturbation mass, kinetic,
 * internal, potential, kinetic+internal+potential}. */
void orc_dycore_energy(orc_dycore *dc, const orc_state *s, double out[6]);
/* Physics coupling (runge_kutta.cpp:46-63): theta' (K/s), u, v (m/s^2)
 * rates added to the slow tendency of every stage; theta's enters before
 * R_pi' is formed.  set_physics copies fixed rates (all NULL: off); a
 * callback is called before each stage's slow tendency with the stage state
 * and zeroed rate arrays to fill (non-zero return: the step returns 103). */
typedef int (*orc_physics_fn)(void *user, int stage, double dt, const orc_state *s, double *theta_rate,
                              double *u_rate, double *v_rate);
void orc_dycore_set_physics(orc_dycore *dc, const double *theta_rate, const double *u_rate, const double *v_rate);
void orc_dycore_set_physics_callback(orc_dycore *dc, orc_physics_fn fn, void *user);
void orc_set_threads(int n);
/* Planted mistakes, for the oracle's own pins only: orc_plant(n) changes one
 * term of the dycore so a test can show that its check fails under that
 * slip.  0 restores the true equations (every factor exactly 1.0, so results
 * are bit-identical to a build without the hook).
 *   1 buoyancy sign flipped          2 dtheta0/dz term halved
 *   3 Gal-Chen metric sign flipped   4 Omega slope-flux sign flipped
 *   5 Omega metric factor zeta in place of (1 - zeta) */
void orc_plant(int which);
/* Face primitives of the dycore (dycore.c), exported for tests:
 * orc_weno_face   WENO5 value at the face right of c, one-division form
 * orc_upwind_weno U times the upwind WENO5 value at the face between q-1, q0
 * orc_up5         5th-order upwind flux U c6 - |U| d5 at the face q0|q1 */
double orc_weno_face(double a, double b, double c, double d, double e);
double orc_upwind_weno(double U, double qm3, double qm2, double qm1, double q0, double q1, double q2);
double orc_up5(double U, double qm2, double qm1, double q0, double q1, double q2, double q3);

/* ---- nest lateral boundary (nest.c) ------------------------------------- */
void orc_davies_profile(int zone, double alpha_max, double *alpha);
/* full = 1: child <- parent interpolation everywhere; full = 0: Davies
 * relaxation of the zone toward pold + frac (pnew - pold). */
void orc_nest_apply(const orc_state *pnew, const orc_state *pold, int pnx, int pny, orc_state *child, int cnx,
                    int cny, int nz, int r, int i0, int j0, int zone, double alpha_max, double frac, int full);
/* Two-way feedback of a child onto its parent (restrict_to_parent role). */
void orc_nest_restrict(const orc_state *child, int cnx, int cny, orc_state *parent, int pnx, int pny, int nz,
                       int r, int i0, int j0);

#ifdef __cplusplus
}
#endif

#endif
