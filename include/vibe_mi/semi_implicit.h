/* vibe_mi/semi_implicit.h -- the reference's semi-implicit integrator on the
 * GPU (C ABI, device-resident state).
 *
 * Replaces, with the reference's arithmetic on its literal equations
 * (vibe_mi/literal.h: FULL prognostic values, bounded x and y, the
 * reference's operation order):
 *   SemiImplicitIntegrator::step           src/time/semi_implicit.cpp:167-226
 *   SemiImplicitIntegrator::assemble_rhs   src/time/semi_implicit.cpp:58-108
 *   SemiImplicitIntegrator::max_stable_dt  src/time/semi_implicit.cpp:228-233
 *   Integrator::stats / reset_stats        include/vibe/time/integrator.hpp:53-68
 *   set_time_filter                        include/vibe/time/semi_implicit.hpp:65
 * One step of state s with time step dt, as the reference performs it:
 *   1. N = Equations::tendencies(s), plus the held physics rates of
 *      vibe_mi_si_set_physics_tendency on theta, qv, u and v;
 *   2. a = (dt/2)^2 (Rd/cv) pi0 cp theta0, b = 1 and
 *      rhs = pi - dt c G + div(a grad pi) (horizontal part, a of the cell),
 *      c = (Rd/cv) pi0 / max(rho0, 1e-8), G = div(rho0 u) + (dt/2) div(rho0 N_u)
 *      (face rho0 the mean of the two cells, neighbours clamped,
 *      dz = max(dzeta (ztop - zs), 1));
 *   3. A pi_new = rhs solved with the configured Helmholtz solver
 *      (vibe_mi/helmholtz.h) to the relative residual tol; its iterations and
 *      residual go to the statistics;
 *   4. theta, qv, rho and qc..qg += dt N (forward Euler);
 *   5. with a time filter alpha > 0, after the first step:
 *      pi = (1 - alpha) pi + alpha pi;
 *   6. the sponge (Damping::apply_sponge, damping.cpp:76-138) with the given
 *      per-level rates: u, v, w relaxed to 0, rho, theta, pi to the
 *      reference state.
 * As in the reference, the solution pi_new updates neither u, v, w nor pi:
 * the reference applies update_from_pressure to a copy of the state that it
 * then drops (semi_implicit.cpp:185-196), so only step 4-6 change the state.
 * pi_new stays readable through vibe_mi_si_pressure.  A non-finite state
 * after the step returns VIBE_MI_ERR_NONFINITE and sets stats.diverged (the
 * reference throws NumericalError, semi_implicit.cpp:222-225).
 *
 * Arrays: the species layout of vibe_mi/literal.h.  Calls on one handle must
 * not overlap.  Return codes as in dycore.h.
 */
#ifndef VIBE_MI_SEMI_IMPLICIT_H
#define VIBE_MI_SEMI_IMPLICIT_H

#include <stddef.h>

#include "vibe_mi/helmholtz.h"
#include "vibe_mi/literal.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  vibe_mi_literal_desc eq; /* grid, reference state, numerics of Equations   */
  const double *sponge;    /* nz per-level rates (Damping's sponge_), NULL: none */
  int solver;              /* VIBE_MI_HELM_BICGSTAB, _REFINE or _MULTIGRID      */
  double tol;              /* numerics.helmholtz_tol                           */
  int max_iter;            /* numerics.helmholtz_max_iter                      */
  double time_filter;      /* set_time_filter alpha (0: off)                   */
} vibe_mi_si_desc;

/* Same fields as the reference's IntegratorStats (integrator.hpp:53-59). */
typedef struct {
  long long steps;
  long long acoustic_substeps; /* 0: no acoustic substeps in this integrator */
  double wall_time;            /* seconds of the last vibe_mi_si_step call   */
  int helmholtz_iterations;    /* summed over the steps                      */
  double helmholtz_residual;   /* relative residual of the last solve        */
  int diverged;
} vibe_mi_si_stats;

typedef struct vibe_mi_si vibe_mi_si;

int vibe_mi_si_create(const vibe_mi_si_desc *desc, vibe_mi_si **out);
int vibe_mi_si_destroy(vibe_mi_si *h);
/* Host state (12 species) to and from the device copy. */
int vibe_mi_si_set_state(vibe_mi_si *h, const double *const s[12]);
int vibe_mi_si_get_state(vibe_mi_si *h, double *const s[12]);
/* Physics rates added to N every step (PhysicsDriver::step's
 * PhysicsTendency, semi_implicit.cpp:175-180): theta and qv cells, u and v
 * faces, host arrays copied and held; a NULL entry is zero, all NULL
 * switches the coupling off. */
int vibe_mi_si_set_physics_tendency(vibe_mi_si *h, const double *theta, const double *qv, const double *u,
                                    const double *v);
/* set_time_filter (semi_implicit.hpp:65): alpha >= 0, 0 switches it off. */
int vibe_mi_si_set_time_filter(vibe_mi_si *h, double alpha);
/* nsteps steps of dt on the device state. */
int vibe_mi_si_step(vibe_mi_si *h, double dt, int nsteps);
int vibe_mi_si_get_stats(vibe_mi_si *h, vibe_mi_si_stats *out);
int vibe_mi_si_reset_stats(vibe_mi_si *h);
/* max_stable_dt of the device state: 0.6 / (advective CFL at dt = 1), or 600
 * when the state is at rest (Advection::cfl_number, advection.cpp:325-340). */
int vibe_mi_si_max_stable_dt(vibe_mi_si *h, double *dt_out);
/* assemble_rhs of the device state for a step dt (its tendencies and the
 * held physics rates included) into host cell arrays a and rhs (either may
 * be NULL); the state is not changed. */
int vibe_mi_si_rhs(vibe_mi_si *h, double dt, double *a, double *rhs);
/* pi_new of the last step's Helmholtz solve (host cell array). */
int vibe_mi_si_pressure(vibe_mi_si *h, double *pi_new);

#ifdef __cplusplus
}
#endif

#endif /* VIBE_MI_SEMI_IMPLICIT_H */
