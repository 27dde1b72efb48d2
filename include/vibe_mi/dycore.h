/* vibe_mi/dycore.h -- C ABI of the MI355X dynamical-core step.
 *
 * Replaces, for one GPU, the reference's time-step interface
 *   Integrator::step(State&, StepContext)        include/vibe/time/integrator.hpp:62-82
 *   RungeKutta3Integrator::step / acoustic_substep  src/time/runge_kutta.cpp:65-118
 *   Equations::tendencies / acoustic_tendencies  include/vibe/dyn/equations.hpp:61-82
 *   AcousticSubstepper::step (HEVI)              include/vibe/time/acoustic_substep.hpp:55-85
 *   Damping::apply_sponge                        src/dyn/damping.cpp:76-138
 *   energy_budget                                src/dyn/diagnostics.cpp:255-277
 *   TimeStepController / Advection::cfl_number   src/time/timestep_control.cpp:23-56
 * and the setup helpers the reference's driver uses to build inputs
 * (src/grid/geometry.cpp, src/dyn/reference_state.cpp,
 * src/driver/initial_conditions.cpp, src/dyn/damping.cpp:24-41).
 *
 * State layout (the build's own; dense, x fastest, periodic in x and y):
 *   u, v, rho', theta', pi' : [nz][ny][nx]   (u: west x-face, v: south y-face)
 *   w                       : [nz+1][ny][nx] (interface below cell k)
 * rho', theta', pi' are perturbations from the reference state rho0,
 * theta0, pi0 (cell centres, [nz][ny][nx]).  All host arrays are FP64.
 *
 * Ownership: a domain handle owns its device buffers (allocated once at
 * creation); step/substep/tendency calls never allocate.  Calls on one
 * handle must not overlap; different handles are independent.
 * Return codes: 0 ok, 2 bad argument, >= 100 device/runtime errors
 * (see c_kernels.h).
 */
#ifndef VIBE_MI_DYCORE_H
#define VIBE_MI_DYCORE_H

#include <stddef.h>
#include <stdint.h>

#include "vibe_mi/c_kernels.h" /* error codes */

#ifdef __cplusplus
extern "C" {
#endif

#define VIBE_MI_FP64 0  /* plain FP64: storage, compute and reductions   */
#define VIBE_MI_MIXED 1 /* FP64 state + RK3 accumulation, FP32 compute,  */
                        /* FP64 compensated reductions                   */
/* FP32-compensated: VIBE_MI_MIXED whose FP32 acoustic (substep) state keeps
 * an 8-bit remainder of u (hi + lo, about 31 significant bits), to which the
 * substeps add u's FP32 increments with an exact two-sum, and whose substeps
 * form the x part of the horizontal mass-flux divergence in FP64: within 1e-4 of FP64
 * after 1 h where VIBE_MI_MIXED is not, at 5 instead of 4 bytes per stored
 * u value. */
#define VIBE_MI_MIXED_COMP 2

typedef struct {
  int nx, ny, nz;
  double dx, dy, ztop;
  const double *zeta;   /* nz+1 normalised interfaces (0 .. 1)          */
  const double *zs;     /* ny*nx terrain height (m)                     */
  const double *rho0;   /* nz*ny*nx reference density                   */
  const double *theta0; /* nz*ny*nx reference potential temperature     */
  const double *pi0;    /* nz*ny*nx reference Exner function            */
  const double *sponge; /* nz Rayleigh rates (1/s)                      */
  double eps_s;         /* HEVI off-centering (0 = Crank-Nicolson)      */
  int precision;        /* VIBE_MI_FP64, _MIXED or _MIXED_COMP          */
} vibe_mi_domain_desc;

typedef struct vibe_mi_domain vibe_mi_domain;

/* Sizes: nx, ny >= 4 (the 3-cell periodic stencils); 3 <= nz <= 103 in
 * FP64 and 3 <= nz <= 128 in FP32-mixed (the tallest column one substep
 * workgroup solves with its factorisation in LDS).
 * VIBE_C_ERR_BAD_ARG also when nx*ny*(nz+1)*8 bytes reach 4 GiB (the kernels
 * address a field with 32-bit byte offsets); larger grids run as several
 * members of a batch. */
int vibe_mi_domain_create(const vibe_mi_domain_desc *desc, vibe_mi_domain **out);
int vibe_mi_domain_destroy(vibe_mi_domain *d);
/* Upload / download the FP64 state (host arrays in the layout above). */
int vibe_mi_domain_set_state(vibe_mi_domain *d, const double *u, const double *v,
                             const double *w, const double *rho, const double *theta,
                             const double *pi);
int vibe_mi_domain_get_state(vibe_mi_domain *d, double *u, double *v, double *w,
                             double *rho, double *theta, double *pi);
/* nsteps WS-RK3 steps of length dt with nsub acoustic substeps per stage
 * (runge_kutta.cpp:77-118).  Returns VIBE_MI_ERR_NONFINITE if the state
 * became non-finite in any stage. */
int vibe_mi_domain_step(vibe_mi_domain *d, double dt, int nsub, int nsteps);
/* Same, timed with HIP events on the domain's stream: total elapsed ms. */
int vibe_mi_domain_step_timed(vibe_mi_domain *d, double dt, int nsub, int nsteps,
                              double *ms_total);
/* Per-kernel time of the last timed call (ms summed over the call). */
int vibe_mi_domain_kernel_times(vibe_mi_domain *d, double *ms_stage, double *ms_substep,
                                int *n_stage, int *n_substep);
/* Slow tendencies of the current state (equations.cpp:60-150 role), copied
 * to host arrays (layout of the state). */
int vibe_mi_domain_slow_tendency(vibe_mi_domain *d, double *ru, double *rv, double *rw,
                                 double *rrho, double *rtheta, double *rpi);
/* One acoustic substep in place (runge_kutta.cpp:65-75 role). */
int vibe_mi_domain_substep(vibe_mi_domain *d, double dtau);

/* Physics coupling: rates added to the slow tendency R of every RK stage,
 * before the substeps (RungeKutta3Integrator::slow_tendency with a
 * PhysicsDriver, src/time/runge_kutta.cpp:46-63; PhysicsTendency,
 * include/vibe/dyn/tendency.hpp:65-80): theta' in K/s, u and v in m/s^2,
 * FP64, in the state layout ([nz][ny][nx]; u at west faces, v at south
 * faces).  The state is dry, so there are no moisture rates.  The heating
 * enters R_theta before R_pi' is formed from R_rho' and R_theta (the
 * linearised equation of state), so pi' follows it.  Steps, batched steps
 * and vibe_mi_domain_slow_tendency all add the rates.
 *
 * set_physics_tendency holds the given rates (copied to the device) for
 * every following stage; NULL for a field means zero; all three NULL
 * switches the coupling off. */
int vibe_mi_domain_set_physics_tendency(vibe_mi_domain *d, const double *theta_rate, const double *u_rate,
                                        const double *v_rate);
/* The device rate buffers (nx*ny*nz FP64 each; allocated and zeroed on the
 * first call, which switches the coupling on): a GPU physics package writes
 * them on vibe_mi_domain_stream(d) between steps. */
int vibe_mi_domain_physics_buffers(vibe_mi_domain *d, double **theta_rate, double **u_rate, double **v_rate);
/* Switches the physics coupling of the following steps off (on = 0) or back
 * on (StepContext::physics_on, integrator.hpp:42-49), keeping the held rates,
 * buffers and callback.  On by default. */
int vibe_mi_domain_physics_enable(vibe_mi_domain *d, int on);
/* Per-stage physics, the reference's call pattern (physics evaluated on each
 * stage's state: slow_tendency at runge_kutta.cpp:88).  Before the
 * slow tendency of each stage the step zeroes the rate buffers and calls
 * fn(user, stage 0..2, dt of the step, stream, stage state, rates).
 * host = 1: state and rates are host arrays (the step synchronises, copies
 * the stage state to the host and uploads the rates after fn returns);
 * host = 0: device pointers, and fn enqueues its work on `stream`.  A
 * non-zero return from fn ends the step with VIBE_MI_ERR_PHYSICS.  fn = NULL
 * removes the callback and switches the coupling off.  Single-domain steps
 * only: vibe_mi_batch_step returns VIBE_C_ERR_BAD_ARG for a member with a
 * callback (a batched step holds the rates of set_physics_tendency). */
typedef int (*vibe_mi_physics_fn)(void *user, int stage, double dt, void *stream, const double *u, const double *v,
                                  const double *w, const double *rho_p, const double *theta_p, const double *pi_p,
                                  double *theta_rate, double *u_rate, double *v_rate);
int vibe_mi_domain_set_physics_callback(vibe_mi_domain *d, vibe_mi_physics_fn fn, void *user, int host);
/* Compensated FP64 domain totals {mass, perturbation mass, kinetic,
 * internal, potential, total} (diagnostics.cpp:255-277). */
int vibe_mi_domain_energy(vibe_mi_domain *d, double out[6]);
/* Advective CFL of the current state (advection.cpp:325-340). */
int vibe_mi_domain_cfl(vibe_mi_domain *d, double dt, double *cfl, double *max_wind);
/* Stability report of the current state for a step dt
 * (TimeStepController::inspect, timestep_control.cpp:23-56): out =
 * {advective CFL, acoustic CFL, vertical CFL, max wind speed, min dz,
 * stable}, stable = 1 when the state is finite, the advective CFL < 1.5 and
 * the vertical CFL < 2, else 0. */
int vibe_mi_domain_stability(vibe_mi_domain *d, double dt, double out[6]);
/* Largest stable large step of the current state
 * (RungeKutta3Integrator::max_stable_dt, runge_kutta.cpp:121-141):
 * min(0.8 / advective CFL at dt = 1, 2 / max over cells of
 * cs_max (1/dx + 1/dy + 1/dz)), cs_max = the largest sqrt(cp/cv Rd max(T, 1))
 * of the state (Equations::max_sound_speed), dz = max(dzeta (ztop - zs), 1). */
int vibe_mi_domain_max_stable_dt(vibe_mi_domain *d, double *dt_out);
/* hipStream_t of the domain, as void*. */
void *vibe_mi_domain_stream(vibe_mi_domain *d);
/* Bytes of device memory the handle holds. */
size_t vibe_mi_domain_device_bytes(vibe_mi_domain *d);

/* ---- the reference's State layout (host only) -------------------------- */
/* One reference field (include/vibe/grid/field.hpp:34-96): nx x ny x nz
 * interior points, `halo` points on every side, one more face along the
 * staggered axis (u: x, v: y, w: z), i fastest:
 *   offset(i, j, k) = ((k + halo) nsy + (j + halo)) nsx + (i + halo),
 *   nsx = nx + 2 halo (+1 for u), nsy = ny + 2 halo (+1 for v),
 *   nsz = nz + 2 halo (+1 for w).
 * species[12] are the 12 fields of a dyn::State in the order of dyn::Species
 * (state.hpp:53-56): u, v, w, rho, theta, pi, qv, qc, qr, qi, qs, qg, as
 * totals; rho0, theta0, pi0 are the ReferenceState's cell fields in the same
 * layout.  Elements of one field (0 for bad arguments; species 0..11). */
size_t vibe_mi_reference_field_size(int nx, int ny, int nz, int halo, int species);
/* Reference State -> this library's perturbation state (dycore.h layout):
 * u, v from faces 0..nx-1 / 0..ny-1 (face nx / ny duplicates face 0 in a
 * periodic domain), w from interfaces 0..nz, rho' = rho - rho0 and likewise
 * theta', pi'.  Halo points are not read; the moisture species
 * (species[6..11], may be NULL) are not read either: the step is dry. */
int vibe_mi_state_from_reference(int nx, int ny, int nz, int halo, const double *const species[12],
                                 const double *rho0, const double *theta0, const double *pi0, double *u, double *v,
                                 double *w, double *rho_p, double *theta_p, double *pi_p);
/* Perturbation state -> reference State totals: u at faces 0..nx (face nx =
 * face 0), v at 0..ny, w at 0..nz, rho = rho0 + rho' and likewise theta, pi
 * on the interior; fill_halo = 1 also fills the lateral halo of every
 * interior level with the periodic values (HaloExchange::exchange_local role,
 * halo.cpp:58-86; the reference state taken at the periodic interior point),
 * 0 leaves halos untouched.  Vertical halo levels and the moisture species
 * are never written. */
int vibe_mi_state_to_reference(int nx, int ny, int nz, int halo, const double *u, const double *v, const double *w,
                               const double *rho_p, const double *theta_p, const double *pi_p, const double *rho0,
                               const double *theta0, const double *pi0, int fill_halo, double *const species[12]);

/* ---- setup helpers (host only; inputs to the step) --------------------- */
/* zeta = k/nz */
int vibe_mi_zeta_uniform(int nz, double *zeta);
/* tanh-stretched layers from dz_bottom to dz_top summing to ztop. */
int vibe_mi_zeta_tanh(int nz, double ztop, double dz_bottom, double dz_top, double *zeta);
/* Witch-of-Agnesi ridge h a^2/(a^2+(x-xc)^2), x at cell centres, xc at the
 * domain centre, uniform in y (geometry.cpp:103-117 is the radial form). */
int vibe_mi_terrain_agnesi_ridge(int nx, int ny, double dx, double h, double a, double *zs);
/* Geometric vertical coordinate (make_stretched_zeta, geometry.cpp:39-60):
 * layer thicknesses first_thickness * stretch^k, normalised; zeta[nz] = 1. */
int vibe_mi_zeta_stretched(int nz, double first_thickness, double stretch, double *zeta);
/* Radial Witch of Agnesi h0 a^2/(a^2 + d^2), d from (xc, yc) to the point
 * (i dx, j dy) (witch_of_agnesi, geometry.cpp:103-117). */
int vibe_mi_terrain_agnesi_bell(int nx, int ny, double dx, double dy, double h0, double a, double xc, double yc,
                                double *zs);
/* In-place terrain smoothing (smooth_terrain, geometry.cpp:70-101): per
 * pass, a 1-2-1 filter of the interior that only lowers points, then the
 * interior capped at its highest neighbour + max_slope (if max_slope > 0);
 * the outer rows and columns keep their values. */
int vibe_mi_terrain_smooth(double *zs, int nx, int ny, int passes, double max_slope);
/* Seeded Gaussian random terrain, std sigma, Gaussian correlation length L,
 * periodic, minimum shifted to 0. */
int vibe_mi_terrain_gaussian(int nx, int ny, double dx, double dy, double sigma, double L,
                             uint64_t seed, double *zs);
/* Reference state at cell centres.  kind 0: constant theta (isothermal-theta,
 * reference_state.cpp:35-59); kind 1: constant Brunt-Vaisala frequency bv. */
int vibe_mi_reference_state(int kind, double theta_sfc, double bv, double p_sfc, int nx,
                            int ny, int nz, double ztop, const double *zeta, const double *zs,
                            double *rho0, double *theta0, double *pi0);
/* Reference state from a (z, theta) sounding: replaces
 * ReferenceState::from_profile (reference_state.cpp:61-110) with the same
 * discrete integration, bit for bit: theta linear in the profile at the
 * interfaces, pi integrated upward with the interface-mean theta, theta0 =
 * that mean, pi0 = pi at the cell's upper interface, rho0 from the EOS.
 * np >= 2 profile points, zp increasing.  Returns 0, or 2 on bad arguments. */
int vibe_mi_reference_state_from_profile(int nx, int ny, int nz, double ztop, const double *zeta,
                                         const double *zs, int np, const double *zp, const double *thp,
                                         double p_sfc, double *rho0, double *theta0, double *pi0);
/* Rayleigh sponge rates per level (damping.cpp:24-41, column (0,0)). */
int vibe_mi_sponge(int nz, double ztop, const double *zeta, double zs00, double start_fraction,
                   double alpha_max, double *alpha);
/* Warm/cold bubble theta' += amp cos^2(pi r/2) (initial_conditions.cpp:136-159). */
int vibe_mi_ic_bubble(int nx, int ny, int nz, double dx, double dy, double ztop,
                      const double *zeta, const double *zs, double xc, double yc, double zc,
                      double rx, double rz, double amp, double *theta);
/* Seeded uniform noise theta' += U(-amp, amp) below zmax (SplitMix64). */
int vibe_mi_ic_noise(int nx, int ny, int nz, double ztop, const double *zeta, const double *zs,
                     double amp, double zmax, uint64_t seed, double *theta);
/* u = u0 + umax * tanh(z / hscale) (hscale <= 0: uniform u0). */
int vibe_mi_ic_wind(int nx, int ny, int nz, double ztop, const double *zeta, const double *zs,
                    double u0, double umax, double hscale, double *u);
/* pi' from the equation of state at fixed rho (enforce_state_equation,
 * initial_conditions.cpp:110-121), dry. */
int vibe_mi_ic_balance_pi(size_t ncell, const double *rho0, const double *theta0,
                          const double *rho, const double *theta, double *pi);

#ifdef __cplusplus
}
#endif

#endif
