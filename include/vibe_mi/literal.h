/* vibe_mi/literal.h -- the reference's LITERAL discrete equations on the GPU.
 *
 * Replaces, on the device and with the reference's exact arithmetic:
 *   Equations::tendencies           src/dyn/equations.cpp:60-150
 *     (Advection::momentum          src/dyn/advection.cpp:250-319,
 *      Advection::scalar_nonconservative               :224-244,
 *      Coriolis::apply              src/dyn/coriolis.cpp:40-74,
 *      Damping::tendencies          src/dyn/damping.cpp:43-74)
 *   Equations::acoustic_tendencies  src/dyn/equations.cpp:152-209
 *   Equations::max_sound_speed      src/dyn/equations.cpp:238-249
 * with every convention of the reference kept: FULL prognostic values,
 * bounded x and y (neighbour indices clamped to the interior), zero halos
 * (the reference's single-rank path never fills them), the reference's
 * operation order.  Results equal the reference's bit for bit.
 *
 * The library's time step (dycore.h) integrates the balanced perturbation
 * equations instead; this entry point exists to tie the device arithmetic
 * to the reference's discrete operators term by term.
 *
 * Arrays (host, FP64, dense, i fastest), in the species order of the
 * reference's State (u, v, w, rho, theta, pi, qv, qc, qr, qi, qs, qg):
 * u [nz][ny][nx+1] (faces 0..nx), v [nz][ny+1][nx] (faces 0..ny),
 * w [nz+1][ny][nx] (faces 0..nz), the others [nz][ny][nx].
 * Return codes as in dycore.h.
 */
#ifndef VIBE_MI_LITERAL_H
#define VIBE_MI_LITERAL_H

#include <stddef.h>

#include "vibe_mi/c_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;
  double dx, dy, z_top;
  const double *zeta;   /* nz+1 normalised interfaces                        */
  const double *zs;     /* ny*nx terrain height (m), NULL: flat              */
  const double *rho0;   /* reference state at cells, [nz][ny][nx]           */
  const double *theta0;
  const double *pi0;
  int order;            /* numerics.advection: 2, 4 or 6 (central2/4/6)      */
  int coriolis;         /* numerics.coriolis: 0 none, 1 fplane, 2 betaplane  */
  double latitude;      /* numerics.fplane_latitude (degrees)                */
  double div_coeff;     /* numerics.divergence_damping (0: off)              */
  int div_order;        /* numerics.divergence_order: 2 (Del^2) or 4 (Del^4) */
  int moist;            /* physics.microphysics != "none": advect qc..qg     */
} vibe_mi_literal_desc;

typedef struct vibe_mi_literal vibe_mi_literal;

/* Uploads the grid and reference state and sizes the device buffers. */
int vibe_mi_literal_create(const vibe_mi_literal_desc *desc, vibe_mi_literal **out);
int vibe_mi_literal_destroy(vibe_mi_literal *h);
/* Equations::tendencies of state s into t (all 12 species written). */
int vibe_mi_literal_tendencies(vibe_mi_literal *h, const double *const s[12], double *const t[12]);
/* Equations::acoustic_tendencies of s into t (u, v, w, rho, pi written; the
 * other species are zero). */
int vibe_mi_literal_acoustic(vibe_mi_literal *h, const double *const s[12], double *const t[12]);
/* Equations::max_sound_speed of s. */
int vibe_mi_literal_max_sound_speed(vibe_mi_literal *h, const double *const s[12], double *out);

#ifdef __cplusplus
}
#endif

#endif /* VIBE_MI_LITERAL_H */
