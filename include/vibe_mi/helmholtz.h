/* vibe_mi/helmholtz.h -- 3-D Helmholtz solver of the semi-implicit path
 * (C ABI, device-resident).
 *
 * Replaces, on the GPU, the reference's 3-D solvers
 *   HelmholtzSolver (set_coefficients / solve / iterations / residual)
 *                                   include/vibe/time/helmholtz.hpp:63-86
 *   apply_operator                  src/time/helmholtz.cpp:173-201
 *   helmholtz_residual_l2           src/time/helmholtz.cpp:211-225
 *   KrylovJacobiSolver (BiCGSTAB)   src/time/helmholtz.cpp:286-379
 *   RefinementSolver (low-precision solve + FP64 residual correction)
 *                                   src/time/helmholtz.cpp:420-476
 *   KrylovMultigridSolver (coarse-grid start of the Krylov solve)
 *                                   src/time/helmholtz.cpp:382-417
 *   multigrid_restrict / multigrid_prolong
 *                                   src/time/helmholtz.cpp:131-164
 * for the operator of the reference,
 *
 *   y = -div(a grad x) + b x,
 *   -div(a grad x) = -[(ae (x_E - x) - aw (x - x_W)) / dx^2
 *                     + (an (x_N - x) - as (x - x_S)) / dy^2
 *                     + (at (x_T - x) - ab (x - x_B)) / dz^2],
 *
 * with face coefficients the mean of the two cell values (ae = (a + a_E)/2,
 * ...), neighbours clamped at every boundary (zero flux: x_E = x at the last
 * column, likewise W, N, S, top and bottom) and dz = max(dzeta_k (ztop - zs),
 * 1) per column and level.  a and b are cell arrays [nz][ny][nx], b > 0 for
 * a definite operator.
 *
 * Solvers (the operator is not symmetric on a stretched grid, hence
 * BiCGSTAB as in the reference), each preconditioned by exact solves of the
 * vertical lines (per-column tridiagonal: the vertical couplings plus the
 * horizontal diagonal), which removes the stiffness of thin layers:
 *   VIBE_MI_HELM_BICGSTAB   FP64 BiCGSTAB to the relative residual tol;
 *   VIBE_MI_HELM_REFINE     FP32 BiCGSTAB inner solves of the FP64 residual
 *                           equation, x += dx, until the FP64 relative
 *                           residual ||f - A x|| / ||f|| < tol (iterative
 *                           refinement);
 *   VIBE_MI_HELM_MULTIGRID  two-grid start: f, a, b and the terrain restricted
 *                           (multigrid_restrict) to the (nx+1)/2 x (ny+1)/2 x nz
 *                           grid of spacing 2 dx, 2 dy, that problem solved to
 *                           the looser of tol and 1e-3, its solution prolonged
 *                           (multigrid_prolong) as the initial guess of the
 *                           FP64 BiCGSTAB on the fine grid.  (The reference's
 *                           KrylovMultigridSolver forms the same start on a
 *                           field of the fine grid's size and its inner solver
 *                           then restarts from zero, so its start has no
 *                           effect; here it is a real coarse grid and is used.)
 * Iterations run without host round trips (scalars stay on the device); the
 * host reads the residual every few iterations.  Calls on one handle must not
 * overlap.  Return codes as in dycore.h.
 */
#ifndef VIBE_MI_HELMHOLTZ_H
#define VIBE_MI_HELMHOLTZ_H

#include <stddef.h>

#include "vibe_mi/c_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VIBE_MI_HELM_BICGSTAB 0
#define VIBE_MI_HELM_REFINE 1
#define VIBE_MI_HELM_MULTIGRID 2

typedef struct {
  int nx, ny, nz;
  double dx, dy, ztop;
  const double *zeta; /* nz+1 normalised interfaces; dzeta_k = zeta[k+1] - zeta[k] */
  const double *zs;   /* ny*nx terrain height (m)                                  */
} vibe_mi_helmholtz_desc;

typedef struct vibe_mi_helmholtz vibe_mi_helmholtz;

int vibe_mi_helmholtz_create(const vibe_mi_helmholtz_desc *desc, vibe_mi_helmholtz **out);
int vibe_mi_helmholtz_destroy(vibe_mi_helmholtz *h);
/* a, b: host cell arrays [nz][ny][nx] (set_coefficients role); also builds
 * the line preconditioner.  VIBE_C_ERR_BAD_ARG if b <= 0 or a < 0 anywhere. */
int vibe_mi_helmholtz_set_coefficients(vibe_mi_helmholtz *h, const double *a, const double *b);
/* y = A x in FP64 on the device (apply_operator), host arrays. */
int vibe_mi_helmholtz_apply(vibe_mi_helmholtz *h, const double *x, double *y);
/* ||A x - f||_2 (helmholtz_residual_l2), host arrays. */
int vibe_mi_helmholtz_residual(vibe_mi_helmholtz *h, const double *f, const double *x, double *res_l2);
/* Solves A x = f from x = 0 (the reference's solvers start from zero; the
 * multigrid kind from its coarse-grid start): kind VIBE_MI_HELM_BICGSTAB,
 * VIBE_MI_HELM_REFINE or VIBE_MI_HELM_MULTIGRID, relative residual tol,
 * at most max_iter BiCGSTAB iterations in total.  iterations and rel_res
 * (either may be NULL) report the iterations taken and the final FP64
 * relative residual; like the reference, not converging within max_iter is
 * not an error (check rel_res). */
int vibe_mi_helmholtz_solve(vibe_mi_helmholtz *h, const double *f, double *x, int kind, double tol, int max_iter,
                            int *iterations, double *rel_res);

/* multigrid_restrict (helmholtz.cpp:131-145): 2:1 full weighting in x and
 * y at every level,
 *   coarse(i,j,k) = 1/4 f(2i,2j,k) + 1/8 [f(2i-1,2j,k) + f(2i+1,2j,k)
 *                   + f(2i,2j-1,k) + f(2i,2j+1,k)] + 1/16 [f(2i-1,2j-1,k)
 *                   + f(2i+1,2j-1,k) + f(2i-1,2j+1,k) + f(2i+1,2j+1,k)],
 * fine indices (k included) clamped to the fine extents, for every point of
 * the coarse field.  fine [fnz][fny][fnx], coarse [cnz][cny][cnx]: DEVICE
 * arrays; runs on `stream` (hipStream_t, NULL: the null stream), allocates
 * nothing.  Results equal the reference's bit for bit. */
int vibe_mi_multigrid_restrict_f64(const double *fine, int fnx, int fny, int fnz, double *coarse, int cnx, int cny,
                                   int cnz, void *stream);
/* multigrid_prolong (helmholtz.cpp:147-164): bilinear interpolation, fine
 * point (i, j) at coarse coordinate (i/2, j/2), coarse indices (k included)
 * clamped, lerp(a, b, w) = a + w (b - a); DEVICE arrays as above. */
int vibe_mi_multigrid_prolong_f64(const double *coarse, int cnx, int cny, int cnz, double *fine, int fnx, int fny,
                                  int fnz, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* VIBE_MI_HELMHOLTZ_H */
