/* vibe_mi/nest.h -- C ABI of the nest boundaries (one-way and two-way).
 *
 * Replaces, on the device, the reference's nest operators
 *   Nest::interpolate_from_parent    src/grid/nest.cpp:52-83
 *   Nest::restrict_to_parent         src/grid/nest.cpp:85-108
 *   Nest::apply_lateral_boundary     src/grid/nest.cpp:110-128
 *   davies_profile                   src/grid/nest.cpp:34-43
 *   vibe_bilinear_f64 (arithmetic)   src/common/c_kernels.c:108-144
 * for child and parent domains that live on the same GPU (vibe_mi/dycore.h
 * handles, batch members included).  The parent is periodic (the build's
 * domains are), so parent indices wrap instead of clamping.
 *
 * Placement: child cell (i, j) has its centre at parent index coordinates
 *   xp = (i0 + i + 0.5) / r - 0.5,  yp = (j0 + j + 0.5) / r - 0.5
 * (nest.cpp:67-68: i0, j0 are offsets counted in child cells).  Cell fields
 * (rho', theta', pi', and w, which sits at cell centres horizontally) use that
 * point; u (west x-face) uses xp = (i0 + i) / r and v (south y-face)
 * yp = (j0 + j) / r, each field at its own staggered position.  Levels map
 * one to one (shared vertical coordinate).  Perturbation fields are
 * interpolated, each domain keeping its own reference state.
 *
 * All arithmetic is FP64 on the FP64 states of both domains.
 */
#ifndef VIBE_MI_NEST_H
#define VIBE_MI_NEST_H

#include "vibe_mi/batch.h"
#include "vibe_mi/dycore.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int ratio;        /* refinement ratio r >= 2 (NestPlacement::ratio)           */
  int i0, j0;       /* child origin offset from the parent origin, child cells */
  int zone;         /* Davies zone width in child points (boundary_zone), 1..16 */
  double alpha_max; /* Davies weight at the outer boundary (davies_profile)   */
} vibe_mi_nest_desc;

/* Davies weights alpha(n) = alpha_max (1 - cos(pi (zone - n) / zone)) / 2,
 * n = 0 .. zone-1 (nest.cpp:34-43), written to alpha[zone]. */
int vibe_mi_davies_profile(int zone, double alpha_max, double *alpha);

/* Child state <- bilinear interpolation of the parent's current state over
 * the whole child (initialisation; nest.cpp:52-83). */
int vibe_mi_nest_interpolate(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd);

/* Davies relaxation of the child's lateral zone (points at distance
 * d = min(i, nx-1-i, j, ny-1-j) < zone from the child's edge) toward the
 * parent state interpolated bilinearly in space and linearly in time:
 *   p = p_old + frac (p_new - p_old),  c <- (1 - alpha(d)) c + alpha(d) p
 * where p_new is the parent's current state and p_old its state before its
 * last step (nest.cpp:110-128).  frac in [0, 1]. */
int vibe_mi_nest_relax(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd,
                       double frac);

/* The same relaxation between members `child` and `parent` of one batch,
 * enqueued on the batch's stream behind its last step with no host
 * synchronisation (vibe_mi_nest_relax synchronises both domains' streams):
 * the batch's next step, and later work on the child's own stream, are
 * ordered after it.  The relaxation between the sub-steps of a nest
 * hierarchy (vibe-model_amd/vibe_mi/nesting.py). */
int vibe_mi_batch_nest_relax(vibe_mi_batch *b, int child, int parent, const vibe_mi_nest_desc *nd, double frac);

/* Two-way feedback (Nest::restrict_to_parent, nest.cpp:85-108): every parent
 * cell covered by whole child cells, (i0/r + i, j0/r + j) for i < nx_child/r
 * and j < ny_child/r, takes the mean of the child state over it: rho', theta',
 * pi' and w the mean of the r x r child values (summed row by row, then times
 * 1/r^2, as the reference), u on the parent's west face the mean of the r
 * child faces on it, v likewise.  The reference averages u and v as cell
 * fields and reads its i0, j0 as parent cells here (as child cells in its
 * interpolation); this entry keeps one placement for all three operators, so
 * i0 and j0 must be multiples of r (VIBE_C_ERR_BAD_ARG otherwise).  Writes
 * the parent's current state; its state before its last step is untouched. */
int vibe_mi_nest_restrict(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd);

/* The same feedback between members of one batch, enqueued on the batch's
 * stream with no host synchronisation. */
int vibe_mi_batch_nest_restrict(vibe_mi_batch *b, int child, int parent, const vibe_mi_nest_desc *nd);

/* A run of parent steps of a nest hierarchy held in one batch: links
 * (child[q], parent[q], descs[q]), q < nlinks, one ratio r for all of them,
 * each child with one parent; level l = depth in the tree; dt0 the parent
 * step.  One parent step is r^L sub-steps: at sub-step s the members of
 * level l step by dt0 / r^l when s % r^(L-l) == 0, all in the same batched
 * launches, then every nest that stepped is relaxed toward its parent
 * (vibe_mi_batch_nest_relax, frac = (m + 1) / r for its m-th step in the
 * parent's step), coarse levels first; with two_way, after the last sub-step
 * every nest feeds back onto its parent (vibe_mi_batch_nest_restrict),
 * finest first.  With graphs on (vibe_mi_batch_set_graph) the whole parent
 * step is captured once per buffer-rotation state and replayed: one graph
 * launch per parent step, one host synchronisation per call.  Bit-identical
 * to the same schedule issued call by call.  Members with a physics callback
 * are rejected (held rates are fine). */
int vibe_mi_batch_nested_step(vibe_mi_batch *b, int nlinks, const int *child, const int *parent,
                              const vibe_mi_nest_desc *descs, int two_way, double dt0, int nsub, int nsteps);
/* The same, timed with HIP events on the batch stream (ms over the call). */
int vibe_mi_batch_nested_step_timed(vibe_mi_batch *b, int nlinks, const int *child, const int *parent,
                                    const vibe_mi_nest_desc *descs, int two_way, double dt0, int nsub, int nsteps,
                                    double *ms_total);

#ifdef __cplusplus
}
#endif

#endif
