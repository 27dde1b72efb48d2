/* vibe_mi/batch.h -- C ABI of the batched dynamical-core step.
 *
 * Many domains of uneven size (ensemble members, nest levels) advanced
 * together on one GPU: each kernel phase of the RK3 step (the fused stage
 * kernel, each acoustic substep) is ONE persistent launch whose workgroups
 * take tiles of every active domain from a device-side work list; the
 * Rayleigh sponge is applied by the last stage kernel and the last substep
 * as they write the step's final state.  The launch sequence of a step can
 * be captured once in a hipGraph and replayed.  In FP64 every member
 * advances bit for bit as it would alone through vibe_mi_domain_step;
 * FP32-mixed members keep the intermediate RK3 stage states in FP32 and
 * agree with it to the mixed-precision tolerance.
 *
 * Replaces, for a set of domains on one GPU, the reference's per-domain
 * driver loop (src/driver/driver.cpp:288, one Integrator::step per domain
 * and step; the nest stub driver.cpp:212-219) with one batched step.
 *
 * Ownership: the batch owns one vibe_mi_domain per member (created from the
 * descriptors), its device work lists, counters and graphs.  Member handles
 * (vibe_mi_batch_member) stay valid until the batch is destroyed; use them
 * with the state / energy / CFL entry points of vibe_mi/dycore.h, never
 * destroy them.  Calls on one batch must not overlap.
 * Return codes as in vibe_mi/dycore.h.
 */
#ifndef VIBE_MI_BATCH_H
#define VIBE_MI_BATCH_H

#include <stddef.h>

#include "vibe_mi/dycore.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vibe_mi_batch vibe_mi_batch;

/* Creates the members from ndom descriptors (all of one precision). */
int vibe_mi_batch_create(const vibe_mi_domain_desc *descs, int ndom, vibe_mi_batch **out);
int vibe_mi_batch_destroy(vibe_mi_batch *b);
/* Member handle idx (0 .. ndom-1), owned by the batch. */
int vibe_mi_batch_member(vibe_mi_batch *b, int idx, vibe_mi_domain **out);
/* nsteps RK3 steps of every active member (active[idx] != 0; NULL = all),
 * member idx with its own dt[idx], nsub acoustic substeps per stage.
 * Returns VIBE_MI_ERR_NONFINITE if any member became non-finite. */
int vibe_mi_batch_step(vibe_mi_batch *b, const double *dt, int nsub, int nsteps,
                       const unsigned char *active);
/* Same, timed with HIP events on the batch stream (total ms). */
int vibe_mi_batch_step_timed(vibe_mi_batch *b, const double *dt, int nsub, int nsteps,
                             const unsigned char *active, double *ms_total);
/* Per-group times of the last timed call made without a graph (ms summed
 * over the call): the stage kernel per stage, and the stage's substep
 * launches, measured with HIP events on the batch stream. */
int vibe_mi_batch_kernel_times(vibe_mi_batch *b, double *ms_stage, double *ms_substep, int *n_stage,
                               int *n_substep);
/* use_graph = 1 (default): a step's launch sequence is captured once per
 * (active set, nsub) in a hipGraph and replayed; 0: launched eagerly. */
int vibe_mi_batch_set_graph(vibe_mi_batch *b, int use_graph);
/* Tile shapes the autotuner picked for the active set of the last step
 * (each active set is tuned on its first step): waves per substep
 * workgroup (4, 8, 16) and rows per stage workgroup (64 x rows columns:
 * 2, 4, 8).  The environment variables
 * VIBE_MI_SUBSTEP_WAVES and VIBE_MI_STAGE_ROWS pin either choice (and skip
 * the tuning cache). */
int vibe_mi_batch_tuning(vibe_mi_batch *b, int *substep_waves, int *stage_rows);
/* Total device bytes held by the batch and its members. */
size_t vibe_mi_batch_device_bytes(vibe_mi_batch *b);

#ifdef __cplusplus
}
#endif

#endif
