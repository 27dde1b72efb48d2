/* vibe_mi/c_kernels.h -- C11 ABI of the MI355X hot kernels.
 *
 * Drop-in for the reference's include/vibe/common/c_kernels.h:24-119.  The six
 * vibe_*_f64 entry points keep the reference's names, byte-identical
 * signatures, argument meaning and return codes; every one of them runs its
 * arithmetic in a HIP kernel on the current GPU.  The reference's contract
 * (c_kernels.h:13-19) holds: no entry point allocates (no hipMalloc, hipFree,
 * stream creation or host heap) and none keeps state that changes its
 * results; all are re-entrant and may be called from concurrent threads.
 *   - Arrays the GPU can dereference (device memory of the current GPU,
 *     hipHostMalloc'd host memory) are read and written in place.
 *   - Pageable host arrays are copied in chunks through a statically sized
 *     device array of the library (256 MiB per GPU, loaded with the library
 *     and held until the process ends), taken under a per-GPU lock for the
 *     length of a call: concurrent calls on one GPU serialise.  Limits of
 *     the staged path: tridiagonal n <= 5.5e6 (f64); flux advection
 *     nx*ny <= 1.9e6; other sizes are unlimited.  Beyond them the call
 *     returns (or, for void functions, reports through vibe_mi_last_error)
 *     VIBE_MI_ERR_TOO_LARGE.
 *   - Work runs on hipStreamPerThread; every call returns with its results
 *     in place.
 * The _f32 twins follow the header convention of c_kernels.h:15-16 (none
 * exist in the reference).  The vibe_mi_*_batched entry points take DEVICE
 * pointers, the caller's workspace and a hipStream_t (passed as void*; NULL
 * is the null stream), and never allocate.
 *
 * Errors: VIBE_C_OK / VIBE_C_ERR_SINGULAR / VIBE_C_ERR_BAD_ARG as in the
 * reference (c_kernels.h:34-36); failures of the HIP runtime itself are
 * reported in a separate range (>= VIBE_MI_ERR_HIP).  Void entry points of the
 * reference stay void; their device failures are readable through
 * vibe_mi_last_error().
 */
#ifndef VIBE_MI_C_KERNELS_H
#define VIBE_MI_C_KERNELS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIBE_C_OK 0
#define VIBE_C_ERR_SINGULAR 1
#define VIBE_C_ERR_BAD_ARG 2
#define VIBE_MI_ERR_HIP 100       /* a HIP runtime call failed            */
#define VIBE_MI_ERR_NO_DEVICE 101 /* no usable GPU is visible             */
#define VIBE_MI_ERR_NONFINITE 102 /* the model state became non-finite    */
#define VIBE_MI_ERR_PHYSICS 103   /* a physics callback returned non-zero */
#define VIBE_MI_ERR_TOO_LARGE 104 /* the input exceeds a fixed scratch size */

/* ---- reference entry points (same names / signatures) ------------------- */

/* replaces c_kernels.h:54-55 (implementation c_kernels.c:18-60).  Solves
 * a[k-1] x[k-1] + b[k] x[k] + c[k] x[k+1] = d[k], k = 0..n-1, with a and c of
 * length n-1 (NULL = zeros), ws of length n.  The reference implementation
 * deviates from this contract (wrong c' index in the forward sweep); this one
 * implements the contract with the textbook Thomas recurrence. */
int vibe_tridiagonal_solve_f64(const double *a, const double *b, const double *c,
                               const double *d, int n, double *x, double *ws);

/* replaces c_kernels.h:64-65 (c_kernels.c:64-104).  q_left = left-biased
 * value at face i+1/2; q_right = mirrored reconstruction (right-biased value
 * at face i-1/2), exactly as the reference computes it. */
void vibe_weno5_reconstruct_f64(double qm2, double qm1, double q0, double q1,
                                double q2, double *q_left, double *q_right);

/* replaces c_kernels.h:74-75 (c_kernels.c:108-144). */
double vibe_bilinear_f64(const double *f, int halo, int nsx, int nsy, int nz,
                         double xi, double yj, int k);

/* replaces c_kernels.h:86-89 (c_kernels.c:148-187). */
double vibe_cfl_scan_f64(const double *u, const double *v, const double *w,
                         const double *dzeta, const double *zs, int nx, int ny,
                         int nz, double dx, double dy, double z_top, double dt,
                         double *max_wind_out, double *min_dz_out);

/* replaces c_kernels.h:97 (c_kernels.c:190-206).  Compensated sum; computed
 * as a fixed-order two-level Neumaier/TwoSum reduction, so the result is
 * deterministic and within a few ulp of the exact sum. */
double vibe_neumaier_sum_f64(const double *x, size_t n);

/* replaces c_kernels.h:109-113 (c_kernels.c:209-286). */
void vibe_advect_scalar_flux_f64(const double *q, const double *rho,
                                 const double *u, const double *v, const double *w,
                                 const double *dzeta, const double *zs, int nx,
                                 int ny, int nz, double dx, double dy,
                                 double z_top, double *dqdt);

/* ---- single-precision twins (c_kernels.h:15-16 naming convention) ------- */
int vibe_tridiagonal_solve_f32(const float *a, const float *b, const float *c,
                               const float *d, int n, float *x, float *ws);
void vibe_weno5_reconstruct_f32(float qm2, float qm1, float q0, float q1,
                                float q2, float *q_left, float *q_right);
double vibe_neumaier_sum_f32(const float *x, size_t n);

/* ---- batched device entry points (no counterpart in the reference) ------ */

/* nsys independent systems of size n, interleaved k-major: element k of
 * system s lives at [k*nsys + s] (a, c: k = 0..n-2; b, d, x: k = 0..n-1).
 * ws (device, n*nsys) is the caller's workspace, as in the reference's
 * contract: rows 0..n-2 hold the forward sweep's c', row n-1 the singular
 * flag.  One system per GPU lane.  Returns VIBE_C_ERR_SINGULAR if any pivot
 * of any system is exactly zero (the other systems are still solved). */
int vibe_mi_tridiagonal_solve_batched_f64(const double *a, const double *b,
                                          const double *c, const double *d,
                                          int n, int nsys, double *x, double *ws,
                                          void *stream);
int vibe_mi_tridiagonal_solve_batched_f32(const float *a, const float *b,
                                          const float *c, const float *d,
                                          int n, int nsys, float *x, float *ws,
                                          void *stream);

/* No counterpart in the reference: the same solve for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It takes no workspace and calls no hipMalloc/hipFree:
 * the singular flag and the forward-sweep workspace live in a statically
 * sized device array of the library (32 MiB per GPU, held until the process
 * ends), so n * nsys may be at most vibe_mi_noalloc_scratch_capacity()
 * (4194272); larger inputs return VIBE_MI_ERR_TOO_LARGE without touching the
 * GPU.  Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other threads, other streams) race and must be serialised by the
 * caller.  The flag is cleared by a kernel at the start of every call.
 * stream NULL: the null stream.  Results equal
 * vibe_mi_tridiagonal_solve_batched_f64 bit for bit. */
int vibe_mi_tridiagonal_solve_batched_noalloc_f64(const double *a, const double *b,
                                                  const double *c, const double *d,
                                                  int n, int nsys, double *x,
                                                  void *stream);
/* n * nsys limit of vibe_mi_tridiagonal_solve_batched_noalloc_f64. */
size_t vibe_mi_noalloc_scratch_capacity(void);

/* q: n reconstruction stencils, q[5*s + 0..4] = (qm2, qm1, q0, q1, q2). */
int vibe_mi_weno5_reconstruct_batched_f64(const double *q, size_t n,
                                          double *q_left, double *q_right,
                                          void *stream);

/* Device-pointer Neumaier sum; result written to *out (host).  Uses the
 * staging array's head for partial sums (under the per-GPU lock); the same
 * fixed-order reduction as vibe_neumaier_sum_f64, so both give the same bits
 * for the same values.  n <= 2.5e10. */
int vibe_mi_neumaier_sum_device_f64(const double *x, size_t n, double *out,
                                    void *stream);

/* Last device-side error of this thread (0 if none). */
int vibe_mi_last_error(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VIBE_MI_C_KERNELS_H */
