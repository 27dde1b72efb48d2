// domain.hpp -- host-side state of one domain (shared by the single-domain
// and the batched paths).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"
#include "geo.hpp"
#include "vibe_mi/dycore.h"
#include "vibe_mi/nest.h"

namespace vibe_mi {

// FP64 prognostic state of one domain (device pointers).
struct FieldSet {
  double *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *th = nullptr, *pi = nullptr;
};
template <typename T>
struct AcousticSet {
  T *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *pi = nullptr;
};
// bf16 remainders of an FP32 acoustic set (compensated FP32-mixed mode)
// Remainders of an FP32 acoustic set (compensated mode): 8-bit for u, bf16
// for v, w and pi'; rho' (carried as chi) has none.
struct AcousticLo {
  int8_t *u = nullptr;
  uint16_t *v = nullptr, *w = nullptr, *pi = nullptr;
};


}  // namespace vibe_mi

struct vibe_mi_domain {
  using FieldSet = vibe_mi::FieldSet;
  template <typename T>
  using AcousticSet = vibe_mi::AcousticSet<T>;
  int nx = 0, ny = 0, nz = 0;
  size_t nxy = 0, ncell = 0, nwf = 0;
  double dx = 0, dy = 0, ztop = 0, eps_s = 0;
  int precision = VIBE_MI_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  size_t bytes = 0;
  // FP64 geometry / reference for diagnostics and the FP64 path
  vibe_mi::Geo<double> g64{};
  vibe_mi::Geo<float> g32{};
  double *zs = nullptr, *sponge = nullptr, *rho0 = nullptr, *theta0 = nullptr, *pi0 = nullptr;
  double *dzeta64 = nullptr, *zc64 = nullptr;  // diagnostics (FP64, per level)
  double *ubg64 = nullptr;                     // FP32-mixed background x-wind per level
  float *ubg32 = nullptr;
  std::vector<double> h_sponge;
  int kmin_sponge = 0;
  FieldSet S[3];  // [0] = s0 of the step; [1], [2] stage outputs
  int cur0 = 0;   // index of s^n
  AcousticSet<double> A64[2];
  AcousticSet<float> A32[2];
  vibe_mi::AcousticLo L32[2];  // compensated mode: remainders of A32
  void *tab = nullptr;  // HEVI tables: 4 slots of hevi_slot_elems(nz, rk) entries (T)
  double tab_dtau[4] = {-1.0, -1.0, -1.0, -1.0};
  int nxr = 1, nyr = 1, rk = 1;
  int nw_fp32 = 4, nw_fp64 = 8;  // waves per substep workgroup (autotuned)
  bool tuned[2] = {false, false};
  int *flag = nullptr;
  // physics coupling: device rate buffers (theta', u, v; allocated on first
  // use), whether the stages add them, and the optional per-stage callback
  double *ph_th = nullptr, *ph_u = nullptr, *ph_v = nullptr;
  bool ph_on = false;
  bool ph_enabled = true;  // StepContext::physics_on: coupling of the next steps
  vibe_mi_physics_fn ph_fn = nullptr;
  void *ph_user = nullptr;
  int ph_host = 0;
  std::vector<double> ph_state, ph_rates;  // host staging of a host-mode callback (allocated once)
  vibe_mi::NSum *part = nullptr;
  double *dres = nullptr;
  unsigned long long *dmax = nullptr;
  // timing
  std::vector<hipEvent_t> ev;
  double ms_stage = 0, ms_sub = 0;
  int n_stage = 0, n_sub = 0;
};


namespace vibe_mi {

// Helpers of the single-domain path used by the batched path (dycore.hip).
int domain_create(const vibe_mi_domain_desc *ds, vibe_mi_domain **out);
// Nest interpolation (full = 1) or Davies relaxation of `child` toward
// `parent` enqueued on `stream`, with no host synchronisation (nest.hip).
int nest_enqueue(vibe_mi_domain *child, const vibe_mi_domain *parent, const vibe_mi_nest_desc *nd, double frac,
                 int full, hipStream_t stream);
// Enqueues the child -> parent feedback (nest.h) on `stream`.
int nest_restrict_enqueue(const vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd,
                          hipStream_t stream);
// Factorise the HEVI matrix of table slot `slot` for dtau (cached per slot).
int domain_ensure_table(vibe_mi_domain *d, int slot, double dtau);
// Base of table slot `slot` (layout of HeviTab, entries of type T).
void *domain_table(vibe_mi_domain *d, int slot);

}  // namespace vibe_mi
