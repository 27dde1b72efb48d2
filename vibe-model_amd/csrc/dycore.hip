// dycore.hip -- single-domain host path and C ABI of the dynamical-core step.
//
// Device code (kernels and their numerics) is in kernels.hpp; this file owns
// the domain handle: geometry upload, buffer rotation, the RK3 driver, the
// substep autotuner and the C entry points of vibe_mi/dycore.h.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "domain.hpp"
#include "kernels.hpp"
#include "vibe_mi/dycore.h"

using namespace vibe_mi;

// ---------------------------------------------------------------------------
// domain handle
// ---------------------------------------------------------------------------

namespace {

template <typename P>
int dalloc(vibe_mi_domain *d, P **out, size_t count) {
  void *p = nullptr;
  const size_t b = count * sizeof(P);
  VMI_HIP(hipMalloc(&p, b > 0 ? b : 1));
  d->allocs.push_back(p);
  d->bytes += b;
  *out = static_cast<P *>(p);
  return 0;
}

template <typename P>
int upload(vibe_mi_domain *d, P **out, const std::vector<P> &h) {
  if (int rc = dalloc(d, out, h.size())) return rc;
  VMI_HIP(hipMemcpy(*out, h.data(), h.size() * sizeof(P), hipMemcpyHostToDevice));
  return 0;
}

// Host-side derived geometry in FP64 (the same expressions as the oracle),
// compacted over directions in which the domain is uniform, then converted to
// the compute precision T.
template <typename T>
int build_geo(vibe_mi_domain *d, const vibe_mi_domain_desc *ds, Geo<T> &g) {
  const int nx = ds->nx, ny = ds->ny, nz = ds->nz;
  const int nxr = d->nxr, nyr = d->nyr, rk = d->rk;
  std::vector<double> zeta(ds->zeta, ds->zeta + nz + 1), zc(nz), rdzeta(nz), rdzw(nz + 1, 0.0);
  for (int k = 0; k < nz; ++k) {
    zc[k] = 0.5 * (zeta[k] + zeta[k + 1]);
    rdzeta[k] = 1.0 / (zeta[k + 1] - zeta[k]);
  }
  for (int k = 1; k < nz; ++k) rdzw[k] = 1.0 / (zc[k] - zc[k - 1]);
  auto H = [&](int i, int j) { return static_cast<size_t>((j + ny) % ny) * nx + (i + nx) % nx; };
  std::vector<double> Du(rk), Dv(rk), sx(rk), sy(rk), rDc(rk), rDu(rk), rDv(rk), metx(rk), mety(rk);
  for (int jr = 0; jr < nyr; ++jr)
    for (int ir = 0; ir < nxr; ++ir) {
      const size_t q = static_cast<size_t>(jr) * nxr + ir;
      const double z = ds->zs[H(ir, jr)], zw = ds->zs[H(ir - 1, jr)], zsth = ds->zs[H(ir, jr - 1)];
      const double dc = ds->ztop - z;
      Du[q] = ds->ztop - 0.5 * (zw + z);
      Dv[q] = ds->ztop - 0.5 * (zsth + z);
      sx[q] = (z - zw) / ds->dx;
      sy[q] = (z - zsth) / ds->dy;
      rDc[q] = 1.0 / dc;
      rDu[q] = 1.0 / Du[q];
      rDv[q] = 1.0 / Dv[q];
      metx[q] = sx[q] * rDu[q];
      mety[q] = sy[q] * rDv[q];
    }
  const size_t nr = static_cast<size_t>(rk) * nz;
  std::vector<double> r0(nr), t0(nr), kap(nr), lam(nr);
  for (int k = 0; k < nz; ++k)
    for (int jr = 0; jr < nyr; ++jr)
      for (int ir = 0; ir < nxr; ++ir) {
        const size_t q = static_cast<size_t>(k) * rk + static_cast<size_t>(jr) * nxr + ir;
        const size_t c = static_cast<size_t>(k) * d->nxy + H(ir, jr);
        r0[q] = ds->rho0[c];
        t0[q] = ds->theta0[c];
        kap[q] = (kRd / kCv) * ds->pi0[c] / ds->rho0[c];
        lam[q] = (kRd / kCv) * ds->pi0[c] / ds->theta0[c];
      }
  // face / interface products of the reference state (same association as
  // the oracle's expressions: cp * ((t_l + t_r) / 2) and D * ((r_l + r_r) / 2))
  std::vector<double> cthu(nr), cthv(nr), rhu(nr), rhv(nr), r0wv(static_cast<size_t>(rk) * (nz + 1), 0.0);
  {
    auto cidx = [&](int ir, int jr) { return static_cast<size_t>((jr + nyr) % nyr) * nxr + (ir + nxr) % nxr; };
    for (int k = 0; k < nz; ++k)
      for (int jr = 0; jr < nyr; ++jr)
        for (int ir = 0; ir < nxr; ++ir) {
          const size_t q = static_cast<size_t>(k) * rk + cidx(ir, jr);
          const size_t qw = static_cast<size_t>(k) * rk + cidx(ir - 1, jr);
          const size_t qs = static_cast<size_t>(k) * rk + cidx(ir, jr - 1);
          const size_t col = cidx(ir, jr);
          cthu[q] = kCp * (0.5 * (t0[qw] + t0[q]));
          cthv[q] = kCp * (0.5 * (t0[qs] + t0[q]));
          rhu[q] = Du[col] * (0.5 * (r0[qw] + r0[q]));
          rhv[q] = Dv[col] * (0.5 * (r0[qs] + r0[q]));
          if (k >= 1) r0wv[q] = 0.5 * (r0[q - rk] + r0[q]);
        }
  }
  std::vector<double> drhu(sizeof(T) == 4 ? nr : 0), drhv(sizeof(T) == 4 ? nr : 0);
  if (sizeof(T) == 4) {
    for (int k = 0; k < nz; ++k)
      for (int jr = 0; jr < nyr; ++jr)
        for (int ir = 0; ir < nxr; ++ir) {
          const size_t q = static_cast<size_t>(k) * rk + static_cast<size_t>(jr) * nxr + ir;
          drhu[q] = rhu[static_cast<size_t>(k) * rk + static_cast<size_t>(jr) * nxr + (ir + 1) % nxr] - rhu[q];
          drhv[q] = rhv[static_cast<size_t>(k) * rk + static_cast<size_t>((jr + 1) % nyr) * nxr + ir] - rhv[q];
        }
  }
  auto cvt = [](const std::vector<double> &a) { return std::vector<T>(a.begin(), a.end()); };
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.hi = nxr > 1 ? 1 : 0;
  g.hj = nyr > 1 ? nxr : 0;
  g.rk = rk;
  g.rdx = static_cast<T>(1.0 / ds->dx);
  g.rdy = static_cast<T>(1.0 / ds->dy);
  T *p;
#define VMI_UP(field, src)                            \
  if (int rc = upload(d, &p, cvt(src))) return rc;    \
  g.field = p;
  VMI_UP(zeta, zeta)
  VMI_UP(zc, zc)
  VMI_UP(rdzeta, rdzeta)
  VMI_UP(rdzw, rdzw)
  VMI_UP(Du, Du)
  VMI_UP(Dv, Dv)
  VMI_UP(sx, sx)
  VMI_UP(sy, sy)
  VMI_UP(rDc, rDc)
  VMI_UP(rDu, rDu)
  VMI_UP(rDv, rDv)
  VMI_UP(metx, metx)
  VMI_UP(mety, mety)
  VMI_UP(rho0, r0)
  VMI_UP(theta0, t0)
  VMI_UP(kap, kap)
  VMI_UP(lam, lam)
  VMI_UP(cthu, cthu)
  VMI_UP(cthv, cthv)
  VMI_UP(rhu, rhu)
  VMI_UP(rhv, rhv)
  VMI_UP(r0w, r0wv)
  if constexpr (sizeof(T) == 4) {  // only the FP32 substep reads the differences
    VMI_UP(drhu, drhu)
    VMI_UP(drhv, drhv)
  } else {
    g.drhu = g.drhv = nullptr;
  }
#undef VMI_UP
  return 0;
}

// Directions in which terrain and reference state are uniform.
void detect_uniform(const vibe_mi_domain_desc *ds, bool &xuni, bool &yuni) {
  const int nx = ds->nx, ny = ds->ny, nz = ds->nz;
  const size_t nxy = static_cast<size_t>(nx) * ny;
  xuni = yuni = true;
  const double *f3[3] = {ds->rho0, ds->theta0, ds->pi0};
  for (int j = 0; j < ny && (xuni || yuni); ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t c = static_cast<size_t>(j) * nx + i;
      if (ds->zs[c] != ds->zs[static_cast<size_t>(j) * nx]) xuni = false;
      if (ds->zs[c] != ds->zs[i]) yuni = false;
    }
  for (int k = 0; k < nz && (xuni || yuni); ++k)
    for (const double *f : f3) {
      const double *pl = f + static_cast<size_t>(k) * nxy;
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const double x = pl[static_cast<size_t>(j) * nx + i];
          if (xuni && x != pl[static_cast<size_t>(j) * nx]) xuni = false;
          if (yuni && x != pl[i]) yuni = false;
        }
    }
}

inline dim3 grid_of(const vibe_mi_domain *d, int nk) {
  return dim3(ceil_div(d->nx, 64), d->ny, nk);
}
constexpr unsigned kBlk = 64;

template <typename T>
int run_stage(vibe_mi_domain *d, const Geo<T> &g, const FieldSet &cur, const FieldSet &s0, double h,
              AcousticSet<T> &a, FieldSet &nxt) {
  const StageIn ci{cur.u, cur.v, cur.w, cur.rho, cur.th, cur.pi};
  const StageIn si{s0.u, s0.v, s0.w, s0.rho, s0.th, s0.pi};
  StageOut<T> so{a.u, a.v, a.w, a.rho, a.pi, nxt.th};
  if (sizeof(T) == 4 && d->precision == VIBE_MI_MIXED_COMP) {
    const AcousticLo &l = d->L32[0];
    so.ul = l.u;
    so.vl = cp_keeps(2) ? l.v : nullptr, so.wl = cp_keeps(4) ? l.w : nullptr, so.pl = cp_keeps(8) ? l.pi : nullptr;
  }
  const dim3 grid(ceil_div(d->nx, 64), ceil_div(d->ny, 4)), block(64, 4);
  if (d->ph_on && d->ph_enabled)
    hipLaunchKernelGGL((k_stage_col<T, T, 4, true>), grid, block, 0, d->stream, g, ci, si, h, so, d->flag,
                       PhysArgs{d->ph_th, d->ph_u, d->ph_v});
  else
    hipLaunchKernelGGL((k_stage_col<T, T, 4, false>), grid, block, 0, d->stream, g, ci, si, h, so, d->flag,
                       PhysArgs{});
  VMI_HIP(hipGetLastError());
  return 0;
}

// Allocates (once) and zeroes the physics rate buffers.
int ensure_physics(vibe_mi_domain *d) {
  if (d->ph_th == nullptr) {
    if (int rc = dalloc(d, &d->ph_th, d->ncell)) return rc;
    if (int rc = dalloc(d, &d->ph_u, d->ncell)) return rc;
    if (int rc = dalloc(d, &d->ph_v, d->ncell)) return rc;
    for (double *p : {d->ph_th, d->ph_u, d->ph_v}) VMI_HIP(hipMemsetAsync(p, 0, d->ncell * 8, d->stream));
    VMI_HIP(hipStreamSynchronize(d->stream));
  }
  return 0;
}

// The per-stage physics call (runge_kutta.cpp:46-63 role): zero the rates,
// hand fn the stage state (host copies or device pointers), keep its rates.
int call_physics(vibe_mi_domain *d, int stage, double dt, const FieldSet &s) {
  const size_t nc = d->ncell * 8;
  for (double *p : {d->ph_th, d->ph_u, d->ph_v}) VMI_HIP(hipMemsetAsync(p, 0, nc, d->stream));
  int rc = 0;
  if (d->ph_host) {
    std::vector<double> &h = d->ph_state, &r = d->ph_rates;
    h.resize(5 * d->ncell + d->nwf);
    r.assign(3 * d->ncell, 0.0);
    double *hu = h.data(), *hv = hu + d->ncell, *hr = hv + d->ncell, *ht = hr + d->ncell, *hp = ht + d->ncell,
           *hw = hp + d->ncell;
    VMI_HIP(hipStreamSynchronize(d->stream));
    VMI_HIP(hipMemcpy(hu, s.u, nc, hipMemcpyDeviceToHost));
    VMI_HIP(hipMemcpy(hv, s.v, nc, hipMemcpyDeviceToHost));
    VMI_HIP(hipMemcpy(hw, s.w, d->nwf * 8, hipMemcpyDeviceToHost));
    VMI_HIP(hipMemcpy(hr, s.rho, nc, hipMemcpyDeviceToHost));
    VMI_HIP(hipMemcpy(ht, s.th, nc, hipMemcpyDeviceToHost));
    VMI_HIP(hipMemcpy(hp, s.pi, nc, hipMemcpyDeviceToHost));
    double *rt = r.data(), *ru = rt + d->ncell, *rv = ru + d->ncell;
    rc = d->ph_fn(d->ph_user, stage, dt, d->stream, hu, hv, hw, hr, ht, hp, rt, ru, rv);
    if (rc == 0) {
      VMI_HIP(hipMemcpy(d->ph_th, rt, nc, hipMemcpyHostToDevice));
      VMI_HIP(hipMemcpy(d->ph_u, ru, nc, hipMemcpyHostToDevice));
      VMI_HIP(hipMemcpy(d->ph_v, rv, nc, hipMemcpyHostToDevice));
    }
  } else {
    rc = d->ph_fn(d->ph_user, stage, dt, d->stream, s.u, s.v, s.w, s.rho, s.th, s.pi, d->ph_th, d->ph_u, d->ph_v);
  }
  return rc == 0 ? 0 : VIBE_MI_ERR_PHYSICS;
}

template <typename T>
T *table_slot(vibe_mi_domain *d, int slot) {
  return static_cast<T *>(d->tab) + static_cast<size_t>(slot) * hevi_slot_elems(d->nz, d->rk);
}

// Factorise the HEVI matrix for dtau into table slot `slot` (cached).
template <typename T>
int ensure_table(vibe_mi_domain *d, const Geo<T> &g, int slot, double dtau) {
  if (d->tab_dtau[slot] == dtau) return 0;
  T *t = table_slot<T>(d, slot);
  const T an = static_cast<T>(0.5 * (1.0 + d->eps_s));
  hipLaunchKernelGGL((k_hevi_factor<T>), dim3(ceil_div(d->rk, 64)), dim3(64), 0, d->stream, g, static_cast<T>(dtau),
                     an, t);
  VMI_HIP(hipGetLastError());
  d->tab_dtau[slot] = dtau;
  return 0;
}

constexpr size_t kMaxLds = 160 * 1024;  // LDS per CU (and per workgroup) on gfx950

template <typename T, typename I, typename O, int NW, int MAXCH, bool CP>
int launch_substep_k(vibe_mi_domain *d, const Geo<T> &g, const SubArgs<T, I, O> &sa) {
  const size_t lds = substep_lds_arrays(CP) * static_cast<size_t>(d->nz - 1) * 64 * sizeof(T);
  // s_mc, s_fh (FP64 in the compensated mode) and s_ph
  const size_t lds_static = ((CP ? 2 * sizeof(double) : 2 * sizeof(T)) + sizeof(T)) * NW * 64;
  if (lds + lds_static > kMaxLds) return VIBE_C_ERR_BAD_ARG;  // column too tall for one workgroup
  static bool attr_set = false;  // once per instantiation
  if (!attr_set) {
    VMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_substep<T, I, O, NW, MAXCH, CP>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds - lds_static)));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_substep<T, I, O, NW, MAXCH, CP>), dim3(ceil_div(d->nx, 64), d->ny), dim3(64 * NW), lds,
                     d->stream, g, sa);
  return 0;
}

// li / lo: remainders of the input and (FP32 output) output set in the
// compensated mode, nullptr otherwise.
template <typename T, typename I, typename O, int NW, int MAXCH>
int launch_substep(vibe_mi_domain *d, const Geo<T> &g, const T *tab, const I *u, const I *v, const I *w,
                   const I *rho, const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, T dtau, T an, T ao,
                   int check, const AcousticLo *li, const AcousticLo *lo) {
  SubArgs<T, I, O> sa{u, v, w, rho, pi, HeviTab<T>::at(tab, d->nz, d->rk), dtau, an, ao,
                      uo, vo, wo, rho_o, pi_o, check, d->flag};
  if constexpr (sizeof(I) == 4) {
    // FP32 substeps carry chi = rho' - pi'/kappa, which the stage writes to
    // A[0]; the last substep forms rho'
    sa.chi = d->A32[0].rho;
  }
  if constexpr (sizeof(T) == 4 && sizeof(I) == 4) {
    if (li != nullptr) {
      sa.ul = li->u, sa.vl = li->v, sa.wl = li->w, sa.pl = li->pi;
      if (lo != nullptr) sa.ulo = lo->u, sa.vlo = lo->v, sa.wlo = lo->w, sa.plo = lo->pi;
      const Geo<double> &g8 = d->g64;
      sa.rhu64 = g8.rhu, sa.rhv64 = g8.rhv, sa.r0w64 = g8.r0w, sa.sx64 = g8.sx, sa.sy64 = g8.sy, sa.zeta64 = g8.zeta;
      sa.rdx64 = g8.rdx, sa.rdy64 = g8.rdy;
      return launch_substep_k<T, I, O, NW, MAXCH, true>(d, g, sa);
    }
  }
  return launch_substep_k<T, I, O, NW, MAXCH, false>(d, g, sa);
}

template <typename T, typename I, typename O, int NW>
int dispatch_substep(vibe_mi_domain *d, const Geo<T> &g, const T *tab, const I *u, const I *v, const I *w,
                     const I *rho, const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, T dtau, T an, T ao, int check,
                     const AcousticLo *li, const AcousticLo *lo) {
  const int ch = (d->nz + NW - 1) / NW;
#define VMI_SUB(M) \
  launch_substep<T, I, O, NW, M>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check, li, lo)
  int rc;
  if constexpr (NW == 16) {
    if (ch <= 4) rc = VMI_SUB(4);
    else if (ch <= 8) rc = VMI_SUB(8);
    else return VIBE_C_ERR_BAD_ARG;
  } else if constexpr (NW == 8) {
    if (ch <= 4) rc = VMI_SUB(4);
    else if (ch <= 8) rc = VMI_SUB(8);
    else if (ch <= 12) rc = VMI_SUB(12);
    else if (ch <= 16) rc = VMI_SUB(16);
    else return VIBE_C_ERR_BAD_ARG;
  } else {
    if (ch <= 8) rc = VMI_SUB(8);
    else if (ch <= 12) rc = VMI_SUB(12);
    else if (ch <= 16) rc = VMI_SUB(16);
    else if (ch <= 24) rc = VMI_SUB(24);
    else if (ch <= 32) rc = VMI_SUB(32);
    else return VIBE_C_ERR_BAD_ARG;
  }
#undef VMI_SUB
  if (rc) return rc;
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T, typename I, typename O>
int substep_nw(vibe_mi_domain *d, const Geo<T> &g, int nw, const T *tab, const I *u, const I *v, const I *w,
               const I *rho, const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, T dtau, T an, T ao, int check,
               const AcousticLo *li = nullptr, const AcousticLo *lo = nullptr) {
  if (nw == 16)
    return dispatch_substep<T, I, O, 16>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check,
                                         li, lo);
  if (nw == 8)
    return dispatch_substep<T, I, O, 8>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check,
                                        li, lo);
  return dispatch_substep<T, I, O, 4>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check, li,
                                      lo);
}

// ---- substep workgroup autotuner --------------------------------------------
// Waves per substep workgroup (4, 8 or 16: levels per wave = nz/waves) trade
// registers per wave, LDS per workgroup and the length of each wave's serial
// march against the idle time of the other waves during the Thomas sweep.  The
// fastest is measured once per (nx, ny, nz, compaction, precision) with HIP
// events on scratch buffers (the state is not touched) and cached for the
// process.  VIBE_MI_SUBSTEP_WAVES=4|8|16 overrides the measurement.
struct TuneKey {
  int nx, ny, nz, rk, prec;
  bool operator<(const TuneKey &o) const {
    return std::tie(nx, ny, nz, rk, prec) < std::tie(o.nx, o.ny, o.nz, o.rk, o.prec);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;

template <typename T>
int autotune_substep(vibe_mi_domain *d, const Geo<T> &g, AcousticSet<T> *A, double dtau) {
  int &slot_nw = sizeof(T) == 8 ? d->nw_fp64 : d->nw_fp32;
  if (d->tuned[sizeof(T) == 8]) return 0;
  if (const char *env = std::getenv("VIBE_MI_SUBSTEP_WAVES")) {
    const int v = std::atoi(env);
    if (v == 4 || v == 8 || v == 16) {
      slot_nw = v;
      d->tuned[sizeof(T) == 8] = true;
      return 0;
    }
  }
  const bool comp = sizeof(T) == 4 && d->precision == VIBE_MI_MIXED_COMP;
  const TuneKey key{d->nx, d->ny, d->nz, d->rk, static_cast<int>(sizeof(T)) + (comp ? 16 : 0)};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      slot_nw = it->second;
      d->tuned[sizeof(T) == 8] = true;
      return 0;
    }
  }
  if (int rc = ensure_table<T>(d, g, 3, dtau)) return rc;
  const T an = static_cast<T>(0.5 * (1.0 + d->eps_s)), ao = static_cast<T>(0.5 * (1.0 - d->eps_s));
  const T *tab = table_slot<T>(d, 3);
  AcousticSet<T> &in = A[0], &out = A[1];
  float best = 1e30f;
  int best_nw = slot_nw;
  for (int nw : {4, 8, 16}) {
    float total = 0.f;
    bool ok = true;
    for (int rep = 0; rep < 4 && ok; ++rep) {
      VMI_HIP(hipEventRecord(d->ev[0], d->stream));
      ok = substep_nw<T, T, T>(d, g, nw, tab, in.u, in.v, in.w, in.rho, in.pi, out.u, out.v, out.w, out.rho, out.pi,
                               static_cast<T>(dtau), an, ao, 0, comp ? &d->L32[0] : nullptr,
                               comp ? &d->L32[1] : nullptr) == 0;
      VMI_HIP(hipEventRecord(d->ev[1], d->stream));
      VMI_HIP(hipEventSynchronize(d->ev[1]));
      float ms = 0.f;
      VMI_HIP(hipEventElapsedTime(&ms, d->ev[0], d->ev[1]));
      if (rep > 0) total += ms;  // first launch warms the instruction cache
    }
    if (ok && total < best) {
      best = total;
      best_nw = nw;
    }
  }
  (void)hipGetLastError();  // a shape that does not fit is simply skipped
  slot_nw = best_nw;
  d->tuned[sizeof(T) == 8] = true;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = best_nw;
  return 0;
}

template <typename T, typename I, typename O>
int run_substep(vibe_mi_domain *d, const Geo<T> &g, int slot, const I *u, const I *v, const I *w, const I *rho,
                const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, double dtau, int check,
                const AcousticLo *li = nullptr, const AcousticLo *lo = nullptr) {
  if (int rc = ensure_table<T>(d, g, slot, dtau)) return rc;
  const T an = static_cast<T>(0.5 * (1.0 + d->eps_s)), ao = static_cast<T>(0.5 * (1.0 - d->eps_s));
  const T *tab = table_slot<T>(d, slot);
  const int nw = sizeof(T) == 8 ? d->nw_fp64 : d->nw_fp32;
  return substep_nw<T, I, O>(d, g, nw, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, static_cast<T>(dtau), an, ao,
                             check, li, lo);
}

// One full RK3 step.  Stage outputs go to S[(cur0+1)%3] / S[(cur0+2)%3].
template <typename T>
int step_impl(vibe_mi_domain *d, const Geo<T> &g, AcousticSet<T> *A, double dt, int nsub, bool timed) {
  const int ns = nsub < 1 ? 1 : nsub;
  const double hk[3] = {dt / 3.0, dt / 2.0, dt};
  const bool comp = sizeof(T) == 4 && d->precision == VIBE_MI_MIXED_COMP;
  if (int rc = autotune_substep<T>(d, g, A, dt / static_cast<double>(ns))) return rc;
  const int i0 = d->cur0, i1 = (d->cur0 + 1) % 3, i2 = (d->cur0 + 2) % 3;
  const FieldSet &s0 = d->S[i0];
  const int outs[3] = {i1, i2, i1};
  int cur = i0;
  for (int st = 0; st < 3; ++st) {
    FieldSet &nxt = d->S[outs[st]];
    if (d->ph_fn != nullptr && d->ph_enabled)
      if (int rc = call_physics(d, st, dt, d->S[cur])) return rc;
    if (timed) VMI_HIP(hipEventRecord(d->ev[0], d->stream));
    if (int rc = run_stage<T>(d, g, d->S[cur], s0, hk[st], A[0], nxt)) return rc;
    if (timed) {
      VMI_HIP(hipEventRecord(d->ev[1], d->stream));
    }
    const double dtau = hk[st] / static_cast<double>(ns);
    int p = 0;
    for (int q = 0; q < ns; ++q) {
      const bool last = (q == ns - 1);
      const AcousticSet<T> &in = A[p];
      if (!last) {
        AcousticSet<T> &out = A[1 - p];
        if (int rc = run_substep<T, T, T>(d, g, st, in.u, in.v, in.w, in.rho, in.pi, out.u, out.v, out.w, out.rho,
                                          out.pi, dtau, 0, comp ? &d->L32[p] : nullptr,
                                          comp ? &d->L32[1 - p] : nullptr))
          return rc;
        p = 1 - p;
      } else {
        if (int rc = run_substep<T, T, double>(d, g, st, in.u, in.v, in.w, in.rho, in.pi, nxt.u, nxt.v, nxt.w,
                                               nxt.rho, nxt.pi, dtau, 1, comp ? &d->L32[p] : nullptr, nullptr))
          return rc;
      }
    }
    if (timed) {
      VMI_HIP(hipEventRecord(d->ev[2], d->stream));
      VMI_HIP(hipEventSynchronize(d->ev[2]));
      float a = 0, b = 0;
      VMI_HIP(hipEventElapsedTime(&a, d->ev[0], d->ev[1]));
      VMI_HIP(hipEventElapsedTime(&b, d->ev[1], d->ev[2]));
      d->ms_stage += a;
      d->ms_sub += b;
      d->n_stage += 1;
      d->n_sub += ns;
    }
    cur = outs[st];
  }
  // sponge on the final state, then it becomes s^n of the next step
  FieldSet &fin = d->S[i1];
  const int nk = d->nz + 1 - d->kmin_sponge;
  if (d->kmin_sponge <= d->nz && nk > 0) {
    hipLaunchKernelGGL(k_sponge, grid_of(d, nk), dim3(kBlk), 0, d->stream, d->nx, d->ny, d->nz, d->kmin_sponge,
                       d->sponge, dt, fin);
    VMI_HIP(hipGetLastError());
  }
  d->cur0 = i1;
  return 0;
}

int alloc_fieldset(vibe_mi_domain *d, FieldSet &f) {
  if (int rc = dalloc(d, &f.u, d->ncell)) return rc;
  if (int rc = dalloc(d, &f.v, d->ncell)) return rc;
  if (int rc = dalloc(d, &f.w, d->nwf)) return rc;
  if (int rc = dalloc(d, &f.rho, d->ncell)) return rc;
  if (int rc = dalloc(d, &f.th, d->ncell)) return rc;
  if (int rc = dalloc(d, &f.pi, d->ncell)) return rc;
  return 0;
}

template <typename T>
int alloc_acoustic(vibe_mi_domain *d, AcousticSet<T> &a) {
  if (int rc = dalloc(d, &a.u, d->ncell)) return rc;
  if (int rc = dalloc(d, &a.v, d->ncell)) return rc;
  if (int rc = dalloc(d, &a.w, d->nwf)) return rc;
  if (int rc = dalloc(d, &a.rho, d->ncell)) return rc;
  if (int rc = dalloc(d, &a.pi, d->ncell)) return rc;
  return 0;
}

void free_all(vibe_mi_domain *d) {
  for (void *p : d->allocs) (void)hipFree(p);
  d->allocs.clear();
  for (auto &e : d->ev) (void)hipEventDestroy(e);
  d->ev.clear();
  if (d->stream) (void)hipStreamDestroy(d->stream);
  d->stream = nullptr;
}

int create_impl(const vibe_mi_domain_desc *ds, vibe_mi_domain *d) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return VIBE_MI_ERR_NO_DEVICE;
  VMI_HIP(hipGetDevice(&d->device));
  d->nx = ds->nx;
  d->ny = ds->ny;
  d->nz = ds->nz;
  d->nxy = static_cast<size_t>(ds->nx) * ds->ny;
  d->ncell = d->nxy * ds->nz;
  d->nwf = d->nxy * (ds->nz + 1);
  d->dx = ds->dx;
  d->dy = ds->dy;
  d->ztop = ds->ztop;
  d->eps_s = ds->eps_s;
  d->precision = ds->precision;
  VMI_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  d->ev.resize(4);
  for (auto &e : d->ev) VMI_HIP(hipEventCreate(&e));
  bool xuni = false, yuni = false;
  detect_uniform(ds, xuni, yuni);
  d->nxr = xuni ? 1 : d->nx;
  d->nyr = yuni ? 1 : d->ny;
  d->rk = d->nxr * d->nyr;
  if (int rc = build_geo<double>(d, ds, d->g64)) return rc;
  if (d->precision != VIBE_MI_FP64)
    if (int rc = build_geo<float>(d, ds, d->g32)) return rc;
  if (int rc = upload(d, &d->ubg64, std::vector<double>(ds->nz, 0.0))) return rc;
  if (int rc = upload(d, &d->ubg32, std::vector<float>(ds->nz, 0.0f))) return rc;
  d->g64.ubg = nullptr;
  d->g64.ubg64 = d->ubg64;
  d->g32.ubg = d->ubg32;
  d->g32.ubg64 = d->ubg64;
  const size_t nc = d->ncell;
  if (int rc = upload(d, &d->zs, std::vector<double>(ds->zs, ds->zs + d->nxy))) return rc;
  {
    std::vector<double> dz(ds->nz), zc(ds->nz);
    for (int k = 0; k < ds->nz; ++k) {
      dz[k] = ds->zeta[k + 1] - ds->zeta[k];
      zc[k] = 0.5 * (ds->zeta[k] + ds->zeta[k + 1]);
    }
    if (int rc = upload(d, &d->dzeta64, dz)) return rc;
    if (int rc = upload(d, &d->zc64, zc)) return rc;
  }
  d->h_sponge.assign(ds->sponge, ds->sponge + ds->nz);
  if (int rc = upload(d, &d->sponge, d->h_sponge)) return rc;
  d->kmin_sponge = d->nz + 1;
  for (int k = 0; k < d->nz; ++k)
    if (d->h_sponge[k] > 0.0) {
      d->kmin_sponge = k;
      break;
    }
  if (int rc = upload(d, &d->rho0, std::vector<double>(ds->rho0, ds->rho0 + nc))) return rc;
  if (int rc = upload(d, &d->theta0, std::vector<double>(ds->theta0, ds->theta0 + nc))) return rc;
  if (int rc = upload(d, &d->pi0, std::vector<double>(ds->pi0, ds->pi0 + nc))) return rc;
  for (auto &f : d->S)
    if (int rc = alloc_fieldset(d, f)) return rc;
  const size_t tsz = d->precision != VIBE_MI_FP64 ? sizeof(float) : sizeof(double);
  if (d->precision != VIBE_MI_FP64) {
    for (auto &a : d->A32)
      if (int rc = alloc_acoustic<float>(d, a)) return rc;
    if (d->precision == VIBE_MI_MIXED_COMP)
      for (auto &l : d->L32) {
        if (int rc = dalloc(d, &l.u, d->ncell)) return rc;
        if (int rc = dalloc(d, &l.v, d->ncell)) return rc;
        if (int rc = dalloc(d, &l.w, d->nwf)) return rc;
        if (int rc = dalloc(d, &l.pi, d->ncell)) return rc;
        VMI_HIP(hipMemset(l.u, 0, d->ncell));
        for (uint16_t *q : {l.v, l.pi}) VMI_HIP(hipMemset(q, 0, d->ncell * 2));
        VMI_HIP(hipMemset(l.w, 0, d->nwf * 2));
      }
  } else {
    for (auto &a : d->A64)
      if (int rc = alloc_acoustic<double>(d, a)) return rc;
  }
  char *p;
  if (int rc = dalloc(d, &p, 4 * hevi_slot_elems(d->nz, d->rk) * tsz)) return rc;
  d->tab = p;
  if (int rc = dalloc(d, &d->flag, 4)) return rc;
  if (int rc = dalloc(d, &d->part, static_cast<size_t>(kRedBlocks) * kEnergyComps)) return rc;
  if (int rc = dalloc(d, &d->dres, 8)) return rc;
  if (int rc = dalloc(d, &d->dmax, 8)) return rc;
  VMI_HIP(hipMemset(d->flag, 0, 4 * sizeof(int)));
  for (auto &f : d->S) {
    VMI_HIP(hipMemset(f.u, 0, nc * 8));
    VMI_HIP(hipMemset(f.v, 0, nc * 8));
    VMI_HIP(hipMemset(f.w, 0, d->nwf * 8));
    VMI_HIP(hipMemset(f.rho, 0, nc * 8));
    VMI_HIP(hipMemset(f.th, 0, nc * 8));
    VMI_HIP(hipMemset(f.pi, 0, nc * 8));
  }
  VMI_HIP(hipDeviceSynchronize());
  return 0;
}

int check_flag(vibe_mi_domain *d) {
  int f = 0;
  VMI_HIP(hipMemcpyAsync(&f, d->flag, sizeof(int), hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  return f ? VIBE_MI_ERR_NONFINITE : 0;
}

int set_device(vibe_mi_domain *d) {
  VMI_HIP(hipSetDevice(d->device));
  return 0;
}

}  // namespace

namespace vibe_mi {

int domain_create(const vibe_mi_domain_desc *ds, vibe_mi_domain **out) { return vibe_mi_domain_create(ds, out); }

int domain_ensure_table(vibe_mi_domain *d, int slot, double dtau) {
  if (d->precision != VIBE_MI_FP64) return ensure_table<float>(d, d->g32, slot, dtau);
  return ensure_table<double>(d, d->g64, slot, dtau);
}

void *domain_table(vibe_mi_domain *d, int slot) {
  if (d->precision != VIBE_MI_FP64) return table_slot<float>(d, slot);
  return table_slot<double>(d, slot);
}

}  // namespace vibe_mi

extern "C" {

int vibe_mi_domain_create(const vibe_mi_domain_desc *ds, vibe_mi_domain **out) {
  if (ds == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  *out = nullptr;
  // the tallest column one substep workgroup solves: its HEVI factorisation
  // (3 x (nz-1) x 64 entries) is staged in LDS next to 4 waves' buffers,
  // which caps FP64 columns at 103 levels; FP32 columns are capped at 128 by
  // the level chunks (at most 16 waves of 8 levels, 8 of 16 or 4 of 32)
  const int nz_max = ds->precision == VIBE_MI_FP64 ? 103 : 128;
  if (ds->nx < 4 || ds->ny < 4 || ds->nz < 3 || ds->nz > nz_max || ds->dx <= 0 || ds->dy <= 0 || ds->ztop <= 0 ||
      ds->zeta == nullptr || ds->zs == nullptr || ds->rho0 == nullptr || ds->theta0 == nullptr ||
      ds->pi0 == nullptr || ds->sponge == nullptr ||
      (ds->precision != VIBE_MI_FP64 && ds->precision != VIBE_MI_MIXED && ds->precision != VIBE_MI_MIXED_COMP))
    return VIBE_C_ERR_BAD_ARG;
  // kernels address a field array with 32-bit byte offsets: the largest one
  // (FP64 w, nz+1 levels) must stay below 4 GiB
  if (static_cast<uint64_t>(ds->nx) * static_cast<uint64_t>(ds->ny) * static_cast<uint64_t>(ds->nz + 1) * 8u >=
      (uint64_t{1} << 32))
    return VIBE_C_ERR_BAD_ARG;
  auto *d = new vibe_mi_domain();
  const int rc = create_impl(ds, d);
  if (rc != 0) {
    free_all(d);
    delete d;
    set_last_error(rc);
    return rc;
  }
  *out = d;
  return 0;
}

int vibe_mi_domain_destroy(vibe_mi_domain *d) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  (void)hipSetDevice(d->device);
  (void)hipStreamSynchronize(d->stream);
  free_all(d);
  delete d;
  return 0;
}

int vibe_mi_domain_set_state(vibe_mi_domain *d, const double *u, const double *v, const double *w,
                             const double *rho, const double *theta, const double *pi) {
  if (d == nullptr || !u || !v || !w || !rho || !theta || !pi) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  const FieldSet &f = d->S[d->cur0];
  const size_t nc = d->ncell * 8;
  if (d->precision != VIBE_MI_FP64) {  // background x-wind: level means of u
    std::vector<double> ub(d->nz);
    std::vector<float> ub32(d->nz);
    for (int k = 0; k < d->nz; ++k) {
      double acc = 0.0;
      const double *pl = u + static_cast<size_t>(k) * d->nxy;
      for (size_t c = 0; c < d->nxy; ++c) acc += pl[c];
      ub[k] = acc / static_cast<double>(d->nxy);
      ub32[k] = static_cast<float>(ub[k]);
      ub[k] = static_cast<double>(ub32[k]);  // the same value in both precisions
    }
    VMI_HIP(hipMemcpy(d->ubg64, ub.data(), ub.size() * sizeof(double), hipMemcpyHostToDevice));
    VMI_HIP(hipMemcpy(d->ubg32, ub32.data(), ub32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  VMI_HIP(hipMemcpyAsync(f.u, u, nc, hipMemcpyHostToDevice, d->stream));
  VMI_HIP(hipMemcpyAsync(f.v, v, nc, hipMemcpyHostToDevice, d->stream));
  VMI_HIP(hipMemcpyAsync(f.w, w, d->nwf * 8, hipMemcpyHostToDevice, d->stream));
  VMI_HIP(hipMemcpyAsync(f.rho, rho, nc, hipMemcpyHostToDevice, d->stream));
  VMI_HIP(hipMemcpyAsync(f.th, theta, nc, hipMemcpyHostToDevice, d->stream));
  VMI_HIP(hipMemcpyAsync(f.pi, pi, nc, hipMemcpyHostToDevice, d->stream));
  VMI_HIP(hipMemsetAsync(d->flag, 0, sizeof(int), d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

int vibe_mi_domain_get_state(vibe_mi_domain *d, double *u, double *v, double *w, double *rho, double *theta,
                             double *pi) {
  if (d == nullptr || !u || !v || !w || !rho || !theta || !pi) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  const FieldSet &f = d->S[d->cur0];
  const size_t nc = d->ncell * 8;
  VMI_HIP(hipMemcpyAsync(u, f.u, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(v, f.v, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(w, f.w, d->nwf * 8, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(rho, f.rho, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(theta, f.th, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(pi, f.pi, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

static int step_common(vibe_mi_domain *d, double dt, int nsub, int nsteps, bool timed, double *ms_total) {
  if (d == nullptr || nsteps < 0 || !(dt > 0.0)) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  if (timed) {
    d->ms_stage = d->ms_sub = 0;
    d->n_stage = d->n_sub = 0;
    VMI_HIP(hipEventRecord(d->ev[3], d->stream));
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (ms_total != nullptr) {
    VMI_HIP(hipEventCreate(&t0));
    VMI_HIP(hipEventCreate(&t1));
    VMI_HIP(hipEventRecord(t0, d->stream));
  }
  for (int n = 0; n < nsteps; ++n) {
    int rc;
    if (d->precision != VIBE_MI_FP64)
      rc = step_impl<float>(d, d->g32, d->A32, dt, nsub, timed);
    else
      rc = step_impl<double>(d, d->g64, d->A64, dt, nsub, timed);
    if (rc) return rc;
  }
  if (ms_total != nullptr) {
    VMI_HIP(hipEventRecord(t1, d->stream));
    VMI_HIP(hipEventSynchronize(t1));
    float ms = 0;
    VMI_HIP(hipEventElapsedTime(&ms, t0, t1));
    *ms_total = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  return check_flag(d);
}

int vibe_mi_domain_step(vibe_mi_domain *d, double dt, int nsub, int nsteps) {
  return step_common(d, dt, nsub, nsteps, false, nullptr);
}

int vibe_mi_domain_step_timed(vibe_mi_domain *d, double dt, int nsub, int nsteps, double *ms_total) {
  return step_common(d, dt, nsub, nsteps, true, ms_total);
}

int vibe_mi_domain_kernel_times(vibe_mi_domain *d, double *ms_stage, double *ms_substep, int *n_stage,
                                int *n_substep) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (ms_stage) *ms_stage = d->ms_stage;
  if (ms_substep) *ms_substep = d->ms_sub;
  if (n_stage) *n_stage = d->n_stage;
  if (n_substep) *n_substep = d->n_sub;
  return 0;
}

int vibe_mi_domain_slow_tendency(vibe_mi_domain *d, double *ru, double *rv, double *rw, double *rrho,
                                 double *rtheta, double *rpi) {
  if (d == nullptr || !ru || !rv || !rw || !rrho || !rtheta || !rpi) return VIBE_C_ERR_BAD_ARG;
  if (d->precision != VIBE_MI_FP64) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  // R = (s0 + 1*R) - s0 would round; instead evaluate with a zero base:
  // use S[next] as a zero field set and h = 1.
  const int iz = (d->cur0 + 2) % 3;
  FieldSet &z = d->S[iz];
  const size_t nc = d->ncell * 8;
  VMI_HIP(hipMemsetAsync(z.u, 0, nc, d->stream));
  VMI_HIP(hipMemsetAsync(z.v, 0, nc, d->stream));
  VMI_HIP(hipMemsetAsync(z.w, 0, d->nwf * 8, d->stream));
  VMI_HIP(hipMemsetAsync(z.rho, 0, nc, d->stream));
  VMI_HIP(hipMemsetAsync(z.pi, 0, nc, d->stream));
  VMI_HIP(hipMemsetAsync(z.th, 0, nc, d->stream));
  const int io = (d->cur0 + 1) % 3;
  FieldSet &o = d->S[io];
  AcousticSet<double> a{o.u, o.v, o.w, o.rho, o.pi};
  if (int rc = run_stage<double>(d, d->g64, d->S[d->cur0], z, 1.0, a, o)) return rc;
  VMI_HIP(hipMemcpyAsync(ru, o.u, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(rv, o.v, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(rw, o.w, d->nwf * 8, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(rrho, o.rho, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(rtheta, o.th, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipMemcpyAsync(rpi, o.pi, nc, hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

int vibe_mi_domain_set_physics_tendency(vibe_mi_domain *d, const double *theta_rate, const double *u_rate,
                                        const double *v_rate) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (theta_rate == nullptr && u_rate == nullptr && v_rate == nullptr) {
    d->ph_on = false;
    return 0;
  }
  if (int rc = set_device(d)) return rc;
  if (int rc = ensure_physics(d)) return rc;
  const size_t nc = d->ncell * 8;
  const double *src[3] = {theta_rate, u_rate, v_rate};
  double *dst[3] = {d->ph_th, d->ph_u, d->ph_v};
  for (int q = 0; q < 3; ++q) {
    if (src[q] != nullptr)
      VMI_HIP(hipMemcpyAsync(dst[q], src[q], nc, hipMemcpyHostToDevice, d->stream));
    else
      VMI_HIP(hipMemsetAsync(dst[q], 0, nc, d->stream));
  }
  VMI_HIP(hipStreamSynchronize(d->stream));
  d->ph_on = true;
  return 0;
}

int vibe_mi_domain_physics_buffers(vibe_mi_domain *d, double **theta_rate, double **u_rate, double **v_rate) {
  if (d == nullptr || theta_rate == nullptr || u_rate == nullptr || v_rate == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  if (int rc = ensure_physics(d)) return rc;
  *theta_rate = d->ph_th;
  *u_rate = d->ph_u;
  *v_rate = d->ph_v;
  d->ph_on = true;
  return 0;
}

int vibe_mi_domain_physics_enable(vibe_mi_domain *d, int on) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  d->ph_enabled = on != 0;
  return 0;
}

int vibe_mi_domain_set_physics_callback(vibe_mi_domain *d, vibe_mi_physics_fn fn, void *user, int host) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (fn == nullptr) {
    d->ph_fn = nullptr;
    d->ph_user = nullptr;
    d->ph_on = false;
    return 0;
  }
  if (int rc = set_device(d)) return rc;
  if (int rc = ensure_physics(d)) return rc;
  d->ph_fn = fn;
  d->ph_user = user;
  d->ph_host = host ? 1 : 0;
  d->ph_on = true;
  return 0;
}

int vibe_mi_domain_substep(vibe_mi_domain *d, double dtau) {
  if (d == nullptr || d->precision != VIBE_MI_FP64) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  FieldSet &s = d->S[d->cur0];
  FieldSet &o = d->S[(d->cur0 + 1) % 3];
  if (int rc = run_substep<double, double, double>(d, d->g64, 3, s.u, s.v, s.w, s.rho, s.pi, o.u, o.v, o.w, o.rho,
                                                   o.pi, dtau, 1))
    return rc;
  VMI_HIP(hipMemcpyAsync(o.th, s.th, d->ncell * 8, hipMemcpyDeviceToDevice, d->stream));
  d->cur0 = (d->cur0 + 1) % 3;
  return check_flag(d);
}

int vibe_mi_domain_energy(vibe_mi_domain *d, double out[6]) {
  if (d == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  const FieldSet &f = d->S[d->cur0];
  const size_t chunk = (d->ncell + kRedBlocks - 1) / kRedBlocks;
  const int blocks = static_cast<int>((d->ncell + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(kRedThreads), 0, d->stream, d->nx, d->ny, d->nz, d->dx,
                     d->dy, d->ztop, d->zs, d->dzeta64, d->zc64, d->rho0, d->theta0, d->pi0, f.u, f.v, f.w, f.rho,
                     f.th, f.pi, chunk, d->part);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, d->stream, d->part, blocks, kEnergyComps, d->dres);
  VMI_HIP(hipGetLastError());
  double r[kEnergyComps];
  VMI_HIP(hipMemcpyAsync(r, d->dres, sizeof(r), hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  for (int q = 0; q < kEnergyComps; ++q) out[q] = r[q];
  out[5] = r[2] + r[3] + r[4];
  return 0;
}

// One pass of k_cfl_state into r[8] (see the kernel): advective CFL, max
// wind, acoustic CFL, vertical CFL, min dz, non-finite flag, max sound speed,
// max of 1/dx + 1/dy + 1/dz (bit patterns of the doubles).
static int stability_pass(vibe_mi_domain *d, double dt, unsigned long long r[8]) {
  if (int rc = set_device(d)) return rc;
  const FieldSet &f = d->S[d->cur0];
  const unsigned long long init[8] = {0ull, 0ull, 0ull, 0ull, 0x7FF0000000000000ull, 0ull, 0ull, 0ull};  // min dz from +inf
  VMI_HIP(hipMemcpyAsync(d->dmax, init, sizeof(init), hipMemcpyHostToDevice, d->stream));
  const unsigned blocks = std::min<size_t>(ceil_div(d->ncell, 256), 2048);
  hipLaunchKernelGGL(k_cfl_state, dim3(blocks), dim3(256), 0, d->stream, d->nx, d->ny, d->nz, d->dx, d->dy, d->ztop,
                     d->zs, d->dzeta64, f.u, f.v, f.w, f.rho, f.th, f.pi, d->theta0, d->pi0, dt, d->dmax);
  VMI_HIP(hipGetLastError());
  VMI_HIP(hipMemcpyAsync(r, d->dmax, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

int vibe_mi_domain_cfl(vibe_mi_domain *d, double dt, double *cfl, double *max_wind) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  unsigned long long r[8];
  if (int rc = stability_pass(d, dt, r)) return rc;
  if (cfl) *cfl = __builtin_bit_cast(double, r[0]);
  if (max_wind) *max_wind = __builtin_bit_cast(double, r[1]);
  return 0;
}

int vibe_mi_domain_stability(vibe_mi_domain *d, double dt, double out[6]) {
  if (d == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  unsigned long long r[8];
  if (int rc = stability_pass(d, dt, r)) return rc;
  out[0] = __builtin_bit_cast(double, r[0]);  // advective CFL
  out[1] = __builtin_bit_cast(double, r[2]);  // acoustic CFL
  out[2] = __builtin_bit_cast(double, r[3]);  // vertical CFL
  out[3] = __builtin_bit_cast(double, r[1]);  // max wind
  out[4] = __builtin_bit_cast(double, r[4]);  // min dz
  // stable: finite state, advective CFL < 1.5 and vertical CFL < 2 (timestep_control.cpp:48)
  out[5] = (r[5] == 0 && out[0] < 1.5 && out[2] < 2.0) ? 1.0 : 0.0;
  return 0;
}

int vibe_mi_domain_max_stable_dt(vibe_mi_domain *d, double *dt_out) {
  if (d == nullptr || dt_out == nullptr) return VIBE_C_ERR_BAD_ARG;
  unsigned long long r[8];
  if (int rc = stability_pass(d, 1.0, r)) return rc;
  // RungeKutta3Integrator::max_stable_dt (runge_kutta.cpp:121-141): the
  // advective limit 0.8 / CFL(dt = 1) and the one-substep acoustic limit
  // 2 / max(cs_max (1/dx + 1/dy + 1/dz)), cs_max the state's largest sound
  // speed; the product's maximum is cs_max times the geometric maximum
  const double adv = __builtin_bit_cast(double, r[0]);
  const double rate = __builtin_bit_cast(double, r[6]) * __builtin_bit_cast(double, r[7]);
  const double dt_adv = adv > 0.0 ? 0.8 / adv : 1e9;
  const double dt_ac = rate > 0.0 ? 2.0 / rate : 1e9;
  *dt_out = std::min(dt_adv, dt_ac);
  return 0;
}

void *vibe_mi_domain_stream(vibe_mi_domain *d) { return d ? static_cast<void *>(d->stream) : nullptr; }

size_t vibe_mi_domain_device_bytes(vibe_mi_domain *d) { return d ? d->bytes : 0; }

}  // extern "C"
