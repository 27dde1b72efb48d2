// batch.hip -- many domains of uneven size advanced by one persistent launch
// per kernel phase, from device-side tile work lists.
//
// Per RK3 step of the active members:
//   kb_reset                   zero this step's work-queue counters
//   for each stage st:
//     kb_stage    64 x TY      fused Omega, slow tendencies and RK3 update
//     nsub x kb_substep rows   fused HEVI acoustic substep
//   kb_advance                 rotate every active member's state buffers
// The Rayleigh sponge is applied as the last stage and the last substep
// write the step's final state; FP32 batches keep the intermediate stage
// states in the FP32 acoustic buffers.
// Each launch is persistent: a grid of about one wave of resident
// workgroups, each claiming tiles from an atomic counter until the list is
// drained (tiles are ordered largest column first so the tail is short).
// Tiles are {member, x-block, y-block}; every member's operands (geometry,
// state buffers, HEVI tables, stage lengths) come from a device descriptor,
// read once per tile into scalar registers.  The tile bodies are the same
// functions as the single-domain kernels', so each member's result is
// identical to stepping it alone.
//
// hipGraph: the whole launch sequence of one step reads everything that
// changes between steps (buffer rotation, dt) from device memory, so it is
// captured once per (active set, nsub) and replayed; counters are cleared by
// a kernel, not a memset node.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "batch_kernels.hpp"
#include "vibe_mi/batch.h"
#include "vibe_mi/nest.h"

// FP32 kernels live in batch_f32_{flat,x,any}.hip (compiled with
// multiply-add contraction)
VMI_BATCH_F32_ALL(extern)

using namespace vibe_mi;

// ---------------------------------------------------------------------------
// batch handle
// ---------------------------------------------------------------------------

namespace {

constexpr int kMaxLaunches = 128;                 // persistent launches per step: 3 (1 + nsub) + 1
constexpr int kMaxCounters = kMaxLaunches * kXcds;  // one counter per XCD segment and launch
constexpr int kMaxNsub = 64;

struct WorkList {
  TileList rows{};       // {member, x-block, j}
  TileList stage[3]{};   // {member, x-block, y-block} for TY = 2, 4, 8
  int *ids = nullptr;    // active members
  int nids = 0;
  std::vector<int> members;  // active members (host copy)
  // tile shapes for this active set (autotuned on its first step): waves per
  // substep workgroup and rows per stage workgroup
  int nw = 8, ty = 4;
  bool tuned = false;
};

}  // namespace

struct vibe_mi_batch {
  int precision = VIBE_MI_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<vibe_mi_domain *> doms;
  void *desc = nullptr;  // BDom<T>[ndom] on the device
  std::vector<unsigned char> hdesc;
  int *ctr = nullptr;
  int *flag = nullptr;
  std::vector<void *> allocs;
  size_t bytes = 0;
  int nzmax = 0;
  std::map<std::string, WorkList> lists;
  std::map<std::string, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_objs;
  bool use_graph = true;
  int nw = 8;  // waves per substep workgroup
  int ty = 4;  // rows per stage workgroup
  int cus = 256;
  int gm = kGeoFlat;  // geometry mode of the batch: the most general of its members
  bool phys = false;  // an active member has physics rates: the stage kernels add them
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // [5]: nest relaxation
  // per-group kernel times of the last timed eager call (ms, summed)
  bool group_timing = false;
  double ms_stage = 0, ms_sub = 0;
  int n_stage = 0, n_sub = 0;
};

namespace {

template <typename P>
int balloc(vibe_mi_batch *b, P **out, size_t count) {
  void *p = nullptr;
  const size_t n = count * sizeof(P);
  VMI_HIP(hipMalloc(&p, n > 0 ? n : 1));
  b->allocs.push_back(p);
  b->bytes += n;
  *out = static_cast<P *>(p);
  return 0;
}

template <typename P>
int bupload(vibe_mi_batch *b, P **out, const std::vector<P> &h) {
  if (int rc = balloc(b, out, h.size())) return rc;
  if (!h.empty()) VMI_HIP(hipMemcpy(*out, h.data(), h.size() * sizeof(P), hipMemcpyHostToDevice));
  return 0;
}

std::string key_of(const std::vector<unsigned char> &act) { return std::string(act.begin(), act.end()); }

// Uploads a tile list with its XCD segments: contiguous runs of near-equal
// cost (cost of a tile = levels of its member).
int upload_list(vibe_mi_batch *b, const std::vector<int4> &t, TileList *out) {
  std::vector<double> cum(t.size() + 1, 0.0);
  for (size_t q = 0; q < t.size(); ++q) cum[q + 1] = cum[q] + b->doms[t[q].x]->nz;
  std::vector<int> seg(kXcds + 1, 0);
  for (int x = 1; x < kXcds; ++x) {
    const double target = cum.back() * x / kXcds;
    seg[x] = static_cast<int>(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
    seg[x] = std::max(seg[x], seg[x - 1]);
  }
  seg[kXcds] = static_cast<int>(t.size());
  int4 *dt = nullptr;
  int *ds = nullptr;
  if (int rc = bupload(b, &dt, t)) return rc;
  if (int rc = bupload(b, &ds, seg)) return rc;
  *out = TileList{dt, ds, static_cast<int>(t.size())};
  return 0;
}

// Builds (or returns the cached) work lists of an active set.  Members are
// visited deepest column first so the longest tiles are claimed first.
int work_list(vibe_mi_batch *b, const std::vector<unsigned char> &act, WorkList **out) {
  const std::string key = key_of(act);
  auto it = b->lists.find(key);
  if (it != b->lists.end()) {
    *out = &it->second;
    return 0;
  }
  std::vector<int> order;
  for (int m = 0; m < static_cast<int>(b->doms.size()); ++m)
    if (act[m]) order.push_back(m);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->doms[x]->nz > b->doms[y]->nz; });
  std::vector<int4> rows, st2, st4, st8;
  for (int m : order) {
    const vibe_mi_domain *d = b->doms[m];
    const int nbx = static_cast<int>(ceil_div(d->nx, 64));
    for (int j = 0; j < d->ny; ++j)
      for (int bx = 0; bx < nbx; ++bx) rows.push_back(make_int4(m, bx, j, 0));
    for (int by = 0; by < static_cast<int>(ceil_div(d->ny, 2)); ++by)
      for (int bx = 0; bx < nbx; ++bx) st2.push_back(make_int4(m, bx, by, 0));
    for (int by = 0; by < static_cast<int>(ceil_div(d->ny, 4)); ++by)
      for (int bx = 0; bx < nbx; ++bx) st4.push_back(make_int4(m, bx, by, 0));
    for (int by = 0; by < static_cast<int>(ceil_div(d->ny, 8)); ++by)
      for (int bx = 0; bx < nbx; ++bx) st8.push_back(make_int4(m, bx, by, 0));
  }
  WorkList w;
  if (int rc = upload_list(b, rows, &w.rows)) return rc;
  if (int rc = upload_list(b, st2, &w.stage[0])) return rc;
  if (int rc = upload_list(b, st4, &w.stage[1])) return rc;
  if (int rc = upload_list(b, st8, &w.stage[2])) return rc;
  if (int rc = bupload(b, &w.ids, order)) return rc;
  w.nids = static_cast<int>(order.size());
  w.members = order;
  *out = &(b->lists[key] = w);
  return 0;
}

// Resident workgroups of a kernel on the whole device (persistent grid size).
template <typename K>
int resident(vibe_mi_batch *b, K kernel, int threads, size_t lds, int ntiles, unsigned *grid) {
  int per_cu = 0;
  VMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
  const long cap = static_cast<long>(std::max(per_cu, 1)) * b->cus;
  *grid = static_cast<unsigned>(std::max(1L, std::min<long>(cap, ntiles)));
  return 0;
}

// Writes the member descriptors for step lengths dt[] (cur0 from the host
// mirror in each member handle) and factorises the HEVI tables.
template <typename T>
int upload_desc(vibe_mi_batch *b, const double *dt, int nsub) {
  const int n = static_cast<int>(b->doms.size());
  std::vector<BDom<T>> h(n);
  for (int m = 0; m < n; ++m) {
    vibe_mi_domain *d = b->doms[m];
    BDom<T> &x = h[m];
    std::memset(static_cast<void *>(&x), 0, sizeof(x));
    if constexpr (sizeof(T) == 4) {
      x.g = d->g32;
      x.A[0] = d->A32[0];
      x.A[1] = d->A32[1];
    } else {
      x.g = d->g64;
      x.A[0] = d->A64[0];
      x.A[1] = d->A64[1];
    }
    for (int q = 0; q < 3; ++q) x.S[q] = d->S[q];
    const double hk[3] = {dt[m] / 3.0, dt[m] / 2.0, dt[m]};
    for (int st = 0; st < 3; ++st) {
      const double dtau = hk[st] / static_cast<double>(nsub);
      if (int rc = domain_ensure_table(d, st, dtau)) return rc;
      x.tab[st] = HeviTab<T>::at(static_cast<const T *>(domain_table(d, st)), d->nz, d->rk);
      x.h[st] = hk[st];
      x.dtau[st] = static_cast<T>(dtau);
    }
    x.an = static_cast<T>(0.5 * (1.0 + d->eps_s));
    x.ao = static_cast<T>(0.5 * (1.0 - d->eps_s));
    x.dt = dt[m];
    x.sponge = d->sponge;
    x.ph = (d->ph_on && d->ph_enabled) ? PhysArgs{d->ph_th, d->ph_u, d->ph_v} : PhysArgs{};
    x.kmin_sponge = d->kmin_sponge;
    x.cur0 = d->cur0;
    x.flag = b->flag;
    if (d->precision == VIBE_MI_MIXED_COMP) {
      x.L[0] = d->L32[0];
      x.L[1] = d->L32[1];
      const Geo<double> &g8 = d->g64;
      x.rhu64 = g8.rhu, x.rhv64 = g8.rhv, x.r0w64 = g8.r0w, x.sx64 = g8.sx, x.sy64 = g8.sy, x.zeta64 = g8.zeta;
      x.rdx64 = g8.rdx, x.rdy64 = g8.rdy;
    }
    VMI_HIP(hipStreamSynchronize(d->stream));
  }
  VMI_HIP(hipMemcpyAsync(b->desc, h.data(), sizeof(BDom<T>) * n, hipMemcpyHostToDevice, b->stream));
  VMI_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

// dry: only set the kernel's LDS attribute (before a graph capture).
template <typename T, typename O, int NW, int GM, bool CP>
int launch_substep_nw(vibe_mi_batch *b, const WorkList &w, int st, int q, int abase, int last, int *ctr, bool dry) {
  const auto *D = static_cast<const BDom<T> *>(b->desc);
  const int ch = (b->nzmax + NW - 1) / NW;
  const size_t lds = substep_lds_arrays(CP) * static_cast<size_t>(b->nzmax - 1) * 64 * sizeof(T);
  // s_mc, s_fh (FP64 in the compensated mode), s_ph and the tile loop's slots
  const size_t lds_static = ((CP ? 2 * sizeof(double) : 2 * sizeof(T)) + sizeof(T)) * NW * 64 + 64;
  if (lds + lds_static > 160 * 1024) return VIBE_C_ERR_BAD_ARG;
#define VMI_BSUB1(M, RO)                                                                                      \
  do {                                                                                                        \
    auto *kern = &kb_substep<T, O, NW, M, GM, CP, RO>;                                                            \
    static bool attr = false;                                                                                 \
    if (!attr) {                                                                                              \
      VMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),                                       \
                                  hipFuncAttributeMaxDynamicSharedMemorySize,                                 \
                                  static_cast<int>(160 * 1024 - lds_static)));                                \
      attr = true;                                                                                            \
    }                                                                                                         \
    if (dry) break;                                                                                           \
    unsigned grid = 0;                                                                                        \
    if (int rc = resident(b, kern, 64 * NW, lds, w.rows.ntiles, &grid)) return rc;                            \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * NW), lds, b->stream, D, w.rows, st, q, abase, last, ctr);  \
  } while (0)
  // FP32 outputs: the stage's last substep (last = 1) is the variant that
  // writes rho' (dry: set both variants' attribute)
#define VMI_BSUB(M)                                 \
  do {                                              \
    if constexpr (sizeof(T) == 4 && sizeof(O) == 4) { \
      if (dry || !last) VMI_BSUB1(M, false);        \
      if (dry || last) VMI_BSUB1(M, true);          \
    } else {                                        \
      VMI_BSUB1(M, (sizeof(O) == 8));               \
    }                                               \
  } while (0)
  if constexpr (NW == 16) {
    if (ch <= 4) VMI_BSUB(4);
    else if (ch <= 8) VMI_BSUB(8);
    else return VIBE_C_ERR_BAD_ARG;
  } else if constexpr (NW == 8) {
    if (ch <= 4) VMI_BSUB(4);
    else if (ch <= 8) VMI_BSUB(8);
    else if (ch <= 12) VMI_BSUB(12);
    else if (ch <= 16) VMI_BSUB(16);
    else return VIBE_C_ERR_BAD_ARG;
  } else {
    if (ch <= 8) VMI_BSUB(8);
    else if (ch <= 12) VMI_BSUB(12);
    else if (ch <= 16) VMI_BSUB(16);
    else if (ch <= 24) VMI_BSUB(24);
    else if (ch <= 32) VMI_BSUB(32);
    else return VIBE_C_ERR_BAD_ARG;
  }
#undef VMI_BSUB
#undef VMI_BSUB1
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T, typename O, int GM, bool CP>
int launch_substep_cp(vibe_mi_batch *b, int nw, const WorkList &w, int st, int q, int abase, int last, int *ctr, bool dry) {
  if (nw == 16) return launch_substep_nw<T, O, 16, GM, CP>(b, w, st, q, abase, last, ctr, dry);
  if (nw == 8) return launch_substep_nw<T, O, 8, GM, CP>(b, w, st, q, abase, last, ctr, dry);
  return launch_substep_nw<T, O, 4, GM, CP>(b, w, st, q, abase, last, ctr, dry);
}

template <typename T, typename O, int GM>
int launch_substep_gm(vibe_mi_batch *b, int nw, const WorkList &w, int st, int q, int abase, int last, int *ctr, bool dry) {
  if constexpr (sizeof(T) == 4)
    if (b->precision == VIBE_MI_MIXED_COMP)
      return launch_substep_cp<T, O, GM, true>(b, nw, w, st, q, abase, last, ctr, dry);
  return launch_substep_cp<T, O, GM, false>(b, nw, w, st, q, abase, last, ctr, dry);
}

template <typename T, typename O>
int launch_substep(vibe_mi_batch *b, int nw, const WorkList &w, int st, int q, int abase, int last, int *ctr,
                   bool dry = false) {
  if (b->gm == kGeoFlat) return launch_substep_gm<T, O, kGeoFlat>(b, nw, w, st, q, abase, last, ctr, dry);
  if (b->gm == kGeoX) return launch_substep_gm<T, O, kGeoX>(b, nw, w, st, q, abase, last, ctr, dry);
  return launch_substep_gm<T, O, kGeoAny>(b, nw, w, st, q, abase, last, ctr, dry);
}

// ain < 0: the stage reads the FP64 current state; otherwise (FP32 batches)
// the FP32 acoustic set A[ain].  The stage writes A[aout].
template <typename T, int TY, int GM, bool CUR32, bool PH>
int launch_stage_k(vibe_mi_batch *b, const WorkList &w, int st, int ain, int aout, int *ctr) {
  const auto *D = static_cast<const BDom<T> *>(b->desc);
  const int li = TY == 2 ? 0 : (TY == 4 ? 1 : 2);
  auto *kern = &kb_stage<T, TY, GM, CUR32, PH>;
  unsigned grid = 0;
  if (int rc = resident(b, kern, 64 * TY, 0, w.stage[li].ntiles, &grid)) return rc;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64, TY), 0, b->stream, D, w.stage[li], st, ain, aout, ctr);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T, int TY, int GM, bool PH>
int launch_stage_ph(vibe_mi_batch *b, const WorkList &w, int st, int ain, int aout, int *ctr) {
  if constexpr (sizeof(T) == 4) {
    if (ain >= 0) return launch_stage_k<T, TY, GM, true, PH>(b, w, st, ain, aout, ctr);
  }
  return launch_stage_k<T, TY, GM, false, PH>(b, w, st, ain, aout, ctr);
}

template <typename T, int TY, int GM>
int launch_stage_ty(vibe_mi_batch *b, const WorkList &w, int st, int ain, int aout, int *ctr) {
  if (b->phys) return launch_stage_ph<T, TY, GM, true>(b, w, st, ain, aout, ctr);
  return launch_stage_ph<T, TY, GM, false>(b, w, st, ain, aout, ctr);
}

template <typename T, int GM>
int launch_stage_gm(vibe_mi_batch *b, int ty, const WorkList &w, int st, int ain, int aout, int *ctr) {
  if (ty == 2) return launch_stage_ty<T, 2, GM>(b, w, st, ain, aout, ctr);
  if (ty == 8) return launch_stage_ty<T, 8, GM>(b, w, st, ain, aout, ctr);
  return launch_stage_ty<T, 4, GM>(b, w, st, ain, aout, ctr);
}

template <typename T>
int launch_stage(vibe_mi_batch *b, int ty, const WorkList &w, int st, int ain, int aout, int *ctr) {
  if (b->gm == kGeoFlat) return launch_stage_gm<T, kGeoFlat>(b, ty, w, st, ain, aout, ctr);
  if (b->gm == kGeoX) return launch_stage_gm<T, kGeoX>(b, ty, w, st, ain, aout, ctr);
  return launch_stage_gm<T, kGeoAny>(b, ty, w, st, ain, aout, ctr);
}

// Enqueues one RK3 step of the active set on the batch stream (capturable:
// kernels only).
template <typename T>
int enqueue_step(vibe_mi_batch *b, const WorkList &w, int nsub) {
  int *ctr = b->ctr;
  int c = 0;
  hipLaunchKernelGGL(kb_reset, dim3(1), dim3(64), 0, b->stream, ctr, kMaxCounters);
  const bool gt = b->group_timing;
  // Acoustic buffers: stage st writes A[aout], its substeps ping-pong from
  // there and end in A[aend].  In FP32 batches the intermediate stage states
  // s^(1), s^(2) only feed the next stage's tendencies, so stages 1 and 2
  // end in FP32 (the next stage reads A[aend] and writes the other buffer);
  // only the step's last substep writes the FP64 state.  FP64 batches keep
  // every stage state in FP64.
  int aout = 0, aend = 0;
  for (int st = 0; st < 3; ++st) {
    const bool cur32 = sizeof(T) == 4 && st > 0;
    const int ain = cur32 ? aend : -1;
    aout = cur32 ? (aend ^ 1) : 0;
    aend = (aout + nsub) & 1;
    if (gt) VMI_HIP(hipEventRecord(b->ev[2], b->stream));
    if (int rc = launch_stage<T>(b, w.ty, w, st, ain, aout, ctr + kXcds * c++)) return rc;
    if (gt) VMI_HIP(hipEventRecord(b->ev[3], b->stream));
    for (int q = 0; q < nsub; ++q) {
      const bool last = q == nsub - 1;
      int rc;
      if constexpr (sizeof(T) == 4) {
        rc = (last && st == 2) ? launch_substep<T, double>(b, w.nw, w, st, q, aout, 1, ctr + kXcds * c++)
                               : launch_substep<T, T>(b, w.nw, w, st, q, aout, last ? 1 : 0, ctr + kXcds * c++);
      } else {
        rc = launch_substep<T, T>(b, w.nw, w, st, q, aout, last ? 1 : 0, ctr + kXcds * c++);
      }
      if (rc) return rc;
    }
    if (gt) {
      VMI_HIP(hipEventRecord(b->ev[4], b->stream));
      VMI_HIP(hipEventSynchronize(b->ev[4]));
      float a = 0.f, x = 0.f;
      VMI_HIP(hipEventElapsedTime(&a, b->ev[2], b->ev[3]));
      VMI_HIP(hipEventElapsedTime(&x, b->ev[3], b->ev[4]));
      b->ms_stage += a;
      b->ms_sub += x;
      b->n_stage += 1;
      b->n_sub += nsub;
    }
  }
  // the sponge is applied by the step's last stage kernel (theta') and last
  // substep (u, v, w, rho', pi') as they write the final state
  hipLaunchKernelGGL((kb_advance<T>), dim3(1), dim3(64), 0, b->stream, static_cast<BDom<T> *>(b->desc), w.ids, w.nids);
  VMI_HIP(hipGetLastError());
  return 0;
}

// Autotuner: waves per substep workgroup and rows per stage workgroup,
// measured on this batch's full work list with HIP events (the substep runs
// on the acoustic scratch buffers, the stage writes the stage-1 output
// buffers, neither touches s^n), cached per batch shape for the process.
std::map<std::string, std::pair<int, int>> g_btune;

template <typename T>
int autotune(vibe_mi_batch *b, WorkList &w) {
  if (w.tuned) {
    b->nw = w.nw;
    b->ty = w.ty;
    return 0;
  }
  // keyed by the active members' shapes (a small active set of a nest
  // hierarchy wants smaller tiles than the whole batch)
  std::string shape = std::to_string(sizeof(T)) + (b->phys ? "p" : "") + (b->precision == VIBE_MI_MIXED_COMP ? "c" : "");
  for (int m : w.members) {
    const vibe_mi_domain *d = b->doms[m];
    shape += "," + std::to_string(d->nx) + "x" + std::to_string(d->ny) + "x" + std::to_string(d->nz) + "r" + std::to_string(d->rk);
  }
  // VIBE_MI_SUBSTEP_WAVES (4, 8, 16) and VIBE_MI_STAGE_ROWS (2, 4, 8) pin a
  // shape instead of measuring it (tests, profiling); pinned shapes are not
  // cached
  const char *env_nw = std::getenv("VIBE_MI_SUBSTEP_WAVES");
  const char *env_ty = std::getenv("VIBE_MI_STAGE_ROWS");
  auto it = (env_nw || env_ty) ? g_btune.end() : g_btune.find(shape);
  if (it != g_btune.end()) {
    w.nw = b->nw = it->second.first;
    w.ty = b->ty = it->second.second;
    w.tuned = true;
    return 0;
  }
  bool pin_nw = false, pin_ty = false;
  if (env_nw) {
    const int v = std::atoi(env_nw);
    if (v == 4 || v == 8 || v == 16) {
      w.nw = v;
      pin_nw = true;
    }
  }
  if (env_ty) {
    const int v = std::atoi(env_ty);
    if (v == 2 || v == 4 || v == 8) {
      w.ty = v;
      pin_ty = true;
    }
  }
  auto time_it = [&](auto &&launch, float *ms) -> int {
    float total = 0.f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(kb_reset, dim3(1), dim3(64), 0, b->stream, b->ctr, kMaxCounters);
      VMI_HIP(hipEventRecord(b->ev[0], b->stream));
      if (int rc = launch()) return rc;
      VMI_HIP(hipEventRecord(b->ev[1], b->stream));
      VMI_HIP(hipEventSynchronize(b->ev[1]));
      float x = 0.f;
      VMI_HIP(hipEventElapsedTime(&x, b->ev[0], b->ev[1]));
      if (rep > 0) total += x;
    }
    *ms = total;
    return 0;
  };
  // two passes over the candidates, the faster of each candidate's two
  // timings kept: the first launches of a process run while clocks and
  // caches settle, which would otherwise penalise whichever shape is tried
  // first
  if (!pin_nw) {
    const int cand[3] = {4, 8, 16};
    float best_of[3] = {1e30f, 1e30f, 1e30f};
    for (int pass = 0; pass < 2; ++pass)
      for (int c = 0; c < 3; ++c) {
        float ms = 0.f;
        const int rc = time_it([&]() { return launch_substep<T, T>(b, cand[c], w, 0, 0, 0, 0, b->ctr); }, &ms);
        if (rc == 0 && ms < best_of[c]) best_of[c] = ms;
      }
    float best = 1e30f;
    for (int c = 0; c < 3; ++c)
      if (best_of[c] < best) {
        best = best_of[c];
        w.nw = cand[c];
      }
    (void)hipGetLastError();
  }
  if (!pin_ty) {
    const int cand[3] = {2, 4, 8};
    float best_of[3] = {1e30f, 1e30f, 1e30f};
    for (int pass = 0; pass < 2; ++pass)
      for (int c = 0; c < 3; ++c) {
        float ms = 0.f;
        const int rc = time_it([&]() { return launch_stage<T>(b, cand[c], w, 0, -1, 0, b->ctr); }, &ms);
        if (rc == 0 && ms < best_of[c]) best_of[c] = ms;
      }
    float best = 1e30f;
    for (int c = 0; c < 3; ++c)
      if (best_of[c] < best) {
        best = best_of[c];
        w.ty = cand[c];
      }
    (void)hipGetLastError();
  }
  w.tuned = true;
  b->nw = w.nw;
  b->ty = w.ty;
  if (!env_nw && !env_ty) g_btune[shape] = {w.nw, w.ty};
  return 0;
}

void free_batch(vibe_mi_batch *b) {
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : b->graph_objs) (void)hipGraphDestroy(g);
  b->graphs.clear();
  b->graph_objs.clear();
  for (void *p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  for (auto *d : b->doms) vibe_mi_domain_destroy(d);
  b->doms.clear();
  for (auto &e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

template <typename T>
int step_batch(vibe_mi_batch *b, const double *dt, int nsub, int nsteps, const unsigned char *active, bool timed,
               double *ms_total) {
  const int n = static_cast<int>(b->doms.size());
  std::vector<unsigned char> act(n, 1);
  if (active)
    for (int m = 0; m < n; ++m) act[m] = active[m] ? 1 : 0;
  for (int m = 0; m < n; ++m)
    if (act[m] && (!(dt[m] > 0.0) || b->doms[m]->ph_fn != nullptr)) return VIBE_C_ERR_BAD_ARG;
  WorkList *w = nullptr;
  if (int rc = work_list(b, act, &w)) return rc;
  if (w->nids == 0 || nsteps == 0) {
    if (ms_total) *ms_total = 0.0;
    return 0;
  }
  if (int rc = upload_desc<T>(b, dt, nsub)) return rc;
  b->phys = false;
  for (int m = 0; m < n; ++m)
    if (act[m] && b->doms[m]->ph_on && b->doms[m]->ph_enabled) b->phys = true;
  if (int rc = autotune<T>(b, *w)) return rc;
  if constexpr (sizeof(T) == 4)
    if (int rc = launch_substep<T, double>(b, w->nw, *w, 0, 0, 0, 1, b->ctr, true)) return rc;
  if (int rc = launch_substep<T, T>(b, w->nw, *w, 0, 0, 0, 0, b->ctr, true)) return rc;
  VMI_HIP(hipMemsetAsync(b->flag, 0, sizeof(int), b->stream));
  hipGraphExec_t exec = nullptr;
  if (b->use_graph) {
    const std::string gk = key_of(act) + "|" + std::to_string(nsub) + (b->phys ? "|phys" : "");
    auto it = b->graphs.find(gk);
    if (it == b->graphs.end()) {
      hipGraph_t graph = nullptr;
      VMI_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_step<T>(b, *w, nsub);
      hipGraph_t captured = nullptr;
      const hipError_t ec = hipStreamEndCapture(b->stream, &captured);
      if (rc) return rc;
      VMI_HIP(ec);
      graph = captured;
      b->graph_objs.push_back(graph);
      VMI_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      b->graphs[gk] = exec;
    } else {
      exec = it->second;
    }
  }
  b->group_timing = timed && exec == nullptr;
  if (timed) {
    b->ms_stage = b->ms_sub = 0;
    b->n_stage = b->n_sub = 0;
  }
  if (timed) VMI_HIP(hipEventRecord(b->ev[0], b->stream));
  for (int s = 0; s < nsteps; ++s) {
    if (exec) {
      VMI_HIP(hipGraphLaunch(exec, b->stream));
    } else if (int rc = enqueue_step<T>(b, *w, nsub)) {
      return rc;
    }
  }
  if (timed) VMI_HIP(hipEventRecord(b->ev[1], b->stream));
  b->group_timing = false;
  int f = 0;
  VMI_HIP(hipMemcpyAsync(&f, b->flag, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  VMI_HIP(hipStreamSynchronize(b->stream));
  if (timed && ms_total) {
    float ms = 0.f;
    VMI_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    *ms_total = ms;
  }
  for (int m = 0; m < n; ++m)
    if (act[m]) b->doms[m]->cur0 = (b->doms[m]->cur0 + nsteps) % 3;
  return f ? VIBE_MI_ERR_NONFINITE : 0;
}

// ---------------------------------------------------------------------------
// Nest hierarchy: whole parent steps replayed from one graph
// ---------------------------------------------------------------------------
// One parent step of a hierarchy of ratio r (levels 0..L) is r^L sub-steps.
// At sub-step s the members of level l step (dt0 / r^l) when
// s % r^(L-l) == 0, all in the same batched launches; after it every nest
// that stepped is relaxed toward its parent (Davies, the parent interpolated
// in time across its current step, frac = (m + 1) / r for the nest's m-th
// step in it), coarse levels first; with two_way, after the last sub-step
// every nest feeds back onto its parent, finest level first.  The whole
// sequence is enqueued on the batch stream and captured once per buffer
// rotation state (the nest operators take the state buffers as arguments,
// and a parent step advances level l's rotation by r^l mod 3), so a run of
// parent steps costs one graph launch each and a single synchronisation.
struct NestLink {
  int child, parent;
  vibe_mi_nest_desc d;
};

template <typename T>
int nested_steps(vibe_mi_batch *b, const std::vector<NestLink> &links, bool two_way, double dt0, int nsub, int nsteps,
                 bool timed, double *ms_total) {
  const int n = static_cast<int>(b->doms.size());
  if (links.empty() || !(dt0 > 0.0)) return VIBE_C_ERR_BAD_ARG;
  const int r = links[0].d.ratio;
  std::vector<int> par(n, -1);
  for (const auto &l : links) {
    if (l.child < 0 || l.child >= n || l.parent < 0 || l.parent >= n || l.child == l.parent || par[l.child] >= 0 ||
        l.d.ratio != r || r < 2)
      return VIBE_C_ERR_BAD_ARG;
    par[l.child] = l.parent;
  }
  std::vector<int> level(n, 0);
  for (int m = 0; m < n; ++m) {
    int depth = 0;
    for (int q = m; par[q] >= 0; q = par[q])
      if (++depth > n) return VIBE_C_ERR_BAD_ARG;  // a cycle
    level[m] = depth;
  }
  const int L = *std::max_element(level.begin(), level.end());
  auto ipow = [](int base, int e) {
    long v = 1;
    for (int q = 0; q < e; ++q) v *= base;
    return v;
  };
  if (ipow(r, L) > 4096) return VIBE_C_ERR_BAD_ARG;
  const int R = static_cast<int>(ipow(r, L));
  std::vector<double> dts(n);
  for (int m = 0; m < n; ++m) {
    dts[m] = dt0 / static_cast<double>(ipow(r, level[m]));
    if (b->doms[m]->ph_fn != nullptr) return VIBE_C_ERR_BAD_ARG;
  }
  // links coarse level first (relaxation order)
  std::vector<NestLink> order = links;
  std::stable_sort(order.begin(), order.end(),
                   [&](const NestLink &a, const NestLink &c) { return level[a.child] < level[c.child]; });
  struct Sub {
    std::vector<unsigned char> act;
    WorkList *w = nullptr;
    std::vector<std::pair<int, double>> relax;  // (index into order, frac)
  };
  std::vector<Sub> plan(R);
  for (int s = 0; s < R; ++s) {
    Sub &sb = plan[s];
    sb.act.assign(n, 0);
    for (int m = 0; m < n; ++m) sb.act[m] = (s % ipow(r, L - level[m]) == 0) ? 1 : 0;
    for (int q = 0; q < static_cast<int>(order.size()); ++q) {
      const int c = order[q].child;
      if (!sb.act[c]) continue;
      const long span = ipow(r, L - level[c]);
      sb.relax.emplace_back(q, static_cast<double>((s / span) % r + 1) / static_cast<double>(r));
    }
  }
  if (int rc = upload_desc<T>(b, dts.data(), nsub)) return rc;
  b->phys = false;
  for (int m = 0; m < n; ++m)
    if (b->doms[m]->ph_on && b->doms[m]->ph_enabled) b->phys = true;
  for (Sub &sb : plan) {
    if (int rc = work_list(b, sb.act, &sb.w)) return rc;
    if (int rc = autotune<T>(b, *sb.w)) return rc;
    if constexpr (sizeof(T) == 4)
      if (int rc = launch_substep<T, double>(b, sb.w->nw, *sb.w, 0, 0, 0, 1, b->ctr, true)) return rc;
    if (int rc = launch_substep<T, T>(b, sb.w->nw, *sb.w, 0, 0, 0, 0, b->ctr, true)) return rc;
  }
  std::string pkey = "nest|" + std::to_string(nsub) + "|" + std::to_string(two_way ? 1 : 0) + (b->phys ? "|p" : "");
  for (const auto &l : links)
    pkey += "|" + std::to_string(l.child) + ">" + std::to_string(l.parent) + ":" + std::to_string(l.d.i0) + "," +
            std::to_string(l.d.j0) + "," + std::to_string(l.d.zone) + "," + std::to_string(l.d.alpha_max);
  VMI_HIP(hipMemsetAsync(b->flag, 0, sizeof(int), b->stream));
  // the parent step's launch sequence with the host's buffer rotation
  // tracked through it (the nest operators read it when enqueued)
  auto enqueue_parent_step = [&]() -> int {
    for (const Sub &sb : plan) {
      if (int rc = enqueue_step<T>(b, *sb.w, nsub)) return rc;
      for (int m = 0; m < n; ++m)
        if (sb.act[m]) b->doms[m]->cur0 = (b->doms[m]->cur0 + 1) % 3;
      for (const auto &rx : sb.relax) {
        const NestLink &l = order[rx.first];
        if (int rc = nest_enqueue(b->doms[l.child], b->doms[l.parent], &l.d, rx.second, 0, b->stream)) return rc;
      }
    }
    if (two_way)
      for (auto it = order.rbegin(); it != order.rend(); ++it)
        if (int rc = nest_restrict_enqueue(b->doms[it->child], b->doms[it->parent], &it->d, b->stream)) return rc;
    return 0;
  };
  if (timed) VMI_HIP(hipEventRecord(b->ev[0], b->stream));
  for (int st = 0; st < nsteps; ++st) {
    std::string gk = pkey + "|rot";
    std::vector<int> rot0(n);
    for (int m = 0; m < n; ++m) {
      rot0[m] = b->doms[m]->cur0;
      gk += std::to_string(rot0[m]);
    }
    if (!b->use_graph) {
      if (int rc = enqueue_parent_step()) return rc;
      continue;
    }
    auto it = b->graphs.find(gk);
    hipGraphExec_t exec = nullptr;
    if (it == b->graphs.end()) {
      VMI_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_parent_step();
      hipGraph_t captured = nullptr;
      const hipError_t ec = hipStreamEndCapture(b->stream, &captured);
      for (int m = 0; m < n; ++m) b->doms[m]->cur0 = rot0[m];  // captured, not run yet
      if (rc) return rc;
      VMI_HIP(ec);
      b->graph_objs.push_back(captured);
      VMI_HIP(hipGraphInstantiate(&exec, captured, nullptr, nullptr, 0));
      b->graphs[gk] = exec;
    } else {
      exec = it->second;
    }
    VMI_HIP(hipGraphLaunch(exec, b->stream));
    for (int m = 0; m < n; ++m) b->doms[m]->cur0 = static_cast<int>((rot0[m] + ipow(r, level[m])) % 3);
  }
  if (timed) VMI_HIP(hipEventRecord(b->ev[1], b->stream));
  int f = 0;
  VMI_HIP(hipMemcpyAsync(&f, b->flag, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  VMI_HIP(hipStreamSynchronize(b->stream));
  if (timed && ms_total) {
    float ms = 0.f;
    VMI_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    *ms_total = ms;
  }
  // later work on a member's own stream is ordered after the batch stream
  VMI_HIP(hipEventRecord(b->ev[5], b->stream));
  for (auto *d : b->doms) VMI_HIP(hipStreamWaitEvent(d->stream, b->ev[5], 0));
  return f ? VIBE_MI_ERR_NONFINITE : 0;
}

int nested_common(vibe_mi_batch *b, int nlinks, const int *child, const int *parent, const vibe_mi_nest_desc *descs,
                  int two_way, double dt0, int nsub, int nsteps, bool timed, double *ms_total) {
  if (b == nullptr || nlinks < 1 || child == nullptr || parent == nullptr || descs == nullptr || nsteps < 0 ||
      nsub < 1 || 3 * (nsub + 1) + 1 > kMaxLaunches || nsub > kMaxNsub)
    return VIBE_C_ERR_BAD_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return VIBE_MI_ERR_HIP;
  for (auto *d : b->doms)
    if (hipStreamSynchronize(d->stream) != hipSuccess) return VIBE_MI_ERR_HIP;
  std::vector<NestLink> links(nlinks);
  for (int q = 0; q < nlinks; ++q) links[q] = NestLink{child[q], parent[q], descs[q]};
  const int rc = b->precision != VIBE_MI_FP64
                     ? nested_steps<float>(b, links, two_way != 0, dt0, nsub, nsteps, timed, ms_total)
                     : nested_steps<double>(b, links, two_way != 0, dt0, nsub, nsteps, timed, ms_total);
  if (rc) set_last_error(rc);
  return rc;
}

}  // namespace

extern "C" {

int vibe_mi_batch_nested_step(vibe_mi_batch *b, int nlinks, const int *child, const int *parent,
                              const vibe_mi_nest_desc *descs, int two_way, double dt0, int nsub, int nsteps) {
  return nested_common(b, nlinks, child, parent, descs, two_way, dt0, nsub, nsteps, false, nullptr);
}

int vibe_mi_batch_nested_step_timed(vibe_mi_batch *b, int nlinks, const int *child, const int *parent,
                                    const vibe_mi_nest_desc *descs, int two_way, double dt0, int nsub, int nsteps,
                                    double *ms_total) {
  return nested_common(b, nlinks, child, parent, descs, two_way, dt0, nsub, nsteps, true, ms_total);
}


int vibe_mi_batch_nest_relax(vibe_mi_batch *b, int child, int parent, const vibe_mi_nest_desc *nd, double frac) {
  if (b == nullptr) return VIBE_C_ERR_BAD_ARG;
  const int n = static_cast<int>(b->doms.size());
  if (child < 0 || child >= n || parent < 0 || parent >= n) return VIBE_C_ERR_BAD_ARG;
  vibe_mi_domain *c = b->doms[child];
  if (int rc = nest_enqueue(c, b->doms[parent], nd, frac, 0, b->stream)) return rc;
  // later work on the child's own stream (state copies) waits for the relaxation
  VMI_HIP(hipEventRecord(b->ev[5], b->stream));
  VMI_HIP(hipStreamWaitEvent(c->stream, b->ev[5], 0));
  return 0;
}

int vibe_mi_batch_nest_restrict(vibe_mi_batch *b, int child, int parent, const vibe_mi_nest_desc *nd) {
  if (b == nullptr) return VIBE_C_ERR_BAD_ARG;
  const int n = static_cast<int>(b->doms.size());
  if (child < 0 || child >= n || parent < 0 || parent >= n) return VIBE_C_ERR_BAD_ARG;
  vibe_mi_domain *p = b->doms[parent];
  if (int rc = nest_restrict_enqueue(b->doms[child], p, nd, b->stream)) return rc;
  VMI_HIP(hipEventRecord(b->ev[5], b->stream));
  VMI_HIP(hipStreamWaitEvent(p->stream, b->ev[5], 0));
  return 0;
}

int vibe_mi_batch_create(const vibe_mi_domain_desc *descs, int ndom, vibe_mi_batch **out) {
  if (descs == nullptr || out == nullptr || ndom <= 0) return VIBE_C_ERR_BAD_ARG;
  *out = nullptr;
  for (int m = 1; m < ndom; ++m)
    if (descs[m].precision != descs[0].precision) return VIBE_C_ERR_BAD_ARG;
  auto *b = new vibe_mi_batch();
  b->precision = descs[0].precision;
  auto fail = [&](int rc) {
    free_batch(b);
    delete b;
    set_last_error(rc);
    return rc;
  };
  for (int m = 0; m < ndom; ++m) {
    vibe_mi_domain *d = nullptr;
    if (int rc = domain_create(&descs[m], &d)) return fail(rc);
    b->doms.push_back(d);
    b->nzmax = std::max(b->nzmax, d->nz);
    const int gm = d->rk == 1 ? kGeoFlat : ((d->nyr == 1 && d->nxr == d->nx) ? kGeoX : kGeoAny);
    b->gm = std::max(b->gm, gm);
  }
  if (hipGetDevice(&b->device) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  b->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  for (auto &e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  const size_t dsz = b->precision != VIBE_MI_FP64 ? sizeof(BDom<float>) : sizeof(BDom<double>);
  unsigned char *p = nullptr;
  if (int rc = balloc(b, &p, dsz * ndom)) return fail(rc);
  b->desc = p;
  if (int rc = balloc(b, &b->ctr, kMaxCounters)) return fail(rc);
  if (int rc = balloc(b, &b->flag, 1)) return fail(rc);
  if (hipMemset(b->flag, 0, sizeof(int)) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  *out = b;
  return 0;
}

int vibe_mi_batch_destroy(vibe_mi_batch *b) {
  if (b == nullptr) return VIBE_C_ERR_BAD_ARG;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  delete b;
  return 0;
}

int vibe_mi_batch_member(vibe_mi_batch *b, int idx, vibe_mi_domain **out) {
  if (b == nullptr || out == nullptr || idx < 0 || idx >= static_cast<int>(b->doms.size())) return VIBE_C_ERR_BAD_ARG;
  *out = b->doms[idx];
  return 0;
}

static int batch_step_common(vibe_mi_batch *b, const double *dt, int nsub, int nsteps, const unsigned char *active,
                             bool timed, double *ms_total) {
  if (b == nullptr || dt == nullptr || nsteps < 0 || nsub < 1 || 3 * (nsub + 1) + 1 > kMaxLaunches ||
      nsub > kMaxNsub)
    return VIBE_C_ERR_BAD_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return VIBE_MI_ERR_HIP;
  for (auto *d : b->doms)
    if (hipStreamSynchronize(d->stream) != hipSuccess) return VIBE_MI_ERR_HIP;
  const int rc = b->precision != VIBE_MI_FP64 ? step_batch<float>(b, dt, nsub, nsteps, active, timed, ms_total)
                                               : step_batch<double>(b, dt, nsub, nsteps, active, timed, ms_total);
  if (rc) set_last_error(rc);
  return rc;
}

int vibe_mi_batch_step(vibe_mi_batch *b, const double *dt, int nsub, int nsteps, const unsigned char *active) {
  return batch_step_common(b, dt, nsub, nsteps, active, false, nullptr);
}

int vibe_mi_batch_step_timed(vibe_mi_batch *b, const double *dt, int nsub, int nsteps, const unsigned char *active,
                             double *ms_total) {
  return batch_step_common(b, dt, nsub, nsteps, active, true, ms_total);
}

int vibe_mi_batch_set_graph(vibe_mi_batch *b, int use_graph) {
  if (b == nullptr) return VIBE_C_ERR_BAD_ARG;
  b->use_graph = use_graph != 0;
  return 0;
}

int vibe_mi_batch_kernel_times(vibe_mi_batch *b, double *ms_stage, double *ms_substep, int *n_stage,
                               int *n_substep) {
  if (b == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (ms_stage) *ms_stage = b->ms_stage;
  if (ms_substep) *ms_substep = b->ms_sub;
  if (n_stage) *n_stage = b->n_stage;
  if (n_substep) *n_substep = b->n_sub;
  return 0;
}

int vibe_mi_batch_tuning(vibe_mi_batch *b, int *substep_waves, int *stage_rows) {
  if (b == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (substep_waves) *substep_waves = b->nw;
  if (stage_rows) *stage_rows = b->ty;
  return 0;
}

size_t vibe_mi_batch_device_bytes(vibe_mi_batch *b) {
  if (b == nullptr) return 0;
  size_t n = b->bytes;
  for (auto *d : b->doms) n += vibe_mi_domain_device_bytes(d);
  return n;
}

}  // extern "C"
