// batch_kernels.hpp -- device side of the batched step: member descriptors,
// tile work lists, the persistent tile loop and the batched kernels.  The
// FP64 kernels are instantiated in batch.hip; the FP32 ones in
// batch_f32_{flat,x,any}.hip,
// which is compiled with multiply-add contraction (the FP64 path keeps
// -ffp-contract=off so it rounds exactly like the CPU oracle).
#pragma once

#include "domain.hpp"
#include "kernels.hpp"

namespace vibe_mi {

// Device descriptor of one member.
template <typename T>
struct BDom {
  Geo<T> g;
  FieldSet S[3];        // FP64 states; S[cur0] = s^n
  AcousticSet<T> A[2];  // acoustic ping-pong buffers
  HeviTab<T> tab[3];    // HEVI factorisation and coefficients of stage st
  double h[3];  // stage lengths dt/3, dt/2, dt
  T dtau[3];    // substep lengths h/nsub
  T an, ao;     // HEVI off-centring weights
  double dt;
  const double *sponge;
  PhysArgs ph;  // physics rates added to every stage's slow tendency (th null: none)
  int kmin_sponge;
  int cur0;  // rotation index, advanced on the device by kb_advance
  int *flag;
  // compensated FP32 mode (null otherwise): remainders of A[0], A[1] and
  // the FP64 geometry of the substep's mass-flux sums
  AcousticLo L[2];
  const double *rhu64, *rhv64, *r0w64, *sx64, *sy64, *zeta64;
  double rdx64, rdy64;
};

// A tile work list split into one contiguous segment per XCD: seg[x] ..
// seg[x+1] are the tiles workgroups on XCD x take first, so that tiles that
// share halo rows are processed under the same L2.
constexpr int kXcds = 8;
struct TileList {
  const int4 *tiles;
  const int *seg;  // kXcds + 1 segment starts, seg[kXcds] = ntiles
  int ntiles;
};

// Persistent tile loop.  Workgroups are dispatched to the XCDs round-robin
// by workgroup id, so workgroup b serves segment b % 8 first, then steals
// from the other segments in turn until every counter is past its segment.
// One lane claims the next tile while the workgroup works on the current
// one; the claimed index is broadcast through LDS.  Every thread of the
// workgroup runs body(t) for the same t.  ctr: kXcds counters, zero at launch.
template <typename F>
__device__ __forceinline__ void tile_loop(int *__restrict__ ctr, const TileList &L, F &&body) {
  __shared__ int s_next[2];
  const bool leader = threadIdx.x == 0 && threadIdx.y == 0;
  const int home = static_cast<int>(blockIdx.x % kXcds);
  int qk = 0;  // segments exhausted so far (leader only)
  auto claim = [&]() {
    while (qk < kXcds) {
      const int x = (home + qk) % kXcds;
      const int t = L.seg[x] + atomicAdd(ctr + x, 1);
      if (t < L.seg[x + 1]) return t;
      ++qk;
    }
    return L.ntiles;
  };
  if (leader) s_next[0] = claim();
  __syncthreads();
  int par = 0;
  for (;;) {
    const int t = __builtin_amdgcn_readfirstlane(s_next[par]);
    if (t >= L.ntiles) break;
    if (leader) s_next[par ^ 1] = claim();
    body(t);
    __syncthreads();
    par ^= 1;
  }
}

__device__ __forceinline__ int stage_out_index(int cur0, int st) { return (cur0 + (st == 1 ? 2 : 1)) % 3; }

static __global__ void __launch_bounds__(64) kb_reset(int *__restrict__ ctr, int n) {
  for (int q = threadIdx.x; q < n; q += 64) ctr[q] = 0;
}

// Rotates the state buffers of the members in ids[0..n): s^{n+1} = S[cur0+1].
template <typename T>
__global__ void __launch_bounds__(64) kb_advance(BDom<T> *__restrict__ D, const int *__restrict__ ids, int n) {
  for (int q = threadIdx.x; q < n; q += 64) {
    BDom<T> &dd = D[ids[q]];
    dd.cur0 = (dd.cur0 + 1) % 3;
  }
}

// Fused stage kernel over tiles {member, x-block, y-block, -} of 64 x TY
// columns; writes u, v, w, rho', pi' of s0 + h R to the acoustic set
// A[aout] and theta' to the stage output.  CUR32: the current state's
// u, v, w, rho', pi' come from the FP32 acoustic set A[ain] the previous
// stage's substeps ended in (FP32-mixed stages 2 and 3), theta' from the
// FP64 stage output; otherwise all from the FP64 state.  PH: the variant
// with the physics coupling, launched when a member of the batch has rates.
template <typename T, int TY, int GM, bool CUR32, bool PH>
__global__ void __launch_bounds__(64 * TY) kb_stage(const BDom<T> *__restrict__ D, TileList L, int st, int ain,
                                                    int aout, int *__restrict__ ctr) {
  tile_loop(ctr, L, [&](int t) {
    const int4 tl = L.tiles[t];
    const BDom<T> &dd = D[tl.x];
    const Geo<T> g = dd.g;
    const int c0 = dd.cur0;
    const FieldSet cur = dd.S[(c0 + st) % 3], s0 = dd.S[c0], nxt = dd.S[stage_out_index(c0, st)];
    const AcousticSet<T> a = dd.A[aout];
    const StageIn si{s0.u, s0.v, s0.w, s0.rho, s0.th, s0.pi};
    StageOut<T> so{a.u, a.v, a.w, a.rho, a.pi, nxt.th};
    if constexpr (sizeof(T) == 4) {  // remainders, compensated mode only (else null)
      const AcousticLo l = dd.L[aout];
      so.ul = l.u;
      so.vl = cp_keeps(2) ? l.v : nullptr, so.wl = cp_keeps(4) ? l.w : nullptr, so.pl = cp_keeps(8) ? l.pi : nullptr;
    }
    SpongeArgs sp;  // the last stage's theta' is the step's final one
    if (st == 2) sp = SpongeArgs{dd.sponge, dd.dt, dd.kmin_sponge};
    if constexpr (CUR32) {
      const AcousticSet<T> ac = dd.A[ain];
      const StageCur<T> ci{ac.u, ac.v, ac.w, ac.rho, ac.pi, cur.th};
      stage_tile<T, T, TY, GM, T, PH>(g, ci, si, dd.h[st], so, dd.flag, tl.y, tl.z, sp, dd.ph);
    } else {
      const StageCur<double> ci{cur.u, cur.v, cur.w, cur.rho, cur.pi, cur.th};
      stage_tile<T, T, TY, GM, double, PH>(g, ci, si, dd.h[st], so, dd.flag, tl.y, tl.z, sp, dd.ph);
    }
  });
}

// Acoustic substep q of stage st over row tiles, from the acoustic set
// A[(abase+q)&1] to A[(abase+q+1)&1].  The last substep of a stage
// (last = 1) checks finiteness; with O = double it writes the FP64 stage
// output S[out] instead.  CP: the compensated FP32 mode (substep_row).  FP32
// substeps carry chi = rho' - pi'/kappa in the stage's output set
// A[abase&1] in place of rho'; RO (a stage's last substep) writes rho'.
template <typename T, typename O, int NW, int MAXCH, int GM, bool CP = false, bool RO = (sizeof(O) == 8)>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(kSubstepWavesPerEu<T, NW, MAXCH, CP>))) kb_substep(const BDom<T> *__restrict__ D, TileList L, int st, int q,
                                                      int abase, int last, int *__restrict__ ctr) {
  tile_loop(ctr, L, [&](int t) {
    const int4 tl = L.tiles[t];
    const BDom<T> &dd = D[tl.x];
    const Geo<T> g = dd.g;
    const AcousticSet<T> in = dd.A[(abase + q) & 1];
    SubArgs<T, T, O> sa{};
    sa.u = in.u;
    sa.v = in.v;
    sa.w = in.w;
    sa.rho = in.rho;
    sa.pi = in.pi;
    sa.chi = dd.A[abase & 1].rho;
    sa.tab = dd.tab[st];
    sa.dtau = dd.dtau[st];
    sa.an = dd.an;
    sa.ao = dd.ao;
    sa.flag = dd.flag;
    if constexpr (sizeof(O) == 8) {
      if (last) {
        const FieldSet o = dd.S[stage_out_index(dd.cur0, st)];
        sa.uo = o.u;
        sa.vo = o.v;
        sa.wo = o.w;
        sa.rho_o = o.rho;
        sa.pi_o = o.pi;
        sa.check = 1;
        if (st == 2) sa.sp = SpongeArgs{dd.sponge, dd.dt, dd.kmin_sponge};  // the step's final state
      } else if constexpr (sizeof(T) == 8) {
        const AcousticSet<T> o = dd.A[(abase + q + 1) & 1];
        sa.uo = o.u;
        sa.vo = o.v;
        sa.wo = o.w;
        sa.rho_o = o.rho;
        sa.pi_o = o.pi;
      }
    } else {
      const AcousticSet<T> o = dd.A[(abase + q + 1) & 1];
      sa.uo = o.u;
      sa.vo = o.v;
      sa.wo = o.w;
      sa.rho_o = o.rho;
      sa.pi_o = o.pi;
      sa.check = last;
    }
    if constexpr (CP) {
      const AcousticLo li = dd.L[(abase + q) & 1];
      sa.ul = li.u, sa.vl = li.v, sa.wl = li.w, sa.pl = li.pi;
      if constexpr (sizeof(O) == 4) {
        const AcousticLo lo = dd.L[(abase + q + 1) & 1];
        sa.ulo = lo.u, sa.vlo = lo.v, sa.wlo = lo.w, sa.plo = lo.pi;
      }
      sa.rhu64 = dd.rhu64, sa.rhv64 = dd.rhv64, sa.r0w64 = dd.r0w64, sa.sx64 = dd.sx64, sa.sy64 = dd.sy64;
      sa.zeta64 = dd.zeta64, sa.rdx64 = dd.rdx64, sa.rdy64 = dd.rdy64;
    }
    substep_row<T, T, O, NW, MAXCH, GM, CP, kSubstepIncrement<T, T, CP>, RO>(g, sa, tl.y, tl.z);
  });
}

}  // namespace vibe_mi

// Explicit instantiation of every FP32 batched kernel shape (X = extern for a
// declaration in batch.hip, empty for the definitions in batch_f32_*.hip).
#define VMI_BATCH_F32_SUBSTEP_GM_RO(X, O, GM, CP, RO)                                                              \
  X template __global__ void vibe_mi::kb_substep<float, O, 16, 4, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 16, 8, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 8, 4, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                 \
  X template __global__ void vibe_mi::kb_substep<float, O, 8, 8, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                 \
  X template __global__ void vibe_mi::kb_substep<float, O, 8, 12, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 8, 16, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 4, 8, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                 \
  X template __global__ void vibe_mi::kb_substep<float, O, 4, 12, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 4, 16, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 4, 24, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);                \
  X template __global__ void vibe_mi::kb_substep<float, O, 4, 32, GM, CP, RO>(VMI_KB_SUBSTEP_ARGS);
// FP32 outputs: without and with rho' (a stage's last substep); FP64
// outputs always write rho'
#define VMI_BATCH_F32_SUBSTEP_F(X, GM, CP) \
  VMI_BATCH_F32_SUBSTEP_GM_RO(X, float, GM, CP, false) VMI_BATCH_F32_SUBSTEP_GM_RO(X, float, GM, CP, true)
#define VMI_BATCH_F32_SUBSTEP_D(X, GM, CP) VMI_BATCH_F32_SUBSTEP_GM_RO(X, double, GM, CP, true)
#define VMI_BATCH_F32_STAGE_GM(X, GM, PH)                                                                     \
  X template __global__ void vibe_mi::kb_stage<float, 2, GM, false, PH>(VMI_KB_STAGE_ARGS);                    \
  X template __global__ void vibe_mi::kb_stage<float, 4, GM, false, PH>(VMI_KB_STAGE_ARGS);                    \
  X template __global__ void vibe_mi::kb_stage<float, 8, GM, false, PH>(VMI_KB_STAGE_ARGS);                    \
  X template __global__ void vibe_mi::kb_stage<float, 2, GM, true, PH>(VMI_KB_STAGE_ARGS);                     \
  X template __global__ void vibe_mi::kb_stage<float, 4, GM, true, PH>(VMI_KB_STAGE_ARGS);                     \
  X template __global__ void vibe_mi::kb_stage<float, 8, GM, true, PH>(VMI_KB_STAGE_ARGS);
// every substep shape of one geometry compaction (plain and compensated)
#define VMI_BATCH_F32_SUBSTEPS(X, GM)                                                                           \
  VMI_BATCH_F32_SUBSTEP_F(X, GM, false)                                                                         \
  VMI_BATCH_F32_SUBSTEP_D(X, GM, false)                                                                         \
  VMI_BATCH_F32_SUBSTEP_F(X, GM, true)                                                                          \
  VMI_BATCH_F32_SUBSTEP_D(X, GM, true)
// every stage shape of one geometry compaction (without and with physics)
#define VMI_BATCH_F32_STAGES(X, GM) VMI_BATCH_F32_STAGE_GM(X, GM, false) VMI_BATCH_F32_STAGE_GM(X, GM, true)
#define VMI_BATCH_F32_GM(X, GM) VMI_BATCH_F32_SUBSTEPS(X, GM) VMI_BATCH_F32_STAGES(X, GM)
#define VMI_KB_SUBSTEP_ARGS const vibe_mi::BDom<float> *__restrict__, vibe_mi::TileList, int, int, int, int, int *__restrict__
#define VMI_KB_STAGE_ARGS const vibe_mi::BDom<float> *__restrict__, vibe_mi::TileList, int, int, int, int *__restrict__
#define VMI_BATCH_F32_ALL(X) \
  VMI_BATCH_F32_GM(X, vibe_mi::kGeoFlat) VMI_BATCH_F32_GM(X, vibe_mi::kGeoX) VMI_BATCH_F32_GM(X, vibe_mi::kGeoAny)
