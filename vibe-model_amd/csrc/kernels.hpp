// kernels.hpp -- device code of the dynamical-core step on gfx950 (shared by
// the single-domain and the batched host paths).
//
// One WS-RK3 step (src/time/runge_kutta.cpp:77-118 of the reference) is:
//   for each stage h in {dt/3, dt/2, dt}:
//     k_stage_col  contravariant vertical velocity Omega and the slow
//                  tendencies R (equations.cpp:60-150 role, balanced form)
//                  fused with the RK3 update s = s0 + h R (FP64 accumulation),
//                  one thread per column of a 64 x TY tile marching up the
//                  levels with the tile's horizontal neighbourhood in LDS
//     nsub x k_substep
//                  forward-backward acoustic substep: horizontal momentum
//                  explicit, w / pi' / rho' vertically implicit (HEVI), the
//                  Thomas sweep of each column on one lane
//   k_sponge       backward-Euler Rayleigh sponge (damping.cpp:76-138)
// Each kernel is a body function over one tile (a column, a row of 64
// columns) plus a thin __global__ wrapper: the single-domain wrappers map
// blockIdx to a tile, the batched wrappers (batch.hip) take tiles of many
// domains from a device work list.
//
// Discrete equations (identical to the FP64 CPU oracle term by term):
//   geometry  D = H - zs, Du/Dv at faces, slopes sx/sy at faces,
//             normalised zeta levels, dzeta = layer thickness, dzw = centre
//             spacing; Gal-Chen metric: z_x|zeta = (1 - zeta) * sx.
//   Omega     w - (1 - zeta_k) (Mc(k-1) + Mc(k))/2 at interior interfaces,
//             Mc = cell-centred u.grad(zs); Omega = 0 at ground and lid.
//   u, v, w   flux-form advective 5th-order upwind transport (D-weighted
//             mass fluxes), - cp theta' grad(pi'), + g theta'/theta0 on w.
//   rho'      - (1/D) div(rho' V) (second-order centred faces).
//   theta'    upwind WENO5 flux form (one reconstruction per face, picked by
//             the sign of the face mass flux) minus theta' div V,
//             - w dtheta0/dz.
//   pi'       kappa R_rho + lambda R_theta (EOS linearised about the
//             reference state: kappa = (Rd/cv) pi0/rho0, lambda = (Rd/cv)
//             pi0/theta0).
//   substep   u -= dtau cp theta0_u PX(pi'), v likewise; then per column
//             rho' += -(dtau/D)[div_h(D rho0 V) + d_zeta(rho0 Omega~)],
//             pi' += kappa * delta rho', w with off-centred implicit
//             vertical pressure gradient; Omega~ = an w_new + ao w_old - M.
//
// Precision: VIBE_MI_FP64 runs every kernel in double.  VIBE_MI_MIXED keeps
// the FP64 state (s0, stage outputs) and does the RK3 accumulation
// s0 + h*R in FP64, computes tendencies and substeps in FP32 on the vector
// ALUs, and sums mass/energy with FP64 compensated reductions.
//
// Compiled with -ffp-contract=off so the FP64 path rounds like the oracle.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "geo.hpp"
#include "stencils.hpp"
#include "domain.hpp"

namespace vibe_mi {

// ---- fused z-marching stage kernel -------------------------------------------
// A workgroup owns a 64 x TY tile of columns, one thread per column, marching
// up the levels.  Every vertical stencil (5th-order momentum, WENO5 theta',
// pressure gradient, Omega averages) reads a register window of the thread's
// own column; horizontal neighbours come from the level's u, v, w, theta',
// rho' staged in LDS with a 3-cell halo.  Omega is computed in place from the
// staged next level (the slope flux Mc of the own, west and south columns is
// carried up), so no separate Omega pass or array exists.  Writes u, v, w,
// rho', pi' (type O) and theta' (FP64) of s = s0 + h R in one pass.  Each
// interface flux is evaluated once: the flux through the top of level k is
// carried to level k+1 as its bottom flux (the same expression on the same
// operands).
struct StageIn {
  const double *u, *v, *w, *rho, *th, *pi;
};
// Current state s^(k) read by a stage: u, v, w, rho', pi' of type CI (FP64
// state, or in FP32-mixed batches the FP32 acoustic set the previous stage's
// substeps ended in, whose u holds u - ubg), theta' always FP64.
template <typename CI>
struct StageCur {
  const CI *u, *v, *w, *rho, *pi;
  const double *th;
};
__host__ __device__ inline StageCur<double> stage_cur(const StageIn &s) {
  return StageCur<double>{s.u, s.v, s.w, s.rho, s.pi, s.th};
}
template <typename O>
struct StageOut {
  O *u, *v, *w, *rho, *pi;
  double *th;
  // compensated FP32 mode (FP32 sets only): remainders of u (8-bit) and v, w,
  // pi' (bf16) (nullptr: plain FP32 or FP64 output); rho' (chi) has none
  int8_t *ul = nullptr;
  uint16_t *vl = nullptr, *wl = nullptr, *pl = nullptr;
};

// ---- compensated FP32 storage -----------------------------------------------
// An acoustic-set value is kept as an FP32 value hi and a remainder lo: for
// v, w and pi' (when the remainder mask names them) the upper 16 bits
// (bfloat16, rounded to nearest even) of its FP32 pattern, hi + lo holding
// about 32 significant bits; for u 8 bits (r8_bits below).  Substeps add FP32
// increments with comp_add, which recovers the rounding error of hi + d
// exactly (Knuth's two-sum) and carries it into the remainder, so no
// increment is lost against a large value.
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return static_cast<uint16_t>((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }
struct F2 {
  float hi, lo;
};
__device__ __forceinline__ F2 comp_add(float hi, float lo, float d) {
#pragma clang fp contract(off)
  const float s = hi + d;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (d - bb);
  const float e = err + lo;
  const float h2 = s + e;
  return F2{h2, e - (h2 - s)};
}
// The 1/kappa of chi = rho' - pi'/kappa (FP32 acoustic sets): the stage and
// the stage's last substep must use the same value, not the exact one; the
// hardware reciprocal of the FP32 kappa (within 1 ulp) leaves rho' = chi +
// pi' r off by (kappa r - 1) times the stage's change of rho', about 1e-7 of
// it, far below the FP32 rounding of rho' itself.
__device__ __forceinline__ double chi_rkap(float kap) { return static_cast<double>(__builtin_amdgcn_rcpf(kap)); }
// u keeps an 8-bit remainder instead: with hi = m 2^e (0.5 <= |m| < 1), the
// remainder lo of a round-to-nearest hi satisfies |lo| <= 2^(e-25) and is
// stored as q = rint(lo 2^(32-e)), |q| <= 128 (clamped to 127), so hi + lo
// keeps about 31 significant bits in 5 bytes.
__device__ __forceinline__ int8_t r8_bits(float hi, float lo) {
  const int e = __builtin_amdgcn_frexp_expf(hi);
  const float q = __builtin_rintf(__builtin_ldexpf(lo, 32 - e));
  return static_cast<int8_t>(fminf(fmaxf(q, -127.0f), 127.0f));
}
__device__ __forceinline__ float r8_val(float hi, int q) {
  return __builtin_ldexpf(static_cast<float>(q), __builtin_amdgcn_frexp_expf(hi) - 32);
}
// Writes x to an acoustic-set array of type O at byte offset off; an FP32
// array with a remainder array also gets the bf16 remainder x - hi.
template <typename O>
__device__ __forceinline__ void put_acoustic(O *p, uint16_t *lo, uint32_t off, double x) {
  if constexpr (sizeof(O) == 4) {
    const float hi = static_cast<float>(x);
    gst(p, off, hi);
    if (lo != nullptr) gst(lo, off >> 1, bf16_bits(static_cast<float>(x - static_cast<double>(hi))));
  } else {
    gst(p, off, static_cast<O>(x));
  }
}
// the same with an 8-bit remainder (u)
template <typename O>
__device__ __forceinline__ void put_acoustic(O *p, int8_t *lo, uint32_t off, double x) {
  if constexpr (sizeof(O) == 4) {
    const float hi = static_cast<float>(x);
    gst(p, off, hi);
    if (lo != nullptr) gst(lo, off >> 2, r8_bits(hi, static_cast<float>(x - static_cast<double>(hi))));
  } else {
    gst(p, off, static_cast<O>(x));
  }
}

// Rayleigh sponge folded into the writes of a step's final state
// (damping.cpp:76-138 with perturbation targets 0): level k >= kmin is
// scaled by 1 / (1 + dt alpha(min(k, nz-1))), the factor sponge_at applies
// (alpha = 0 gives exactly 1).  kmin beyond nz: no sponge.
struct SpongeArgs {
  const double *alpha = nullptr;
  double dt = 0.0;
  int kmin = 0x7fffffff;
};
// Physics rates added to the slow tendency R of every stage before the
// substeps (the coupling of src/time/runge_kutta.cpp:46-63): theta' in K/s,
// u and v in m/s^2, FP64 in the state layout.  th == nullptr: no physics.
// The heating enters R_theta before R_pi' is formed from R_rho' and R_theta
// (the linearised equation of state), so pi' follows it.
struct PhysArgs {
  const double *th = nullptr, *u = nullptr, *v = nullptr;
};

__device__ __forceinline__ double sponge_scale(const SpongeArgs &sp, int k, int nz, double x) {
  if (k < sp.kmin) return x;
  const double a = C(sp.alpha)[k < nz - 1 ? k : nz - 1];
  const double f = 1.0 / (1.0 + sp.dt * a);
  return f * x;
}

// PH: the physics coupling is compiled in (a member without rates passes
// th == nullptr); without it the stage kernel carries no trace of it.
template <typename T, typename O, int TY, int GM, typename CI = double, bool PH = false>
__device__ __forceinline__ void stage_tile(const Geo<T> &g, const StageCur<CI> &cur, const StageIn &s0, double h,
                                           const StageOut<O> &out, int *__restrict__ flag, int bx, int by,
                                           const SpongeArgs &sp = SpongeArgs{}, const PhysArgs &ph = PhysArgs{}) {
  // levels k and k+1 of u, v, w, theta', rho' over the tile and a 3-cell
  // halo, in a ring of three LDS buffers (level k+2 is filled meanwhile);
  // level k+1 gives Omega at the top face of level k
  constexpr int HX = 64 + 6, HY = TY + 6, AREA = HX * HY, NT = 64 * TY;
  constexpr int NSLOT = (AREA + NT - 1) / NT;
  constexpr int FU = 0, FV = 1, FW = 2, FT = 3, FR = 4;
  __shared__ T s_h[3][5][AREA];
  const int lx = static_cast<int>(threadIdx.x), ly = static_cast<int>(threadIdx.y), tid = ly * 64 + lx;
  const int nx = g.nx, ny = g.ny, nz = g.nz, rk = g.rk;
  const size_t nxy = static_cast<size_t>(nx) * ny;
  const int i_raw = bx * 64 + lx, j_raw = by * TY + ly;
  const bool act = i_raw < nx && j_raw < ny;  // threads outside the domain only stage the halo
  const int i = i_raw < nx ? i_raw : nx - 1, j = j_raw < ny ? j_raw : ny - 1;
  const CI *__restrict__ U = cur.u;
  const CI *__restrict__ V = cur.v;
  const CI *__restrict__ W = cur.w;
  const CI *__restrict__ RH = cur.rho;
  const double *__restrict__ TH = cur.th;
  const CI *__restrict__ PI = cur.pi;
  // FP32 current state: its u holds u - ubg (the background is added back)
  constexpr bool cur32 = sizeof(CI) == 4;
  constexpr uint32_t SC = sizeof(CI);
  auto ubg_at = [&](int k) { return C(g.ubg)[k]; };
  // element offsets (within a level) of this thread's halo slots
  uint32_t hoff[NSLOT];
#pragma unroll
  for (int q = 0; q < NSLOT; ++q) {
    const int e = tid + q * NT < AREA ? tid + q * NT : AREA - 1;
    const int hx = e % HX, hy = e / HX;
    const int gi = ((bx * 64 - 3 + hx) % nx + nx) % nx, gj = ((by * TY - 3 + hy) % ny + ny) % ny;
    hoff[q] = static_cast<uint32_t>(gj) * static_cast<uint32_t>(nx) + static_cast<uint32_t>(gi);
  }
  // level strides in bytes of the FP64 fields and of the current-state
  // fields (fields stay below 4 GiB)
  const uint32_t lsD = static_cast<uint32_t>(nxy) * 8u, lsC = static_cast<uint32_t>(nxy) * SC;
  // the halo values of the next level are loaded into registers before the
  // current level's arithmetic and written to the other buffer after it;
  // they keep their storage type until the write (a conversion right after
  // the load would wait for the load at once)
  CI pre[4][NSLOT];  // u, v, w, rho'
  double preT[NSLOT];  // theta'
  T pre_ubg = T(0);
  auto load_level = [&](int k) {
    const uint32_t b = static_cast<uint32_t>(k) * lsC, bt = static_cast<uint32_t>(k) * lsD;
    const CI *src[4] = {U, V, W, RH};
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int q = 0; q < NSLOT; ++q) pre[f][q] = gld(src[f], hoff[q] * SC + b);
#pragma unroll
    for (int q = 0; q < NSLOT; ++q) preT[q] = gld(TH, hoff[q] * 8u + bt);
    if constexpr (cur32) pre_ubg = ubg_at(k);
    (void)pre_ubg;
  };
  auto store_level = [&](int buf) {
    constexpr int dst[4] = {FU, FV, FW, FR};
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int q = 0; q < NSLOT; ++q)
        if (q + 1 < NSLOT || tid + q * NT < AREA)
          s_h[buf][dst[f]][tid + q * NT] = (f == 0 && cur32) ? T(pre[f][q]) + pre_ubg : T(pre[f][q]);
#pragma unroll
    for (int q = 0; q < NSLOT; ++q)
      if (q + 1 < NSLOT || tid + q * NT < AREA) s_h[buf][FT][tid + q * NT] = T(preT[q]);
  };
  const int lc = (ly + 3) * HX + lx + 3;  // own column in the halo tile
  const uint32_t unx = static_cast<uint32_t>(nx);
  const uint32_t cC = static_cast<uint32_t>(j) * unx + static_cast<uint32_t>(i);
  const uint32_t cW = static_cast<uint32_t>(j) * unx + static_cast<uint32_t>(wrapi(i - 1, nx));
  const uint32_t cS = static_cast<uint32_t>(wrapi(j - 1, ny)) * unx + static_cast<uint32_t>(i);
  // column geometry (compacted)
  const int gC = gcolm<GM>(g, i, j), gW = gcolm<GM>(g, i - 1, j), gE = gcolm<GM>(g, i + 1, j),
            gS = gcolm<GM>(g, i, j - 1), gN = gcolm<GM>(g, i, j + 1), gWN = gcolm<GM>(g, i - 1, j + 1),
            gES = gcolm<GM>(g, i + 1, j - 1);
  constexpr bool xslope = GM != kGeoFlat, yslope = GM == kGeoAny;  // nonzero slope / metric terms
  const T DuW = C(g.Du)[gW], DuC = C(g.Du)[gC], DuE = C(g.Du)[gE], DuES = C(g.Du)[gES];
  const T DvS = C(g.Dv)[gS], DvC = C(g.Dv)[gC], DvN = C(g.Dv)[gN], DvWN = C(g.Dv)[gWN], DvW = C(g.Dv)[gW];
  const T rDuC = C(g.rDu)[gC], rDvC = C(g.rDv)[gC], rD = C(g.rDc)[gC], mxC = C(g.metx)[gC], myC = C(g.mety)[gC];
  const T DuS = C(g.Du)[gS];
  // slopes of the columns whose Omega the tile needs (own, west, south)
  const T sxC = C(g.sx)[gC], sxE = C(g.sx)[gE], sxW = C(g.sx)[gW], sxS = C(g.sx)[gS], sxES = C(g.sx)[gES];
  const T syC = C(g.sy)[gC], syN = C(g.sy)[gN], syW = C(g.sy)[gW], syWN = C(g.sy)[gWN], syS = C(g.sy)[gS];

  const uint32_t cCb = cC * 8u, cWb = cW * 8u, cSb = cS * 8u, oCb = cC * static_cast<uint32_t>(sizeof(O));
  const uint32_t lsO = static_cast<uint32_t>(nxy) * static_cast<uint32_t>(sizeof(O));
  const uint32_t tCb = static_cast<uint32_t>(gC) * static_cast<uint32_t>(sizeof(T)),
                 lsT = static_cast<uint32_t>(rk) * static_cast<uint32_t>(sizeof(T));
  auto ldd = [&](const double *f, int k, uint32_t cb) { return T(gld(f, cb + static_cast<uint32_t>(k) * lsD)); };
  // current-state fields at a level (cb: the column's FP64 byte offset);
  // ldu adds the background wind back to an FP32 u
  auto ldc = [&](const CI *f, int k, uint32_t cb) {
    return T(gld(f, (cb / 8u) * SC + static_cast<uint32_t>(k) * lsC));
  };
  auto ldu = [&](int k, uint32_t cb) {
    if constexpr (cur32)
      return ldc(U, k, cb) + ubg_at(k);
    else
      return ldc(U, k, cb);
  };
  // s0 of the level being updated, loaded one level ahead
  double S0[6];
  auto ld_s0 = [&](int k) {
    const uint32_t b = cCb + static_cast<uint32_t>(k) * lsD;
    S0[0] = gld(s0.u, b);
    S0[1] = gld(s0.v, b);
    S0[2] = gld(s0.w, b);
    S0[3] = gld(s0.rho, b);
    S0[4] = gld(s0.th, b);
    S0[5] = gld(s0.pi, b);
  };
  // reference theta0 of the own column at levels k-1, k, k+1 (clamped), the
  // static-stability factor gamma of interface k, kappa and lambda of level
  // k; the next level's values are loaded one level ahead
  auto th0 = [&](int kq) { return gldc(g.theta0, tCb + static_cast<uint32_t>(clampk(kq, nz - 1)) * lsT); };
  T thm = T(0), th00 = th0(0), thp = th0(1);
  T gamk = (th0(1) - th0(0)) * rD * C(g.rdzw)[1];  // gamma(0) = gamma(1): interface index clamped to 1..nz-1
  T kapk = gldc(g.kap, tCb), lamk = gldc(g.lam, tCb);
  T uw[7], vw[7], ww[7], tw[7], rw[3], pw[3], pWw[3], pSw[3];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int kc = clampk(q - 3, nz - 1);
    uw[q] = ldu(kc, cCb);
    vw[q] = ldc(V, kc, cCb);
    tw[q] = ldd(TH, kc, cCb);
    ww[q] = ldc(W, clampk(q - 3, nz), cCb);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int kc = clampk(q - 1, nz - 1);
    rw[q] = ldc(RH, kc, cCb);
    pw[q] = ldc(PI, kc, cCb);
    pWw[q] = ldc(PI, kc, cWb);
    pSw[q] = ldc(PI, kc, cSb);
  }
  // terrain slope flux Mc (cell centres) of the own, west and south columns
  // at the staged level in buffer `buf` (the Omega kernel's expression)
  T mcC = T(0), mcW = T(0), mcS = T(0);
  auto slope_flux = [&](int buf, T &mc_c, T &mc_w, T &mc_s) {
    const T *__restrict__ Hu = s_h[buf][FU] + lc;
    const T *__restrict__ Hv = s_h[buf][FV] + lc;
    if constexpr (!xslope) {
      mc_c = mc_w = mc_s = T(0);
    } else if constexpr (!yslope) {
      mc_c = T(0.5) * (Hu[0] * sxC + Hu[1] * sxE);
      mc_w = T(0.5) * (Hu[-1] * sxW + Hu[0] * sxC);
      mc_s = T(0.5) * (Hu[-HX] * sxS + Hu[1 - HX] * sxES);
    } else {
      mc_c = T(0.5) * (Hu[0] * sxC + Hu[1] * sxE) + T(0.5) * (Hv[0] * syC + Hv[HX] * syN);
      mc_w = T(0.5) * (Hu[-1] * sxW + Hu[0] * sxC) + T(0.5) * (Hv[-1] * syW + Hv[HX - 1] * syWN);
      mc_s = T(0.5) * (Hu[-HX] * sxS + Hu[1 - HX] * sxES) + T(0.5) * (Hv[-HX] * syS + Hv[0] * syC);
    }
  };
  T omC0 = T(0);  // Omega at the bottom face of the current level (0 at the ground)
  T uE_prev = T(0), vN_prev = T(0);
  // z-face fluxes of the level above, reused as the next level's lower face
  T cu_O = T(0), cu_f = T(0), cv_O = T(0), cv_f = T(0), cw_O = T(0), cw_f = T(0);
  T cm_z = T(0), cr_z = T(0), ct_z = T(0);
  int bad = 0;
  if (act) put_acoustic(out.w, out.wl, oCb, gld(s0.w, cCb));  // interface 0 keeps its value (R_w = 0)
  ld_s0(0);

  load_level(0);
  store_level(0);
  if (nz > 1) {
    load_level(1);
    store_level(1);
  }
  __syncthreads();
  slope_flux(0, mcC, mcW, mcS);
  for (int k = 0; k < nz; ++k) {
    const uint32_t ko = oCb + static_cast<uint32_t>(k) * lsO;
    const int buf = k % 3;
    // levels k and k+1 staged; level k-1's reads of buffer (k+2)%3 are done
    __syncthreads();
    if (k + 2 < nz) load_level(k + 2);
    // Omega at interface k+1 of the own, west and south columns (0 at the lid)
    T ow[3], oW[2], oS[2];
    ow[1] = omC0;
    ow[2] = oW[1] = oS[1] = T(0);
    if (k + 1 < nz) {
      const int b1 = (k + 1) % 3;
      T m1C, m1W, m1S;
      slope_flux(b1, m1C, m1W, m1S);
      const T zf = T(1) - C(g.zeta)[k + 1];
      const T *__restrict__ Hw1 = s_h[b1][FW] + lc;
      ow[2] = Hw1[0] - zf * (T(0.5) * (mcC + m1C));
      oW[1] = Hw1[-1] - zf * (T(0.5) * (mcW + m1W));
      oS[1] = Hw1[-HX] - zf * (T(0.5) * (mcS + m1S));
      mcC = m1C;
      mcW = m1W;
      mcS = m1S;
    }
    omC0 = ow[2];
    // horizontal neighbours at level k, from the staged tile
    const T *__restrict__ Hu = s_h[buf][FU] + lc;
    const T *__restrict__ Hv = s_h[buf][FV] + lc;
    const T *__restrict__ Hw = s_h[buf][FW] + lc;
    const T *__restrict__ Ht = s_h[buf][FT] + lc;
    const T *__restrict__ Hr = s_h[buf][FR] + lc;
    T ux[7], uy[7], vx[7], vy[7], tx[7], ty[7], wx[7], wy[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) {
      ux[d] = Hu[d - 3];
      uy[d] = Hu[(d - 3) * HX];
      vx[d] = Hv[d - 3];
      vy[d] = Hv[(d - 3) * HX];
      tx[d] = Ht[d - 3];
      ty[d] = Ht[(d - 3) * HX];
      wx[d] = Hw[d - 3];
      wy[d] = Hw[(d - 3) * HX];
    }
    const T vWN = Hv[HX - 1], uES = Hu[1 - HX];
    const T rWst = Hr[-1], rE = Hr[1], rSth = Hr[-HX], rN = Hr[HX];
    const T omz = T(1) - C(g.zc)[k];
    const T rdzk = C(g.rdzeta)[k];
    // vertical pi' gradients (0 at ground and lid)
    auto gpair = [&](const T *p) {
      const T glo = (k >= 1) ? (p[1] - p[0]) * C(g.rdzw)[k] : T(0);
      const T ghi = (k + 1 <= nz - 1) ? (p[2] - p[1]) * C(g.rdzw)[k + 1] : T(0);
      return T(0.5) * (glo + ghi);
    };
    const T GcC = xslope ? gpair(pw) : T(0), GcW = xslope ? gpair(pWw) : T(0), GcS = yslope ? gpair(pSw) : T(0);

    // ---------------- u at the west face ----------------
    {
      const T U0 = T(0.5) * (DuW * ux[2] + DuC * ux[3]);
      const T U1 = T(0.5) * (DuC * ux[3] + DuE * ux[4]);
      const T fx0 = up5<T>(U0, ux[0], ux[1], ux[2], ux[3], ux[4], ux[5]);
      const T fx1 = up5<T>(U1, ux[1], ux[2], ux[3], ux[4], ux[5], ux[6]);
      const T V0 = T(0.5) * (DvW * vx[2] + DvC * vw[3]);
      const T V1 = T(0.5) * (DvWN * vWN + DvN * vy[4]);
      const T fy0 = up5<T>(V0, uy[0], uy[1], uy[2], uy[3], uy[4], uy[5]);
      const T fy1 = up5<T>(V1, uy[1], uy[2], uy[3], uy[4], uy[5], uy[6]);
      const T O0 = cu_O, fz0 = cu_f;  // interface k (0 at the ground)
      T O1 = T(0), fz1 = T(0);
      if (k + 1 < nz) {
        O1 = T(0.5) * (oW[1] + ow[2]);
        fz1 = up5<T>(O1, uw[1], uw[2], uw[3], uw[4], uw[5], uw[6]);
      }
      cu_O = O1;
      cu_f = fz1;
      const T adv = -(fx1 - fx0) * g.rdx - (fy1 - fy0) * g.rdy - (fz1 - fz0) * rdzk +
                    uw[3] * ((U1 - U0) * g.rdx + (V1 - V0) * g.rdy + (O1 - O0) * rdzk);
      const T thu = T(0.5) * (tx[2] + tw[3]);
      T px = (pw[1] - pWw[1]) * g.rdx;
      if constexpr (xslope) px = px - (omz * mxC) * (T(0.5) * (GcW + GcC));
      const T R = adv * rDuC - T(kCp) * thu * px;
      double Rd = static_cast<double>(R);
      if constexpr (PH)
        if (ph.th != nullptr) Rd = Rd + gld(ph.u, cCb + static_cast<uint32_t>(k) * lsD);
      double un = S0[0] + h * Rd;
      if constexpr (sizeof(T) == 4) un = un - C(g.ubg64)[k];  // FP32 acoustic buffers hold u - ubg
      if (act) put_acoustic(out.u, out.ul, ko, un);
    }
    // ---------------- v at the south face ----------------
    {
      const T U0 = T(0.5) * (DuS * uy[2] + DuC * uw[3]);
      const T U1 = T(0.5) * (DuES * uES + DuE * ux[4]);
      const T fx0 = up5<T>(U0, vx[0], vx[1], vx[2], vx[3], vx[4], vx[5]);
      const T fx1 = up5<T>(U1, vx[1], vx[2], vx[3], vx[4], vx[5], vx[6]);
      const T V0 = T(0.5) * (DvS * vy[2] + DvC * vy[3]);
      const T V1 = T(0.5) * (DvC * vy[3] + DvN * vy[4]);
      const T fy0 = up5<T>(V0, vy[0], vy[1], vy[2], vy[3], vy[4], vy[5]);
      const T fy1 = up5<T>(V1, vy[1], vy[2], vy[3], vy[4], vy[5], vy[6]);
      const T O0 = cv_O, fz0 = cv_f;
      T O1 = T(0), fz1 = T(0);
      if (k + 1 < nz) {
        O1 = T(0.5) * (oS[1] + ow[2]);
        fz1 = up5<T>(O1, vw[1], vw[2], vw[3], vw[4], vw[5], vw[6]);
      }
      cv_O = O1;
      cv_f = fz1;
      const T adv = -(fx1 - fx0) * g.rdx - (fy1 - fy0) * g.rdy - (fz1 - fz0) * rdzk +
                    vw[3] * ((U1 - U0) * g.rdx + (V1 - V0) * g.rdy + (O1 - O0) * rdzk);
      const T thv = T(0.5) * (ty[2] + tw[3]);
      T py = (pw[1] - pSw[1]) * g.rdy;
      if constexpr (yslope) py = py - (omz * myC) * (T(0.5) * (GcS + GcC));
      const T R = adv * rDvC - T(kCp) * thv * py;
      double Rd = static_cast<double>(R);
      if constexpr (PH)
        if (ph.th != nullptr) Rd = Rd + gld(ph.v, cCb + static_cast<uint32_t>(k) * lsD);
      if (act) put_acoustic(out.v, out.vl, ko, S0[1] + h * Rd);
    }
    // ---------------- w at interface k (1..nz-1) ----------------
    // flux of w through the cell centre below / above interface k
    const T Oc1 = T(0.5) * (ow[1] + ow[2]);
    const T fzc1 = up5<T>(Oc1, ww[1], ww[2], ww[3], ww[4], ww[5], ww[6]);
    if (k >= 1) {
      const T U0 = T(0.5) * (DuC * uw[2] + DuC * uw[3]);
      const T U1 = T(0.5) * (DuE * uE_prev + DuE * ux[4]);
      const T fx0 = up5<T>(U0, wx[0], wx[1], wx[2], wx[3], wx[4], wx[5]);
      const T fx1 = up5<T>(U1, wx[1], wx[2], wx[3], wx[4], wx[5], wx[6]);
      const T V0 = T(0.5) * (DvC * vw[2] + DvC * vw[3]);
      const T V1 = T(0.5) * (DvN * vN_prev + DvN * vy[4]);
      const T fy0 = up5<T>(V0, wy[0], wy[1], wy[2], wy[3], wy[4], wy[5]);
      const T fy1 = up5<T>(V1, wy[1], wy[2], wy[3], wy[4], wy[5], wy[6]);
      const T Oc0 = cw_O, fz0 = cw_f;
      const T fz1 = fzc1;
      const T rz = C(g.rdzw)[k];
      const T adv = -(fx1 - fx0) * g.rdx - (fy1 - fy0) * g.rdy - (fz1 - fz0) * rz +
                    ww[3] * ((U1 - U0) * g.rdx + (V1 - V0) * g.rdy + (Oc1 - Oc0) * rz);
      const T thw = T(0.5) * (tw[2] + tw[3]);
      const T th0w = T(0.5) * (thm + th00);
      const T pz = (pw[1] - pw[0]) * rz * rD;
      const T R = adv * rD - T(kCp) * thw * pz + quot<T>(T(kG) * thw, th0w);
      if (act) put_acoustic(out.w, out.wl, ko, S0[2] + h * static_cast<double>(R));
    }
    cw_O = Oc1;
    cw_f = fzc1;
    // ---------------- cell: rho', theta', pi' ----------------
    {
      const T mfx0 = DuC * uw[3], mfx1 = DuE * ux[4];
      const T frx0 = DuC * (T(0.5) * (rWst + rw[1])) * uw[3];
      const T frx1 = DuE * (T(0.5) * (rw[1] + rE)) * ux[4];
      const T ftx0 = DuC * upwind_weno<T>(uw[3], tx[0], tx[1], tx[2], tx[3], tx[4], tx[5]);
      const T ftx1 = DuE * upwind_weno<T>(ux[4], tx[1], tx[2], tx[3], tx[4], tx[5], tx[6]);
      const T mfy0 = DvC * vw[3], mfy1 = DvN * vy[4];
      const T fry0 = DvC * (T(0.5) * (rSth + rw[1])) * vw[3];
      const T fry1 = DvN * (T(0.5) * (rw[1] + rN)) * vy[4];
      const T fty0 = DvC * upwind_weno<T>(vw[3], ty[0], ty[1], ty[2], ty[3], ty[4], ty[5]);
      const T fty1 = DvN * upwind_weno<T>(vy[4], ty[1], ty[2], ty[3], ty[4], ty[5], ty[6]);
      const T mfz0 = cm_z, frz0 = cr_z, ftz0 = ct_z;
      T mfz1 = T(0), frz1 = T(0), ftz1 = T(0);
      if (k + 1 < nz) {
        mfz1 = ow[2];
        frz1 = (T(0.5) * (rw[1] + rw[2])) * ow[2];
        ftz1 = upwind_weno<T>(ow[2], tw[1], tw[2], tw[3], tw[4], tw[5], tw[6]);
      }
      cm_z = mfz1;
      cr_z = frz1;
      ct_z = ftz1;
      const T rr = -((frx1 - frx0) * g.rdx + (fry1 - fry0) * g.rdy + (frz1 - frz0) * rdzk) * rD;
      const T div = (mfx1 - mfx0) * g.rdx + (mfy1 - mfy0) * g.rdy + (mfz1 - mfz0) * rdzk;
      const T fd = (ftx1 - ftx0) * g.rdx + (fty1 - fty0) * g.rdy + (ftz1 - ftz0) * rdzk;
      // gamma(k+1) with the interface index clamped to 1..nz-1
      const T gam1 = (k + 1 <= nz - 1) ? (thp - th00) * rD * C(g.rdzw)[k + 1] : gamk;
      T rt = -(fd - tw[3] * div) * rD - T(0.5) * (ww[3] * gamk + ww[4] * gam1);
      if constexpr (PH)
        if (ph.th != nullptr) rt = rt + static_cast<T>(gld(ph.th, cCb + static_cast<uint32_t>(k) * lsD));
      const T rp = kapk * rr + lamk * rt;
      const double rn = S0[3] + h * static_cast<double>(rr);
      const double tn = S0[4] + h * static_cast<double>(rt);
      const double pn = S0[5] + h * static_cast<double>(rp);
      if (act) {
        // FP32 acoustic sets carry chi = rho' - pi'/kappa in place of rho':
        // the substeps change rho' and pi' only together (d pi' = kappa
        // d rho' per cell), so chi is constant over a stage's substeps and
        // the stage's last substep recovers rho' = chi + pi'/kappa
        if constexpr (sizeof(O) == 4)
          put_acoustic(out.rho, static_cast<uint16_t *>(nullptr), ko, rn - pn * chi_rkap(kapk));
        else
          put_acoustic(out.rho, static_cast<uint16_t *>(nullptr), ko, rn);
        gst(out.th, cCb + static_cast<uint32_t>(k) * lsD, sponge_scale(sp, k, nz, tn));
        put_acoustic(out.pi, out.pl, ko, pn);
        bad |= !isfinite(tn);
      }
      gamk = gam1;
    }
    // ---------------- advance the vertical windows ----------------
    uE_prev = ux[4];
    vN_prev = vy[4];
    const int kn = clampk(k + 4, nz - 1), kw = clampk(k + 4, nz);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      uw[q] = uw[q + 1];
      vw[q] = vw[q + 1];
      tw[q] = tw[q + 1];
      ww[q] = ww[q + 1];
    }
    uw[6] = ldu(kn, cCb);
    vw[6] = ldc(V, kn, cCb);
    tw[6] = ldd(TH, kn, cCb);
    ww[6] = ldc(W, kw, cCb);
    const int k2 = clampk(k + 2, nz - 1);
    rw[0] = rw[1];
    rw[1] = rw[2];
    rw[2] = ldc(RH, k2, cCb);
    pw[0] = pw[1];
    pw[1] = pw[2];
    pw[2] = ldc(PI, k2, cCb);
    pWw[0] = pWw[1];
    pWw[1] = pWw[2];
    pWw[2] = ldc(PI, k2, cWb);
    pSw[0] = pSw[1];
    pSw[1] = pSw[2];
    pSw[2] = ldc(PI, k2, cSb);
    // operands of level k+1 from s0 and the reference state (unconditional,
    // clamped: a load under a branch would be waited for where it joins)
    {
      const int k1 = min(k + 1, nz - 1);
      ld_s0(k1);
      thm = th00;
      th00 = thp;
      thp = th0(k + 2);
      kapk = gldc(g.kap, tCb + static_cast<uint32_t>(k1) * lsT);
      lamk = gldc(g.lam, tCb + static_cast<uint32_t>(k1) * lsT);
    }
    if (k + 2 < nz) store_level((k + 2) % 3);
  }
  if (act) put_acoustic(out.w, out.wl, oCb + static_cast<uint32_t>(nz) * lsO, gld(s0.w, cCb + static_cast<uint32_t>(nz) * lsD));
  if (bad) atomicOr(flag, 1);
}

template <typename T, typename O, int TY, bool PH>
__global__ void __launch_bounds__(64 * TY) k_stage_col(Geo<T> g, StageIn cur, StageIn s0, double h, StageOut<O> out,
                                                        int *__restrict__ flag, PhysArgs ph) {
  stage_tile<T, O, TY, kGeoAny, double, PH>(g, stage_cur(cur), s0, h, out, flag, blockIdx.x, blockIdx.y,
                                            SpongeArgs{}, ph);
}

// ---- HEVI factorisation table ----------------------------------------------
// The tridiagonal matrix of the vertically implicit w system depends only on
// the reference state, the column depth and dtau, so it is factorised once
// per (stage, column): beta_m = a_m/pivot_m, 1/pivot_m and c'_m, stored
// [m][compact column].  One thread per compact column.
template <typename T>
__global__ void __launch_bounds__(64) k_hevi_factor(Geo<T> g, T dtau, T an, T *__restrict__ tab) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.rk) return;
  const int nz = g.nz, n = nz - 1, rk = g.rk;
  const size_t per = hevi_per(nz, rk);
  T *__restrict__ tA = tab;
  T *__restrict__ tINV = tab + per;
  T *__restrict__ tCP = tab + 2 * per;
  T *__restrict__ tGK = tab + 3 * per;
  T *__restrict__ tCK = tab + 4 * per;
  const T dtD = dtau * C(g.rDc)[c];
  const T an2 = an * an;
  auto ck = [&](int k) { return C(g.kap)[k * rk + c] * dtD * C(g.rdzeta)[k]; };
  auto r0w = [&](int k) { return C(g.r0w)[k * rk + c]; };
  for (int k = 0; k < nz; ++k) tCK[k * rk + c] = ck(k);
  T cp_prev = T(0);
  for (int m = 0; m < n; ++m) {
    const int k = m + 1;
    const T th0w = T(0.5) * (C(g.theta0)[(k - 1) * rk + c] + C(g.theta0)[k * rk + c]);
    const T gk = dtD * T(kCp) * th0w * C(g.rdzw)[k];
    const T ga = gk * an2;
    const T ckk = ck(k), ckm = ck(k - 1);
    const T a = (m > 0) ? -ga * (ckm * r0w(k - 1)) : T(0);
    const T b = T(1) + ga * ((ckk + ckm) * r0w(k));
    const T cc = (m < n - 1) ? -ga * (ckk * r0w(k + 1)) : T(0);
    const T den = (m == 0) ? b : b - a * cp_prev;
    const T inv = T(1) / den;
    const T cp = cc * inv;
    tA[m * rk + c] = a * inv;
    tINV[m * rk + c] = inv;
    tCP[m * rk + c] = cp;
    tGK[m * rk + c] = gk;
    cp_prev = cp;
  }
}

// ---- fused acoustic substep -------------------------------------------------
// One workgroup = 64 consecutive columns (one per lane) x NW waves; wave wv
// owns the levels [wv*ch, wv*ch+ch).  Phase A (all waves, their levels):
// horizontal momentum forward step with the old pi', slope flux Mc, rho0 mass
// divergence, explicit rho'/pi' parts E, Q and the old-w part of the w RHS.
// Phases B: vertical fluxes with the old w and the explicit pi' estimate,
// exchanging the chunk-boundary values through LDS.  Phase C (wave 0): the
// serial Thomas sweeps of the 64 columns over the RHS held in LDS, with the
// stage's precomputed factorisation.  Phase D (all waves): vertical flux
// divergence closes rho' and pi', and w is written.  State traffic per cell:
// read u, v, w, rho', pi' once, write them once.
// Substep operands of one domain (device pointers and scalars).
template <typename T, typename I, typename O>
struct SubArgs {
  const I *u, *v, *w, *rho, *pi;  // acoustic state in
  HeviTab<T> tab;                 // HEVI factorisation and coefficients of this stage
  T dtau, an, ao;
  O *uo, *vo, *wo, *rho_o, *pi_o;  // acoustic state out
  int check;                       // 1: flag non-finite outputs
  int *flag;
  SpongeArgs sp = {};              // FP64 outputs only: sponge of the step's final state
  // compensated FP32 mode: remainders of the state in and (FP32 outputs)
  // out
  const int8_t *ul = nullptr;
  const uint16_t *vl = nullptr, *wl = nullptr, *pl = nullptr;
  int8_t *ulo = nullptr;
  uint16_t *vlo = nullptr, *wlo = nullptr, *plo = nullptr;
  // FP32 state: chi = rho' - pi'/kappa of the stage, read by the stage's
  // last substep only (rho_o != nullptr)
  const I *chi = nullptr;
  // compensated mode: FP64 geometry for the mass-flux sums
  const double *rhu64 = nullptr, *rhv64 = nullptr, *r0w64 = nullptr, *sx64 = nullptr, *sy64 = nullptr,
               *zeta64 = nullptr;
  double rdx64 = 0.0, rdy64 = 0.0;
};

// Thomas sweep operands: the compensated kernels' sweeping wave reads beta
// and c' from the HEVI table in global memory (prefetched 32 rows ahead) and
// only the RHS is staged in LDS, (nz-1)*64 values per workgroup instead of
// three times that, so three 4-wave or two 8-wave workgroups fit per CU at
// their register counts; the other kernels stage all three in LDS in phase
// B (measured faster for them: their occupancy is bound by registers).
__host__ __device__ constexpr bool substep_thomas_global(bool cp) { return cp; }
// LDS arrays of (nz-1)*64 values each a substep workgroup stages
__host__ __device__ constexpr int substep_lds_arrays(bool cp) { return substep_thomas_global(cp) ? 1 : 3; }
// Compensated mode: the substep fields that carry a remainder (bits: 1 u,
// 2 v, 4 w, 8 pi'); a field without one is rounded to FP32 at every
// substep as in the plain mode.  Compile-time switch, u alone by default:
// at 256x256x96 over 1 h the worst field is 5.86e-5 of FP64 with u alone
// and 5.83e-5 with all four, against 1.01e-4 with v, w and pi' but not u.
#ifndef VMI_CP_MASK
#define VMI_CP_MASK 1
#endif
// Remainder arrays a compensated stage writes: u, plus the fields VMI_CP_MASK
// names (chi = rho' - pi'/kappa, read once by the stage's last substep, is
// stored in FP32 alone).
__host__ __device__ constexpr bool cp_keeps(int bit) { return (VMI_CP_MASK & bit) != 0; }
// Waves per SIMD the substep kernels are compiled for: FP32 shapes with up
// to 12 levels per wave (plain about 106 VGPRs, compensated about 127) are
// held to 128 VGPRs (four waves per SIMD), compensated shapes with up to 24
// levels per wave to 168 (three; the variant writing the FP64 state would
// take 171 and two); the other shapes are register-limited by their own
// size.
// With remainders on pi' or w (increment forms) the compensated kernels take
// about 175 VGPRs and are not held.
template <typename T, int NW, int MAXCH, bool CP = false>
inline constexpr int kSubstepWavesPerEu =
    (sizeof(T) == 4 && MAXCH <= 12 && (!CP || (VMI_CP_MASK & 12) == 0))               ? 4
    : (sizeof(T) == 4 && CP && MAXCH <= 24 && (VMI_CP_MASK & 12) == 0) ? 3
                                                                        : 1;

// Increment form of the plain FP32 substep (compile-time switch
// VMI_FP32_INCREMENT): rho', pi' and w updated as old value + increment.
#ifndef VMI_FP32_INCREMENT
#define VMI_FP32_INCREMENT 0
#endif
// Compensated mode: which mass-flux sums are formed in FP64 from FP64
// geometry (compile-time switch; bit 1: the horizontal divergence in FP32,
// bit 2: the slope flux and the vertical mass flux in FP32, bit 4: the y
// part of the horizontal divergence in FP32; 0: all FP64).  Default 6: the
// x part of the horizontal divergence in FP64, the rest in FP32.  Worst
// field at 256x256x96 after 1 h: 1.0e-4 with the whole horizontal
// divergence in FP32, 5.85e-5 with everything in FP64, 6.1e-5 with the
// slope and vertical fluxes in FP32, 6.9e-5 with the y part in FP32 as well
// (one FP64 load per level less, 1 % faster at 1024x1024x96).
#ifndef VMI_COMP_SUMS
#define VMI_COMP_SUMS 6
#endif
template <typename T, typename I, bool CP>
inline constexpr bool kSubstepIncrement = VMI_FP32_INCREMENT != 0 && sizeof(T) == 4 && sizeof(I) == 4 && !CP;

// Row of 64 columns starting at x-block bx, row j; all 64*NW threads of the
// workgroup call it together (it synchronises the workgroup).
//
// Memory access is software-pipelined: every global operand of a level (the
// state, pi' two levels ahead, the face and interface products, the HEVI
// table) is loaded one level before the level's arithmetic, and a level's
// loads are issued before the arithmetic of the level before it.  The
// vector-memory counter retires in issue order, so the wait for level k's
// operands never waits for level k+1's loads: the memory latency of one level
// overlaps the arithmetic of the previous one.  The level loops are fully
// unrolled and the operand buffers are indexed by (level mod 2) or (level
// mod 3), which the unrolling turns into fixed registers (no copies).
// Addresses are (array base, per-lane 32-bit column offset + uniform level
// offset); the 32-bit sum is formed next to each access, which keeps the
// scalar-base + vector-offset addressing form.
//
// CP (compensated FP32 mode, T = I = float): u carries an 8-bit remainder
// (r8_bits) and its FP32 increment is added to hi + lo exactly (comp_add);
// the x part of the horizontal mass-flux divergence is formed in FP64 from
// the full new winds and FP64 face products (F64H, F64Y for the y part
// too); the Thomas sweep reads beta and c'
// from global memory (substep_thomas_global).  VMI_CP_MASK can give v, w and
// pi' bf16 remainders as well; with one on pi' or w, pi' and w are updated in
// increment form (IN): w from the increment system A dw = -gk (ao dpi'_old +
// an dP^), where P^ is the pi' estimate with the full old-w vertical flux
// rho0_w (w - M).  With w_new = w + dw this is the same system as
// A w_new = w - gk (ao dpi'_old + an dP^ao) of the plain form (P^ao with the
// ao part of the old-w flux), and the closing flux rho0_w (w - M) + an rho0_w dw
// equals rho0_w (an w_new + ao w - M).
// RO (FP32 state): this substep writes rho' (a stage's last substep; always
// with FP64 outputs).
template <typename T, typename I, typename O, int NW, int MAXCH, int GM, bool CP = false, bool INC = false,
          bool RO = (sizeof(O) == 8)>
__device__ __forceinline__ void substep_row(const Geo<T> &g, const SubArgs<T, I, O> &sa, int bx, int j) {
  static_assert(!CP || (sizeof(T) == 4 && sizeof(I) == 4), "compensated storage is FP32");
  static_assert(!INC || (sizeof(T) == 4 && sizeof(I) == 4), "the increment form is an FP32 form");
  // RM: the compensated mode (bf16 remainders of the fields VMI_CP_MASK
  // names and FP64 mass-flux sums); IN: pi' and w updated as increments of
  // the old state (each rounded once at full magnitude), which the
  // compensated mode needs when pi' or w carries a remainder
  constexpr bool RM = CP;
  constexpr bool IN = INC || (CP && (VMI_CP_MASK & 12) != 0);
  // F64H / F64M: the compensated mode's horizontal mass-flux divergence /
  // slope and vertical mass fluxes in FP64 from FP64 geometry (compile-time
  // switch VMI_COMP_SUMS; otherwise the FP32 forms of the plain mode)
  constexpr bool F64H = CP && (VMI_COMP_SUMS & 1) == 0;
  // F64Y: the y part of that divergence in FP64 as well (compile-time
  // switch VMI_COMP_SUMS bit 4 keeps it in the plain FP32 form)
  constexpr bool F64Y = F64H && (VMI_COMP_SUMS & 4) == 0;
  constexpr bool F64M = CP && (VMI_COMP_SUMS & 2) == 0;
  // RX (FP32 state): rho' is not carried through the substeps.  Every
  // substep changes pi' by kappa times its change of rho' (Q - pi' =
  // kappa (E - rho') and the closing adds vd to rho', kappa vd to pi'), so
  // the stage writes chi = rho' - pi'/kappa and only the stage's last
  // substep (rho_o != nullptr) forms rho' = chi + pi'/kappa.
  constexpr bool RX = sizeof(I) == 4;
  // RM: which fields keep their remainder through the substep
  constexpr bool LU = RM && (VMI_CP_MASK & 1), LV = RM && (VMI_CP_MASK & 2), LW = RM && (VMI_CP_MASK & 4),
                 LP = RM && (VMI_CP_MASK & 8);
  const I *__restrict__ u = sa.u;
  const I *__restrict__ v = sa.v;
  const I *__restrict__ w = sa.w;
  const I *__restrict__ rho = sa.rho;
  const I *__restrict__ pi = sa.pi;
  const T dtau = sa.dtau, an = sa.an, ao = sa.ao;
  O *__restrict__ uo = sa.uo;
  O *__restrict__ vo = sa.vo;
  O *__restrict__ wo = sa.wo;
  O *__restrict__ rho_o = sa.rho_o;
  O *__restrict__ pi_o = sa.pi_o;
  const int check = sa.check;
  int *__restrict__ flag = sa.flag;
  extern __shared__ unsigned char smem_raw[];
  const int nlds = (g.nz - 1) * 64;
  T *s_td = reinterpret_cast<T *>(smem_raw);  // [nz-1][64] old-w RHS part -> RHS -> d' -> w
  constexpr bool TG = substep_thomas_global(CP);
  T *s_be = s_td + (TG ? 0 : nlds);           // [nz-1][64] beta (not with TG)
  T *s_cp = s_be + (TG ? 0 : nlds);           // [nz-1][64] c' (not with TG)
  // RM: the slope and vertical mass fluxes are formed in FP64
  using TF = std::conditional_t<F64M, double, T>;
  __shared__ TF s_mc[NW][64], s_fh[NW][64];
  __shared__ T s_ph[NW][64];
  const int lane = threadIdx.x & 63;
  // wave index: uniform within the wavefront, so level offsets are scalars
  const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nx = g.nx, ny = g.ny, nz = g.nz, rk = g.rk;
  const uint32_t nxy = static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny);
  const int i_raw = bx * 64 + lane;
  const bool active = i_raw < nx;
  const int i = active ? i_raw : nx - 1;
  const int ch = (nz + NW - 1) / NW;
  const int k0 = wv * ch;
  const int nlev = max(0, min(nz, k0 + ch) - k0);
  const int iW = wrapi(i - 1, nx), iE = wrapi(i + 1, nx), jS = wrapi(j - 1, ny), jN = wrapi(j + 1, ny);
  const uint32_t unx = static_cast<uint32_t>(nx);
  constexpr uint32_t SI = sizeof(I), SO = sizeof(O), ST = sizeof(T);
  // per-lane byte offsets of the five state columns within a level
  const uint32_t cC = static_cast<uint32_t>(j) * unx + i;
  const uint32_t vC = cC * SI, vW = (static_cast<uint32_t>(j) * unx + iW) * SI,
                 vE = (static_cast<uint32_t>(j) * unx + iE) * SI, vS = (static_cast<uint32_t>(jS) * unx + i) * SI,
                 vN = (static_cast<uint32_t>(jN) * unx + i) * SI, oC = cC * SO;
  const uint32_t gC = gcolm<GM>(g, i, j), gE = gcolm<GM>(g, iE, j), gN = gcolm<GM>(g, i, jN);
  const uint32_t tC = gC * ST, tE = gE * ST, tN = gN * ST;
  // level strides in bytes (uniform)
  const uint32_t lsI = nxy * SI, lsO = nxy * SO, lsT = static_cast<uint32_t>(rk) * ST;
  const uint32_t lC = cC * 2u, lsL = nxy * 2u;  // bf16 remainders (CP)
  const uint32_t lE = (static_cast<uint32_t>(j) * unx + iE) * 2u, lN = (static_cast<uint32_t>(jN) * unx + i) * 2u;
  const uint32_t l8C = cC, l8E = lE >> 1, lsL8 = nxy;  // 8-bit u remainders (CP)
  const uint32_t tC8 = gC * 8u, tE8 = gE * 8u, tN8 = gN * 8u, lsT8 = static_cast<uint32_t>(rk) * 8u;
  constexpr bool xslope = GM != kGeoFlat, yslope = GM == kGeoAny;  // nonzero slope / metric terms
  constexpr bool colE = GM != kGeoFlat, colN = GM == kGeoAny;     // east / north geometry column differs
  const T rD = C(g.rDc)[gC];
  const T dtD = dtau * rD;
  const T sx_w = C(g.sx)[gC], sx_e = C(g.sx)[gE], sy_s = C(g.sy)[gC], sy_n = C(g.sy)[gN];
  const T mx_w = C(g.metx)[gC], mx_e = C(g.metx)[gE], my_s = C(g.mety)[gC], my_n = C(g.mety)[gN];
  const T cpdt = dtau;

  // per level of this wave: explicit rho' and pi' parts and the old-w part
  // of the vertical mass flux (the old-w part of the w RHS goes to s_td)
  T E[MAXCH], Q[MAXCH];
  TF Fh[MAXCH];
  TF mc_first = TF(0), mc_last = TF(0), mc0 = TF(0), aw_first = TF(0), r0w_first = TF(0);

  // ---------------- phase A ----------------
  if (nlev > 0) {
    auto ld5 = [&](int k, T *p) {
      const uint32_t s = static_cast<uint32_t>(k) * lsI;
      p[0] = T(gld(pi, vC + s));
      p[1] = T(gld(pi, vW + s));
      p[2] = T(gld(pi, vE + s));
      p[3] = T(gld(pi, vS + s));
      p[4] = T(gld(pi, vN + s));
    };
    auto ld_state = [&](int k, T *x, int *qx) {
      const uint32_t s = static_cast<uint32_t>(k) * lsI;
      x[0] = T(gld(u, vC + s));
      x[1] = T(gld(u, vE + s));
      x[2] = T(gld(v, vC + s));
      x[3] = T(gld(v, vN + s));
      x[4] = T(gld(w, vC + s));
      if constexpr (!RX) x[5] = T(gld(rho, vC + s));
      if constexpr (RM) {
        const uint32_t sl = static_cast<uint32_t>(k) * lsL;
        // u: raw 8-bit remainders, decoded at use against their hi values
        const uint32_t sl8 = static_cast<uint32_t>(k) * lsL8;
        qx[0] = LU ? static_cast<int>(gld(sa.ul, l8C + sl8)) : 0;
        qx[1] = LU ? static_cast<int>(gld(sa.ul, l8E + sl8)) : 0;
        x[7] = LV ? bf16_val(gld(sa.vl, lC + sl)) : T(0);
        x[9] = LV ? bf16_val(gld(sa.vl, lN + sl)) : T(0);
        x[10] = LW ? bf16_val(gld(sa.wl, lC + sl)) : T(0);
      }
    };
    // face and interface products of a level: cp theta0 and D rho0 at the
    // west/east u faces and south/north v faces, kappa, gk and r0w of the
    // interface below the level
    enum { GUW, GUE, GVS, GVN, HUW, HUE, HVS, HVN, KAP, GKB, R0B, NG };
    // RM: the mass-flux products D rho0 at the faces and r0w come in FP64
    // (qd: HUW, HUE, HVS, HVN, R0B) in place of their FP32 values
    auto ld_geo = [&](int k, T *q, double *qd) {
      const uint32_t s = static_cast<uint32_t>(k) * lsT;
      q[GUW] = gldc(g.cthu, tC + s);
      q[GUE] = colE ? gldc(g.cthu, tE + s) : q[GUW];
      q[GVS] = gldc(g.cthv, tC + s);
      q[GVN] = colN ? gldc(g.cthv, tN + s) : q[GVS];
      if constexpr (sizeof(T) == 4) {
        const uint32_t s8 = static_cast<uint32_t>(k) * lsT8;
        if constexpr (F64H) {
          qd[0] = gldc(sa.rhu64, tC8 + s8);
          qd[1] = colE ? gldc(sa.rhu64, tE8 + s8) : qd[0];
          if constexpr (F64Y) {
            qd[2] = gldc(sa.rhv64, tC8 + s8);
            qd[3] = colN ? gldc(sa.rhv64, tN8 + s8) : qd[2];
          } else {
            q[HVN] = gldc(g.rhv, (colN ? tN : tC) + s);
            q[HVS] = colN ? gldc(g.drhv, tC + s) : T(0);
          }
        } else {
          // FP32: the east / north face products and their differences across
          // the cell (zero where the geometry is uniform in that direction)
          q[HUE] = gldc(g.rhu, (colE ? tE : tC) + s);
          q[HUW] = colE ? gldc(g.drhu, tC + s) : T(0);
          q[HVN] = gldc(g.rhv, (colN ? tN : tC) + s);
          q[HVS] = colN ? gldc(g.drhv, tC + s) : T(0);
        }
        if constexpr (F64M)
          qd[4] = gldc(sa.r0w64, tC8 + s8);
        else
          q[R0B] = gldc(g.r0w, tC + s);
      } else {
        q[HUW] = gldc(g.rhu, tC + s);
        q[HUE] = colE ? gldc(g.rhu, tE + s) : q[HUW];
        q[HVS] = gldc(g.rhv, tC + s);
        q[HVN] = colN ? gldc(g.rhv, tN + s) : q[HVS];
        q[R0B] = gldc(g.r0w, tC + s);
      }
      q[KAP] = gldc(g.kap, tC + s);
      // unconditional (level 0 reads entries it does not use): a load under a
      // branch would be waited for where the branches join
      q[GKB] = gldc(sa.tab.GK, tC + (k >= 1 ? s - lsT : 0u));
    };
    // pi' of the five columns at levels k (slot c%3) and k+1 (slot (c+1)%3);
    // the vertical gradient of the interface below k is carried in gl.
    T P[3][5], gl[5];
    ld5(k0, P[0]);
    T pmC = P[0][0];
    if (k0 >= 1) {
      T pm[5];
      ld5(k0 - 1, pm);
      pmC = pm[0];
#pragma unroll
      for (int q = 0; q < 5; ++q) gl[q] = (P[0][q] - pm[q]) * C(g.rdzw)[k0];
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) gl[q] = T(0);
    }
    ld5(min(k0 + 1, nz - 1), P[1]);
    T X[2][RM ? 11 : 6], Gq[2][NG];
    int QX[2][2] = {{0, 0}, {0, 0}};
    double Gd[2][RM ? 5 : 1];
    ld_state(k0, X[0], QX[0]);
    ld_geo(k0, Gq[0], Gd[0]);
    TF mc_prev = TF(0);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      const int cs = c & 1;
      // operands of level k+1 (and pi' of k+2), issued before level k's
      // arithmetic.  Unconditional (the level is clamped to the wave's last
      // one): a load under the level guard would be waited for where the
      // guard's branches join.
      {
        const int kn = k0 + min(c + 1, nlev - 1);
        ld_state(kn, X[cs ^ 1], QX[cs ^ 1]);
        ld_geo(kn, Gq[cs ^ 1], Gd[cs ^ 1]);
        ld5(min(kn + 1, nz - 1), P[(c + 2) % 3]);
      }
      if (c < nlev) {
        const int k = k0 + c;
        const int p0 = c % 3, p1 = (c + 1) % 3;
        T *xs = X[cs];
        if constexpr (LU) {  // the u remainders of the own and east columns
          xs[6] = r8_val(xs[0], QX[cs][0]);
          xs[8] = r8_val(xs[1], QX[cs][1]);
        }
        const T *gs = Gq[cs];
        const double *gd = Gd[cs];
        const T *pc = P[p0];
        const T *pn = P[p1];
        const uint32_t so = static_cast<uint32_t>(k) * lsO;
        const T rz_hi = C(g.rdzw)[k + 1];
        T Gc[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          if ((q <= 2 && xslope) || yslope) {  // vertical gradients only where a metric term uses them
            const T ghi = (k + 1 <= nz - 1) ? (pn[q] - pc[q]) * rz_hi : T(0);
            Gc[q] = T(0.5) * (gl[q] + ghi);
            gl[q] = ghi;
          } else {
            Gc[q] = T(0);
          }
        }
        const T omz = T(1) - C(g.zc)[k];
        T px_w = (pc[0] - pc[1]) * g.rdx, px_e = (pc[2] - pc[0]) * g.rdx;
        T py_s = (pc[0] - pc[3]) * g.rdy, py_n = (pc[4] - pc[0]) * g.rdy;
        if constexpr (xslope) {
          px_w = px_w - (omz * mx_w) * (T(0.5) * (Gc[1] + Gc[0]));
          px_e = px_e - (omz * mx_e) * (T(0.5) * (Gc[0] + Gc[2]));
        }
        if constexpr (yslope) {
          py_s = py_s - (omz * my_s) * (T(0.5) * (Gc[3] + Gc[0]));
          py_n = py_n - (omz * my_n) * (T(0.5) * (Gc[0] + Gc[4]));
        }
        // FP32 buffers hold u - ubg (see Geo::ubg); un_* are the full winds
        F2 cu{}, cv{};
        if constexpr (RM) {
          cu = comp_add(xs[0], xs[6], -(cpdt * (gs[GUW] * px_w)));
          cv = comp_add(xs[2], xs[7], -(cpdt * (gs[GVS] * py_s)));
        }
        const T du_w = RM ? cu.hi : xs[0] - cpdt * (gs[GUW] * px_w);
        const T du_e = xs[1] - cpdt * (gs[GUE] * px_e);
        T un_w = du_w, un_e = du_e;
        if constexpr (sizeof(T) == 4) {
          const T ubk = C(g.ubg)[k];
          un_w = ubk + du_w;
          un_e = ubk + du_e;
        }
        const T vn_s = RM ? cv.hi : xs[2] - cpdt * (gs[GVS] * py_s);
        const T vn_n = xs[3] - cpdt * (gs[GVN] * py_n);
        if (active) {
          if constexpr (sizeof(O) == 8) {
            double uf = static_cast<double>(du_w), vf = static_cast<double>(vn_s);
            if constexpr (RM) {
              uf += static_cast<double>(cu.lo);
              vf += static_cast<double>(cv.lo);
            }
            if constexpr (sizeof(T) == 4) uf += C(g.ubg64)[k];
            gst(uo, oC + so, sponge_scale(sa.sp, k, nz, uf));
            gst(vo, oC + so, sponge_scale(sa.sp, k, nz, vf));
          } else {
            gst(uo, oC + so, static_cast<O>(du_w));
            gst(vo, oC + so, static_cast<O>(vn_s));
            if constexpr (RM) {
              const uint32_t sl = static_cast<uint32_t>(k) * lsL;
              if constexpr (LU) gst(sa.ulo, l8C + static_cast<uint32_t>(k) * lsL8, r8_bits(cu.hi, cu.lo));
              if constexpr (LV) gst(sa.vlo, lC + sl, bf16_bits(cv.lo));
            }
          }
        }
        TF mc = TF(0);
        T hd;
        // F64H, F64M: full new winds hi + lo (+ background) in FP64; the face
        // mass fluxes D rho0 u (about 1e5) are differenced across the cell in
        // FP64 from FP64 geometry (F64H), the slope flux formed in FP64 (F64M)
        double uwd = 0.0, ued = 0.0, vsd = 0.0, vnd = 0.0;
        if constexpr (F64H || F64M) {
          const double ub = C(g.ubg64)[k];
          uwd = ub + static_cast<double>(cu.hi) + static_cast<double>(cu.lo);
          ued = ub + static_cast<double>(xs[1]) + static_cast<double>(xs[8]) +
                static_cast<double>(-(cpdt * (gs[GUE] * px_e)));
          vsd = static_cast<double>(cv.hi) + static_cast<double>(cv.lo);
          vnd = static_cast<double>(xs[3]) + static_cast<double>(xs[9]) +
                static_cast<double>(-(cpdt * (gs[GVN] * py_n)));
        }
        if constexpr (F64M) {
          if constexpr (yslope)
            mc = 0.5 * (uwd * C(sa.sx64)[gC] + ued * C(sa.sx64)[gE]) + 0.5 * (vsd * C(sa.sy64)[gC] + vnd * C(sa.sy64)[gN]);
          else if constexpr (xslope)
            mc = 0.5 * (uwd * C(sa.sx64)[gC] + ued * C(sa.sx64)[gE]);
        } else {
          if constexpr (yslope)
            mc = T(0.5) * (un_w * sx_w + un_e * sx_e) + T(0.5) * (vn_s * sy_s + vn_n * sy_n);
          else if constexpr (xslope)
            mc = T(0.5) * (un_w * sx_w + un_e * sx_e);
        }
        if constexpr (F64H) {
          if constexpr (F64Y) {
            const double hdd = (gd[1] * ued - gd[0] * uwd) * sa.rdx64 + (gd[3] * vnd - gd[2] * vsd) * sa.rdy64;
            hd = static_cast<T>(hdd);
          } else {  // the y part in the plain FP32 difference form
            const T hdy = (gs[HVN] * (vn_n - vn_s) + gs[HVS] * vn_s) * g.rdy;
            hd = static_cast<T>((gd[1] * ued - gd[0] * uwd) * sa.rdx64 + static_cast<double>(hdy));
          }
        } else {
          if constexpr (sizeof(T) == 4) {
            // D rho0 u differenced across the cell without forming the large
            // products: rhu_e (u_e - u_w) + (rhu_e - rhu_w) u_w, where u_e - u_w
            // comes from the stored deviations (the background cancels) and,
            // with u remainders, from their difference as well
            T dux = du_e - du_w, dvy = vn_n - vn_s;
            if constexpr (RM) {
              // RM: the new winds of both faces exactly as hi + lo (the west
              // and south ones are cu, cv), so the differences carry no
              // rounding of the face values
              const F2 ce = comp_add(xs[1], xs[8], -(cpdt * (gs[GUE] * px_e)));
              const F2 cn = comp_add(xs[3], xs[9], -(cpdt * (gs[GVN] * py_n)));
              dux = (ce.hi - cu.hi) + (ce.lo - cu.lo);
              dvy = (cn.hi - cv.hi) + (cn.lo - cv.lo);
            }
            hd = (gs[HUE] * dux + gs[HUW] * un_w) * g.rdx + (gs[HVN] * dvy + gs[HVS] * vn_s) * g.rdy;
          } else {
            hd = (gs[HUE] * un_e - gs[HUW] * un_w) * g.rdx + (gs[HVN] * vn_n - gs[HVS] * vn_s) * g.rdy;
          }
        }
        const T cq = gs[KAP] * dtD;
        // IN: E and Q hold the increments of rho' and pi'
        if constexpr (!RX) E[c] = IN ? -(dtD * hd) : xs[5] - dtD * hd;
        Q[c] = IN ? -(cq * hd) : pc[0] - cq * hd;
        const T wk = xs[4];
        Fh[c] = TF(0);
        if (k >= 1) {
          s_td[(k - 1) * 64 + lane] = IN ? -(gs[GKB] * ao * (pc[0] - pmC)) : wk - gs[GKB] * ao * (pc[0] - pmC);
          TF awk, r0k;
          // IN: the full old-w flux; otherwise its ao part
          if constexpr (F64M) {
            const double wd = static_cast<double>(wk) + static_cast<double>(xs[10]);
            awk = IN ? wd : static_cast<double>(ao) * wd;
            r0k = gd[4];
          } else {
            awk = IN ? wk : ao * wk;
            r0k = gs[R0B];
          }
          if (c > 0) {
            const TF zk = F64M ? TF(C(sa.zeta64)[k]) : TF(C(g.zeta)[k]);
            const TF Mw = (TF(1) - zk) * (TF(0.5) * (mc_prev + mc));
            Fh[c] = r0k * (awk - Mw);
          } else {
            aw_first = awk;
            r0w_first = r0k;
          }
        }
        if (c == 0) mc_first = mc;
        if (k == 0) mc0 = mc;
        mc_prev = mc;
        pmC = pc[0];
      }
    }
    mc_last = mc_prev;
    s_mc[wv][lane] = mc_last;
  }
  __syncthreads();
  // ---------------- phase B: boundary vertical flux, P^, RHS ----------------
  // P^ = Q - ck (Fh_{k+1} - Fh_k) is the pi' estimate with the old-w fluxes;
  // the RHS of interface k is the old-w part minus gk an (P^_k - P^_{k-1}),
  // scaled by 1/pivot for the factored sweep.
  if (nlev > 0) {
    if (k0 >= 1) {
      const TF zk = F64M ? TF(C(sa.zeta64)[k0]) : TF(C(g.zeta)[k0]);
      const TF Mw = (TF(1) - zk) * (TF(0.5) * (s_mc[wv - 1][lane] + mc_first));
      Fh[0] = r0w_first * (aw_first - Mw);
    }
    s_fh[wv][lane] = Fh[0];
  }
  __syncthreads();
  if (nlev > 0) {
    const TF fh_next_chunk = (k0 + nlev < nz) ? s_fh[wv + 1][lane] : TF(0);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      if (c == nlev - 1) {
        const int k = k0 + c;
        const T ck = gldc(sa.tab.CK, tC + static_cast<uint32_t>(k) * lsT);
        T qf = Q[c];
        if constexpr (IN) qf = T(gld(pi, vC + static_cast<uint32_t>(k) * lsI)) + qf;
        s_ph[wv][lane] = qf - ck * static_cast<T>(fh_next_chunk - Fh[c]);
      }
    }
  }
  __syncthreads();
  if (nlev > 0) {
    // ck of a level and gk, 1/pivot, beta, c' of the interface below it
    auto ld_tab = [&](int k, T *t) {
      const uint32_t s = static_cast<uint32_t>(k) * lsT;
      t[0] = gldc(sa.tab.CK, tC + s);
      const uint32_t sm = tC + (k >= 1 ? s - lsT : 0u);  // level 0: read, unused
      t[1] = gldc(sa.tab.GK, sm);
      t[2] = gldc(sa.tab.INV, sm);
      if constexpr (!TG) {
        t[3] = gldc(sa.tab.A, sm);
        t[4] = gldc(sa.tab.CP, sm);
      }
      if constexpr (IN) t[5] = T(gld(pi, vC + static_cast<uint32_t>(k) * lsI));
    };
    const TF fh_next_chunk = (k0 + nlev < nz) ? s_fh[wv + 1][lane] : TF(0);
    T php = (k0 >= 1) ? s_ph[wv - 1][lane] : T(0);
    // operands two levels ahead (the level's own work is short), unconditional
    T TB[3][IN ? 6 : 5];
    ld_tab(k0, TB[0]);
    ld_tab(k0 + min(1, nlev - 1), TB[1]);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      ld_tab(k0 + min(c + 2, nlev - 1), TB[(c + 2) % 3]);
      if (c < nlev) {
        const int k = k0 + c;
        const T *tb = TB[c % 3];
        const TF fhn = (c + 1 < nlev) ? Fh[(c + 1) < MAXCH ? c + 1 : c] : fh_next_chunk;
        const T ph = (IN ? tb[5] + Q[c] : Q[c]) - tb[0] * static_cast<T>(fhn - Fh[c]);
        if (k >= 1) {
          const int m = k - 1;
          const T td = s_td[m * 64 + lane] - tb[1] * an * (ph - php);
          s_td[m * 64 + lane] = td * tb[2];
          if constexpr (!TG) {
            s_be[m * 64 + lane] = tb[3];
            s_cp[m * 64 + lane] = tb[4];
          }
        }
        php = ph;
      }
    }
  }
  __syncthreads();
  // ---------------- phase C: Thomas sweeps (wave 0) ----------------
  // RHS, beta and c' are in LDS (staged by all waves in phase B).  Rows
  // (interfaces 1..nz-1) after the first run as a short remainder followed by
  // whole blocks of B rows; each block's operands are fetched while the
  // previous block runs, into two register sets that alternate, so only the
  // dependent multiply-subtract per row stays on the critical path and the
  // rows of a block carry no guards.
  if (TG && wv == 0) {
    // beta and c' from the table in global memory: D blocks of B rows in
    // flight ahead of the block being swept (ring slots fixed by unrolling)
    constexpr int B = 8, D = 4;
    const int n = nz - 1;
    const int r = (n - 1) % B, nbf = (n - 1) / B;
    T *__restrict__ td = s_td + lane;
    auto tabv = [&](const T *f, int m) { return gldc(f, tC + static_cast<uint32_t>(m) * lsT); };
    T ring[D][B];
    // forward elimination: d'_0 = RHS_0, d'_m = RHS_m - beta_m d'_{m-1}
    T dp = td[0];
    auto fetch_f = [&](int blk, T *o) {
      const int m0 = r + 1 + min(blk, nbf - 1) * B;
#pragma unroll
      for (int q = 0; q < B; ++q) o[q] = tabv(sa.tab.A, m0 + q);
    };
    // the remainder rows 1..r first, then the first D-1 blocks (in that
    // order: the loads retire in issue order)
    T rem[B];
#pragma unroll
    for (int q = 0; q < B; ++q) rem[q] = tabv(sa.tab.A, min(1 + q, n - 1));
    if (nbf > 0) {
#pragma unroll
      for (int sl = 0; sl < D - 1; ++sl) fetch_f(sl, ring[sl]);
    }
#pragma unroll
    for (int q = 0; q < B; ++q) {
      if (q < r) {
        dp = td[(1 + q) * 64] - rem[q] * dp;
        td[(1 + q) * 64] = dp;
      }
    }
    for (int blk = 0; blk < nbf; blk += D) {
#pragma unroll
      for (int sl = 0; sl < D; ++sl) {
        if (blk + sl < nbf) {
          fetch_f(blk + sl + D - 1, ring[(sl + D - 1) % D]);
          const int m0 = r + 1 + (blk + sl) * B;
          T v[B];
#pragma unroll
          for (int q = 0; q < B; ++q) v[q] = td[(m0 + q) * 64];
#pragma unroll
          for (int q = 0; q < B; ++q) {
            dp = v[q] - ring[sl][q] * dp;
            td[(m0 + q) * 64] = dp;
          }
        }
      }
    }
    // back substitution: x_{n-1} = d'_{n-1}, x_m = d'_m - c'_m x_{m+1}
    auto fetch_b = [&](int blk, T *o) {
      const int m0 = n - 2 - r - min(blk, nbf - 1) * B;
#pragma unroll
      for (int q = 0; q < B; ++q) o[q] = tabv(sa.tab.CP, m0 - q);
    };
#pragma unroll
    for (int q = 0; q < B; ++q) rem[q] = tabv(sa.tab.CP, max(n - 2 - q, 0));
    if (nbf > 0) {
#pragma unroll
      for (int sl = 0; sl < D - 1; ++sl) fetch_b(sl, ring[sl]);
    }
    T x = dp;
#pragma unroll
    for (int q = 0; q < B; ++q) {
      if (q < r) {
        x = td[(n - 2 - q) * 64] - rem[q] * x;
        td[(n - 2 - q) * 64] = x;
      }
    }
    for (int blk = 0; blk < nbf; blk += D) {
#pragma unroll
      for (int sl = 0; sl < D; ++sl) {
        if (blk + sl < nbf) {
          fetch_b(blk + sl + D - 1, ring[(sl + D - 1) % D]);
          const int m0 = n - 2 - r - (blk + sl) * B;
          T v[B];
#pragma unroll
          for (int q = 0; q < B; ++q) v[q] = td[(m0 - q) * 64];
#pragma unroll
          for (int q = 0; q < B; ++q) {
            x = v[q] - ring[sl][q] * x;
            td[(m0 - q) * 64] = x;
          }
        }
      }
    }
  }
  if (!TG && wv == 0) {
    constexpr int B = 8;
    const int n = nz - 1;
    const int r = (n - 1) % B, nbf = (n - 1) / B;
    T *__restrict__ td = s_td + lane;
    const T *__restrict__ be = s_be + lane;
    const T *__restrict__ cpv = s_cp + lane;
    T o0[B], v0[B], o1[B], v1[B];
    // forward elimination: d'_0 = RHS_0, d'_m = RHS_m - beta_m d'_{m-1}
    T dp = td[0];
    for (int m = 1; m <= r; ++m) {
      dp = td[m * 64] - be[m * 64] * dp;
      td[m * 64] = dp;
    }
    auto fetch_f = [&](int blk, T *o, T *v) {
      const int m0 = r + 1 + blk * B;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        o[q] = be[(m0 + q) * 64];
        v[q] = td[(m0 + q) * 64];
      }
    };
    auto run_f = [&](int blk, const T *o, const T *v) {
      const int m0 = r + 1 + blk * B;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        dp = v[q] - o[q] * dp;
        td[(m0 + q) * 64] = dp;
      }
    };
    if (nbf > 0) {
      fetch_f(0, o0, v0);
      for (int blk = 0;; blk += 2) {
        fetch_f(min(blk + 1, nbf - 1), o1, v1);
        run_f(blk, o0, v0);
        if (blk + 1 >= nbf) break;
        fetch_f(min(blk + 2, nbf - 1), o0, v0);
        run_f(blk + 1, o1, v1);
        if (blk + 2 >= nbf) break;
      }
    }
    // back substitution: x_{n-1} = d'_{n-1}, x_m = d'_m - c'_m x_{m+1}
    T x = dp;
    for (int m = n - 2; m > n - 2 - r; --m) {
      x = td[m * 64] - cpv[m * 64] * x;
      td[m * 64] = x;
    }
    auto fetch_b = [&](int blk, T *o, T *v) {
      const int m0 = n - 2 - r - blk * B;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        o[q] = cpv[(m0 - q) * 64];
        v[q] = td[(m0 - q) * 64];
      }
    };
    auto run_b = [&](int blk, const T *o, const T *v) {
      const int m0 = n - 2 - r - blk * B;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        x = v[q] - o[q] * x;
        td[(m0 - q) * 64] = x;
      }
    };
    if (nbf > 0) {
      fetch_b(0, o0, v0);
      for (int blk = 0;; blk += 2) {
        fetch_b(min(blk + 1, nbf - 1), o1, v1);
        run_b(blk, o0, v0);
        if (blk + 1 >= nbf) break;
        fetch_b(min(blk + 2, nbf - 1), o0, v0);
        run_b(blk + 1, o1, v1);
        if (blk + 2 >= nbf) break;
      }
    }
  }
  __syncthreads();
  // ---------------- phase D: close rho', pi', write w ----------------
  if (nlev > 0) {
    int bad = 0;
    // r0w of the interface above a level and kappa of the level
    auto ld_d = [&](int k, T *t, double *t8) {
      const uint32_t s = static_cast<uint32_t>(k) * lsT;
      t[0] = gldc(g.r0w, tC + (s + lsT));  // r0w has nz+1 levels
      t[1] = gldc(g.kap, tC + s);
      if constexpr (IN) {  // the old state of the level: rho', pi', w (and RM: their remainders, FP64 r0w above)
        const uint32_t si = static_cast<uint32_t>(k) * lsI;
        if constexpr (!RX) t[2] = T(gld(rho, vC + si));
        t[4] = T(gld(pi, vC + si));
        t[6] = T(gld(w, vC + si));
        if constexpr (RM) {
          const uint32_t sl = static_cast<uint32_t>(k) * lsL;
          t[5] = LP ? bf16_val(gld(sa.pl, lC + sl)) : T(0);
          t[7] = LW ? bf16_val(gld(sa.wl, lC + sl)) : T(0);
        } else {
          t[3] = t[5] = t[7] = T(0);
        }
      }
      if constexpr (F64M) t8[0] = gldc(sa.r0w64, tC8 + static_cast<uint32_t>(k + 1) * lsT8);
      if constexpr (RX && RO) {  // chi of the level
        t[2] = T(gld(sa.chi, vC + static_cast<uint32_t>(k) * lsI));
        t[3] = T(0);  // chi carries no remainder
      }
    };
    const TF fh_next_chunk = (k0 + nlev < nz) ? s_fh[wv + 1][lane] : TF(0);
    T DD[3][(IN || (RX && RO)) ? 8 : 2];
    double DD8[3][1];
    ld_d(k0, DD[0], DD8[0]);
    ld_d(k0 + min(1, nlev - 1), DD[1], DD8[1]);
    // flux through the chunk's lower interface
    TF flo = TF(0);
    if (k0 >= 1) {
      if constexpr (F64M)
        flo = Fh[0] + (static_cast<double>(an) * gldc(sa.r0w64, tC8 + static_cast<uint32_t>(k0) * lsT8)) *
                          static_cast<double>(s_td[(k0 - 1) * 64 + lane]);
      else
        flo = Fh[0] + (an * gldc(g.r0w, tC + static_cast<uint32_t>(k0) * lsT)) * s_td[(k0 - 1) * 64 + lane];
    }
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      ld_d(k0 + min(c + 2, nlev - 1), DD[(c + 2) % 3], DD8[(c + 2) % 3]);  // two levels ahead, unconditional
      if (c < nlev) {
        const int k = k0 + c;
        const T *dd = DD[c % 3];
        const uint32_t so = static_cast<uint32_t>(k) * lsO;
        const TF fhn = (c + 1 < nlev) ? Fh[(c + 1) < MAXCH ? c + 1 : c] : fh_next_chunk;
        TF fhi = TF(0);
        if (k + 1 < nz) {
          if constexpr (F64M)
            fhi = fhn + (static_cast<double>(an) * DD8[c % 3][0]) * static_cast<double>(s_td[k * 64 + lane]);
          else
            fhi = fhn + (an * dd[0]) * s_td[k * 64 + lane];
        }
        const T vd = -dtD * (static_cast<T>(fhi - flo) * C(g.rdzeta)[k]);
        F2 cr{}, cpn{}, cw{};
        // IN: increments of rho', pi' and w, each added to the old value once
        const T dr = RX ? T(0) : E[c] + vd, dp = Q[c] + dd[1] * vd;
        if constexpr (RM && IN) {
          if constexpr (!RX) cr = comp_add(dd[2], dd[3], E[c] + vd);
          cpn = comp_add(dd[4], dd[5], Q[c] + dd[1] * vd);
          if (k >= 1) {
            cw = comp_add(dd[6], dd[7], s_td[(k - 1) * 64 + lane]);
          } else {
            cw.hi = static_cast<float>(mc0);
            cw.lo = static_cast<float>(mc0 - static_cast<double>(cw.hi));
          }
        }
        T rn = T(0), pn, wn;
        if constexpr (RM && IN) {
          rn = cr.hi, pn = cpn.hi, wn = cw.hi;
        } else if constexpr (IN) {
          rn = dd[2] + dr;
          pn = dd[4] + dp;
          wn = (k >= 1) ? dd[6] + s_td[(k - 1) * 64 + lane] : static_cast<T>(mc0);
        } else {
          if constexpr (!RX) rn = E[c] + vd;
          pn = Q[c] + dd[1] * vd;
          wn = (k >= 1) ? s_td[(k - 1) * 64 + lane] : static_cast<T>(mc0);
        }
        // RX, RO: rho' = chi + pi'/kappa in FP64 from the full new pi'
        double rx = 0.0;
        if constexpr (RX && RO) {
          double pv = static_cast<double>(pn);
          if constexpr (RM && IN)
            pv += static_cast<double>(cpn.lo);
          else if constexpr (IN)
            pv = static_cast<double>(dd[4]) + static_cast<double>(dp);
          rx = (static_cast<double>(dd[2]) + static_cast<double>(dd[3])) + pv * chi_rkap(dd[1]);
        }
        if (active) {
          if constexpr (sizeof(O) == 8) {
            double rf = static_cast<double>(rn), pf = static_cast<double>(pn), wf = static_cast<double>(wn);
            if constexpr (RM && IN) {
              rf += static_cast<double>(cr.lo);
              pf += static_cast<double>(cpn.lo);
              wf += static_cast<double>(cw.lo);
            } else if constexpr (IN) {  // the FP64 output takes the increment unrounded
              rf = static_cast<double>(dd[2]) + static_cast<double>(dr);
              pf = static_cast<double>(dd[4]) + static_cast<double>(dp);
              if (k >= 1) wf = static_cast<double>(dd[6]) + static_cast<double>(s_td[(k - 1) * 64 + lane]);
            }
            if constexpr (RX) rf = rx;
            gst(rho_o, oC + so, sponge_scale(sa.sp, k, nz, rf));
            gst(pi_o, oC + so, sponge_scale(sa.sp, k, nz, pf));
            gst(wo, oC + so, sponge_scale(sa.sp, k, nz, wf));
          } else {
            const uint32_t sl = static_cast<uint32_t>(k) * lsL;
            if constexpr (RX && RO) {
              const float rh = static_cast<float>(rx);
              gst(rho_o, oC + so, static_cast<O>(rh));
            } else if constexpr (!RX) {
              gst(rho_o, oC + so, static_cast<O>(rn));
            }
            gst(pi_o, oC + so, static_cast<O>(pn));
            gst(wo, oC + so, static_cast<O>(wn));
            if constexpr (RM) {
              if constexpr (LP) gst(sa.plo, lC + sl, bf16_bits(cpn.lo));
              if constexpr (LW) gst(sa.wlo, lC + sl, bf16_bits(cw.lo));
            }
          }
        }
        if constexpr (RX && RO)
          bad |= !(isfinite(rx) && isfinite(pn) && isfinite(wn));
        else if constexpr (RX)
          bad |= !(isfinite(pn) && isfinite(wn));
        else
          bad |= !(isfinite(rn) && isfinite(pn) && isfinite(wn));
        flo = fhi;
      }
    }
    if (k0 + nlev == nz && active) {
      const uint32_t s = static_cast<uint32_t>(nz);
      double wt = static_cast<double>(T(gld(w, vC + s * lsI)));
      uint16_t wtl = 0;
      if constexpr (RM) {
        if constexpr (LW) wtl = gld(sa.wl, lC + s * lsL);
        wt += static_cast<double>(bf16_val(wtl));
      }
      if constexpr (sizeof(O) == 8) {
        gst(wo, oC + s * lsO, sponge_scale(sa.sp, nz, nz, wt));
      } else {
        gst(wo, oC + s * lsO, static_cast<O>(T(gld(w, vC + s * lsI))));
        if constexpr (LW) gst(sa.wlo, lC + s * lsL, wtl);
      }
    }
    if (check && bad && active) atomicOr(flag, 1);
  }
}

template <typename T, typename I, typename O, int NW, int MAXCH, bool CP = false>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(kSubstepWavesPerEu<T, NW, MAXCH, CP>))) k_substep(Geo<T> g, SubArgs<T, I, O> sa) {
  substep_row<T, I, O, NW, MAXCH, kGeoAny, CP, kSubstepIncrement<T, I, CP>>(g, sa, blockIdx.x, blockIdx.y);
}

// ---- sponge (damping.cpp:76-138 with perturbation targets 0) -------------

__device__ __forceinline__ void sponge_at(int nx, int ny, int nz, const double *__restrict__ alpha, double dt,
                                          const FieldSet &s, int i, int j, int k) {
  VMI_GLOBAL double *__restrict__ u = G(s.u);
  VMI_GLOBAL double *__restrict__ v = G(s.v);
  VMI_GLOBAL double *__restrict__ w = G(s.w);
  VMI_GLOBAL double *__restrict__ rho = G(s.rho);
  VMI_GLOBAL double *__restrict__ th = G(s.th);
  VMI_GLOBAL double *__restrict__ pi = G(s.pi);
  const double a = G(alpha)[k < nz - 1 ? k : nz - 1];
  if (a <= 0.0) return;
  const double f = 1.0 / (1.0 + dt * a);
  const size_t c = (static_cast<size_t>(k) * ny + j) * nx + i;
  w[c] = f * w[c];
  if (k < nz) {
    u[c] = f * u[c];
    v[c] = f * v[c];
    rho[c] = f * rho[c];
    th[c] = f * th[c];
    pi[c] = f * pi[c];
  }
}

static __global__ void __launch_bounds__(256) k_sponge(int nx, int ny, int nz, int kmin, const double *__restrict__ alpha,
                                                double dt, FieldSet s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = kmin + static_cast<int>(blockIdx.z);
  if (i >= nx || k > nz) return;
  sponge_at(nx, ny, nz, alpha, dt, s, i, j, k);
}

// ---- diagnostics ------------------------------------------------------------

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;
constexpr int kEnergyComps = 5;

// Per-cell terms of energy_budget (diagnostics.cpp:255-277), compensated.
static __global__ void __launch_bounds__(kRedThreads) k_energy_partial(
    int nx, int ny, int nz, double dx, double dy, double H, const double *__restrict__ zs,
    const double *__restrict__ dzeta, const double *__restrict__ zcc, const double *__restrict__ rho0,
    const double *__restrict__ th0, const double *__restrict__ pi0, const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, const double *__restrict__ rho,
    const double *__restrict__ th, const double *__restrict__ pi, size_t chunk, NSum *__restrict__ part) {
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * nz;
  const size_t beg = static_cast<size_t>(blockIdx.x) * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  NSum acc[kEnergyComps];
  for (int q = 0; q < kEnergyComps; ++q) acc[q] = NSum{0.0, 0.0};
  for (size_t c = beg + threadIdx.x; c < end; c += kRedThreads) {
    const size_t k = c / nxy, col = c - k * nxy;
    const int j = static_cast<int>(col / nx), i = static_cast<int>(col - static_cast<size_t>(j) * nx);
    const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
    const double z0 = zs[col];
    const double dv = dx * dy * ((H - z0) / H) * dzeta[k] * H;
    const double r = rho0[c] + rho[c];
    const double uc = 0.5 * (u[c] + u[k * nxy + static_cast<size_t>(j) * nx + ip]);
    const double vc = 0.5 * (v[c] + v[k * nxy + static_cast<size_t>(jp) * nx + i]);
    const double wc = 0.5 * (w[c] + w[c + nxy]);
    const double T = (th0[c] + th[c]) * (pi0[c] + pi[c]);
    const double z = z0 + zcc[k] * (H - z0);
    acc[0].add(r * dv);
    acc[1].add(rho[c] * dv);
    acc[2].add(0.5 * r * (uc * uc + vc * vc + wc * wc) * dv);
    acc[3].add(r * kCv * T * dv);
    acc[4].add(r * kG * z * dv);
  }
  __shared__ NSum sh[kRedThreads];
  for (int q = 0; q < kEnergyComps; ++q) {
    sh[threadIdx.x] = acc[q];
    __syncthreads();
    for (int hh = kRedThreads / 2; hh > 0; hh >>= 1) {
      if (threadIdx.x < static_cast<unsigned>(hh)) sh[threadIdx.x].merge(sh[threadIdx.x + hh]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[static_cast<size_t>(q) * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

static __global__ void __launch_bounds__(kRedThreads) k_red_final(const NSum *__restrict__ part, int nparts, int ncomp,
                                                           double *__restrict__ out) {
  __shared__ NSum sh[kRedThreads];
  for (int comp = 0; comp < ncomp; ++comp) {
    NSum acc{0.0, 0.0};
    for (int p = threadIdx.x; p < nparts; p += kRedThreads) acc.merge(part[comp * nparts + p]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int hh = kRedThreads / 2; hh > 0; hh >>= 1) {
      if (threadIdx.x < static_cast<unsigned>(hh)) sh[threadIdx.x].merge(sh[threadIdx.x + hh]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[comp] = sh[0].s + sh[0].c;
    __syncthreads();
  }
}

// Stability diagnostics of the state (TimeStepController::inspect,
// timestep_control.cpp:23-56; its advective CFL is Advection::cfl_number,
// advection.cpp:325-340), one pass: res = {advective CFL, max wind, acoustic
// CFL, vertical CFL} as maxima, {min dz} as a minimum (slot 4, start +inf)
// and a non-finite flag (slot 5), each an FP64 bit pattern combined with
// integer atomics (non-negative doubles order like their bits).
static __global__ void __launch_bounds__(256) k_cfl_state(int nx, int ny, int nz, double dx, double dy, double H,
                                                   const double *__restrict__ zs, const double *__restrict__ dzeta,
                                                   const double *__restrict__ u, const double *__restrict__ v,
                                                   const double *__restrict__ w, const double *__restrict__ rho,
                                                   const double *__restrict__ th, const double *__restrict__ pi,
                                                   const double *__restrict__ th0,
                                                   const double *__restrict__ pi0, double dt,
                                                   unsigned long long *__restrict__ res) {
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * nz;
  double worst = 0.0, wind = 0.0, cac = 0.0, cvert = 0.0, mdz = __builtin_huge_val(), csmax = 0.0, geo = 0.0;
  int bad = 0;
  for (size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
       c += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t k = c / nxy, col = c - k * nxy;
    const int j = static_cast<int>(col / nx), i = static_cast<int>(col - static_cast<size_t>(j) * nx);
    const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1;
    const double dz = fmax(dzeta[k] * (H - zs[col]), 1.0);
    const double uu = 0.5 * (u[c] + u[k * nxy + static_cast<size_t>(j) * nx + ip]);
    const double vv = 0.5 * (v[c] + v[k * nxy + static_cast<size_t>(jp) * nx + i]);
    const double ww = 0.5 * (w[c] + w[c + nxy]);
    worst = fmax(worst, dt * (fabs(uu) / dx + fabs(vv) / dy + fabs(ww) / dz));
    wind = fmax(wind, sqrt(uu * uu + vv * vv + ww * ww));
    mdz = fmin(mdz, dz);
    const double t = (th0[c] + th[c]) * (pi0[c] + pi[c]);
    const double cs = sqrt((kCp / kCv) * kRd * fmax(t, 1.0));
    cac = fmax(cac, dt * cs * (1.0 / dx + 1.0 / dy + 1.0 / dz));
    cvert = fmax(cvert, dt * (fabs(ww) + cs) / dz);
    csmax = fmax(csmax, cs);                       // Equations::max_sound_speed
    geo = fmax(geo, 1.0 / dx + 1.0 / dy + 1.0 / dz);  // max_stable_dt's acoustic rate per unit cs
    bad |= !(isfinite(u[c]) && isfinite(v[c]) && isfinite(w[c]) && isfinite(rho[c]) && isfinite(th[c]) &&
             isfinite(pi[c]));
    if (k == static_cast<size_t>(nz) - 1) bad |= !isfinite(w[c + nxy]);
  }
  __shared__ double s[7][256];
  __shared__ int sb[256];
  s[0][threadIdx.x] = worst;
  s[1][threadIdx.x] = wind;
  s[2][threadIdx.x] = cac;
  s[3][threadIdx.x] = cvert;
  s[4][threadIdx.x] = mdz;
  s[5][threadIdx.x] = csmax;
  s[6][threadIdx.x] = geo;
  sb[threadIdx.x] = bad;
  __syncthreads();
  for (int hh = 128; hh > 0; hh >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(hh)) {
      for (int q = 0; q < 7; ++q)
        if (q != 4) s[q][threadIdx.x] = fmax(s[q][threadIdx.x], s[q][threadIdx.x + hh]);
      s[4][threadIdx.x] = fmin(s[4][threadIdx.x], s[4][threadIdx.x + hh]);
      sb[threadIdx.x] |= sb[threadIdx.x + hh];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // non-negative doubles order like their bit patterns
    for (int q = 0; q < 4; ++q) atomicMax(&res[q], static_cast<unsigned long long>(__double_as_longlong(s[q][0])));
    atomicMin(&res[4], static_cast<unsigned long long>(__double_as_longlong(s[4][0])));
    if (sb[0]) atomicMax(&res[5], 1ull);
    atomicMax(&res[6], static_cast<unsigned long long>(__double_as_longlong(s[5][0])));
    atomicMax(&res[7], static_cast<unsigned long long>(__double_as_longlong(s[6][0])));
  }
}

}  // namespace vibe_mi
