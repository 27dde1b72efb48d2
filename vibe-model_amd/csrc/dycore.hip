// dycore.hip -- the nonhydrostatic dynamical-core step on gfx950.
//
// One WS-RK3 step (src/time/runge_kutta.cpp:77-118 of the reference) is:
//   for each stage h in {dt/3, dt/2, dt}:
//     k_omega      contravariant vertical velocity of the current state
//     k_stage_*    slow tendencies R (equations.cpp:60-150 role, balanced form)
//                  fused with the RK3 update s = s0 + h R (FP64 accumulation)
//     nsub x (k_sub_h, k_sub_v)
//                  forward-backward acoustic substep, horizontal momentum
//                  explicit, w / pi' / rho' vertically implicit (HEVI) with
//                  one Thomas column solve per lane
//   k_sponge       backward-Euler Rayleigh sponge (damping.cpp:76-138)
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
//   theta'    WENO5 + LLF flux form minus theta' div V, - w dtheta0/dz.
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
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "geo.hpp"
#include "stencils.hpp"
#include "vibe_mi/dycore.h"

namespace vibe_mi {

// ---- kinematics -------------------------------------------------------------

template <typename T, typename S>
__device__ __forceinline__ T mc_at(const Geo<T> &g, const Acc &A, const S *__restrict__ u, const S *__restrict__ v,
                                   int i, int j, int k) {
  return T(0.5) * (T(u[A.at(i, j, k)]) * g.sx[gcol(g, i, j)] + T(u[A.at(i + 1, j, k)]) * g.sx[gcol(g, i + 1, j)]) +
         T(0.5) * (T(v[A.at(i, j, k)]) * g.sy[gcol(g, i, j)] + T(v[A.at(i, j + 1, k)]) * g.sy[gcol(g, i, j + 1)]);
}

template <typename T, typename S>
__global__ void __launch_bounds__(256) k_omega(Geo<T> g, const S *__restrict__ u, const S *__restrict__ v,
                                               const S *__restrict__ w, T *__restrict__ om) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const Acc A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  T o = T(0);
  if (k > 0 && k < g.nz)
    o = T(w[A.at(i, j, k)]) -
        (T(1) - g.zeta[k]) * (T(0.5) * (mc_at<T, S>(g, A, u, v, i, j, k - 1) + mc_at<T, S>(g, A, u, v, i, j, k)));
  om[A.at(i, j, k)] = o;
}

// ---- pressure-gradient operators (Gal-Chen, along constant z) --------------

template <typename T, typename S>
__device__ __forceinline__ T gz(const Geo<T> &g, const Acc &A, const S *__restrict__ pi, int i, int j, int k) {
  if (k <= 0 || k >= g.nz) return T(0);
  return (T(pi[A.at(i, j, k)]) - T(pi[A.at(i, j, k - 1)])) * g.rdzw[k];
}
template <typename T, typename S>
__device__ __forceinline__ T gzc(const Geo<T> &g, const Acc &A, const S *__restrict__ pi, int i, int j, int k) {
  return T(0.5) * (gz<T, S>(g, A, pi, i, j, k) + gz<T, S>(g, A, pi, i, j, k + 1));
}
template <typename T, typename S>
__device__ __forceinline__ T px_op(const Geo<T> &g, const Acc &A, const S *__restrict__ pi, int i, int j, int k) {
  const T dpx = (T(pi[A.at(i, j, k)]) - T(pi[A.at(i - 1, j, k)])) * g.rdx;
  const T met = (T(1) - g.zc[k]) * g.metx[gcol(g, i, j)];
  return dpx - met * (T(0.5) * (gzc<T, S>(g, A, pi, i - 1, j, k) + gzc<T, S>(g, A, pi, i, j, k)));
}
template <typename T, typename S>
__device__ __forceinline__ T py_op(const Geo<T> &g, const Acc &A, const S *__restrict__ pi, int i, int j, int k) {
  const T dpy = (T(pi[A.at(i, j, k)]) - T(pi[A.at(i, j - 1, k)])) * g.rdy;
  const T met = (T(1) - g.zc[k]) * g.mety[gcol(g, i, j)];
  return dpy - met * (T(0.5) * (gzc<T, S>(g, A, pi, i, j - 1, k) + gzc<T, S>(g, A, pi, i, j, k)));
}

// ---- fused z-marching stage kernel -------------------------------------------
// One thread per column (i, j), marching up the levels.  Every vertical stencil
// (5th-order momentum, WENO5 theta', pressure gradient, Omega averages) reads
// a register window of the thread's own column; horizontal neighbours are read
// at the current level only, so a workgroup's 64 x 4 tile of columns shares
// them through L1/L2 while it marches.  Writes u, v, w, rho', pi' (type O) and
// theta' (FP64) of s = s0 + h R in one pass.  Each interface flux is evaluated
// once: the flux through the top of level k is carried to level k+1 as its
// bottom flux (the same expression on the same operands).
struct StageIn {
  const double *u, *v, *w, *rho, *th, *pi;
};
template <typename O>
struct StageOut {
  O *u, *v, *w, *rho, *pi;
  double *th;
};

template <typename T, typename O>
__global__ void __launch_bounds__(256) k_stage_col(Geo<T> g, StageIn cur, StageIn s0, const T *__restrict__ om,
                                                   double h, StageOut<O> out, int *__restrict__ flag) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz, rk = g.rk;
  const size_t nxy = static_cast<size_t>(nx) * ny;
  const double *__restrict__ U = cur.u;
  const double *__restrict__ V = cur.v;
  const double *__restrict__ W = cur.w;
  const double *__restrict__ RH = cur.rho;
  const double *__restrict__ TH = cur.th;
  const double *__restrict__ PI = cur.pi;
  uint32_t xi[7], yj[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    xi[d] = static_cast<uint32_t>(wrapi(i + d - 3, nx));
    yj[d] = static_cast<uint32_t>(wrapi(j + d - 3, ny)) * static_cast<uint32_t>(nx);
  }
  const uint32_t cC = yj[3] + xi[3], cW = yj[3] + xi[2], cS = yj[2] + xi[3];
  const uint32_t cE = yj[3] + xi[4], cN = yj[4] + xi[3];
  const uint32_t cWN = yj[4] + xi[2], cES = yj[2] + xi[4];
  // column geometry (compacted)
  const int gC = gcol(g, i, j), gW = gcol(g, i - 1, j), gE = gcol(g, i + 1, j), gS = gcol(g, i, j - 1),
            gN = gcol(g, i, j + 1), gWN = gcol(g, i - 1, j + 1), gES = gcol(g, i + 1, j - 1);
  const T DuW = g.Du[gW], DuC = g.Du[gC], DuE = g.Du[gE], DuES = g.Du[gES];
  const T DvS = g.Dv[gS], DvC = g.Dv[gC], DvN = g.Dv[gN], DvWN = g.Dv[gWN], DvW = g.Dv[gW];
  const T rDuC = g.rDu[gC], rDvC = g.rDv[gC], rD = g.rDc[gC], mxC = g.metx[gC], myC = g.mety[gC];

  auto ldd = [&](const double *f, int k, uint32_t c) { return T(ld32(f, static_cast<size_t>(k) * nxy, c)); };
  auto ldo = [&](int k, uint32_t c) { return ld32(om, static_cast<size_t>(k) * nxy, c); };
  T uw[7], vw[7], ww[7], tw[7], rw[3], pw[3], pWw[3], pSw[3], ow[3], oW[2], oS[2];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int kc = clampk(q - 3, nz - 1);
    uw[q] = ldd(U, kc, cC);
    vw[q] = ldd(V, kc, cC);
    tw[q] = ldd(TH, kc, cC);
    ww[q] = ldd(W, clampk(q - 3, nz), cC);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int kc = clampk(q - 1, nz - 1);
    rw[q] = ldd(RH, kc, cC);
    pw[q] = ldd(PI, kc, cC);
    pWw[q] = ldd(PI, kc, cW);
    pSw[q] = ldd(PI, kc, cS);
    ow[q] = ldo(clampk(q - 1, nz), cC);
  }
  oW[0] = ldo(0, cW);
  oW[1] = ldo(1, cW);
  oS[0] = ldo(0, cS);
  oS[1] = ldo(1, cS);
  T uE_prev = T(0), vN_prev = T(0);
  // z-face fluxes of the level above, reused as the next level's lower face
  T cu_O = T(0), cu_f = T(0), cv_O = T(0), cv_f = T(0), cw_O = T(0), cw_f = T(0);
  T cm_z = T(0), cr_z = T(0), ct_z = T(0);
  int bad = 0;
  st32(out.w, 0, cC, static_cast<O>(ld32(s0.w, 0, cC)));  // interface 0 keeps its value (R_w = 0)

  for (int k = 0; k < nz; ++k) {
    const size_t kb = static_cast<size_t>(k) * nxy;
    const int rb = k * rk;
    // horizontal neighbours at level k
    T ux[7], uy[7], vx[7], vy[7], tx[7], ty[7], wx[7], wy[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) {
      if (d == 3) continue;
      ux[d] = T(ld32(U, kb, yj[3] + xi[d]));
      uy[d] = T(ld32(U, kb, yj[d] + xi[3]));
      vx[d] = T(ld32(V, kb, yj[3] + xi[d]));
      vy[d] = T(ld32(V, kb, yj[d] + xi[3]));
      tx[d] = T(ld32(TH, kb, yj[3] + xi[d]));
      ty[d] = T(ld32(TH, kb, yj[d] + xi[3]));
      wx[d] = T(ld32(W, kb, yj[3] + xi[d]));
      wy[d] = T(ld32(W, kb, yj[d] + xi[3]));
    }
    ux[3] = uy[3] = uw[3];
    vx[3] = vy[3] = vw[3];
    tx[3] = ty[3] = tw[3];
    wx[3] = wy[3] = ww[3];
    const T vWN = T(ld32(V, kb, cWN)), uES = T(ld32(U, kb, cES));
    const T rWst = T(ld32(RH, kb, cW)), rE = T(ld32(RH, kb, cE)), rSth = T(ld32(RH, kb, cS)), rN = T(ld32(RH, kb, cN));
    const T omz = T(1) - g.zc[k];
    const T rdzk = g.rdzeta[k];
    // vertical pi' gradients (0 at ground and lid)
    auto gpair = [&](const T *p) {
      const T glo = (k >= 1) ? (p[1] - p[0]) * g.rdzw[k] : T(0);
      const T ghi = (k + 1 <= nz - 1) ? (p[2] - p[1]) * g.rdzw[k + 1] : T(0);
      return T(0.5) * (glo + ghi);
    };
    const T GcC = gpair(pw), GcW = gpair(pWw), GcS = gpair(pSw);

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
      const T px = (pw[1] - pWw[1]) * g.rdx - (omz * mxC) * (T(0.5) * (GcW + GcC));
      const T R = adv * rDuC - T(kCp) * thu * px;
      st32(out.u, kb, cC, static_cast<O>(ld32(s0.u, kb, cC) + h * static_cast<double>(R)));
    }
    // ---------------- v at the south face ----------------
    {
      const T U0 = T(0.5) * (g.Du[gcol(g, i, j - 1)] * uy[2] + DuC * uw[3]);
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
      const T py = (pw[1] - pSw[1]) * g.rdy - (omz * myC) * (T(0.5) * (GcS + GcC));
      const T R = adv * rDvC - T(kCp) * thv * py;
      st32(out.v, kb, cC, static_cast<O>(ld32(s0.v, kb, cC) + h * static_cast<double>(R)));
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
      const T rz = g.rdzw[k];
      const T adv = -(fx1 - fx0) * g.rdx - (fy1 - fy0) * g.rdy - (fz1 - fz0) * rz +
                    ww[3] * ((U1 - U0) * g.rdx + (V1 - V0) * g.rdy + (Oc1 - Oc0) * rz);
      const T thw = T(0.5) * (tw[2] + tw[3]);
      const T th0w = T(0.5) * (g.theta0[rb - rk + gC] + g.theta0[rb + gC]);
      const T pz = (pw[1] - pw[0]) * rz * rD;
      const T R = adv * rD - T(kCp) * thw * pz + T(kG) * thw / th0w;
      st32(out.w, kb, cC, static_cast<O>(ld32(s0.w, kb, cC) + h * static_cast<double>(R)));
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
      auto gam = [&](int kq) {
        const int kk = kq < 1 ? 1 : (kq > nz - 1 ? nz - 1 : kq);
        return (g.theta0[kk * rk + gC] - g.theta0[(kk - 1) * rk + gC]) * rD * g.rdzw[kk];
      };
      const T rt = -(fd - tw[3] * div) * rD - T(0.5) * (ww[3] * gam(k) + ww[4] * gam(k + 1));
      const T rp = g.kap[rb + gC] * rr + g.lam[rb + gC] * rt;
      const double rn = ld32(s0.rho, kb, cC) + h * static_cast<double>(rr);
      const double tn = ld32(s0.th, kb, cC) + h * static_cast<double>(rt);
      const double pn = ld32(s0.pi, kb, cC) + h * static_cast<double>(rp);
      st32(out.rho, kb, cC, static_cast<O>(rn));
      st32(out.th, kb, cC, tn);
      st32(out.pi, kb, cC, static_cast<O>(pn));
      bad |= !isfinite(tn);
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
    uw[6] = ldd(U, kn, cC);
    vw[6] = ldd(V, kn, cC);
    tw[6] = ldd(TH, kn, cC);
    ww[6] = ldd(W, kw, cC);
    const int k2 = clampk(k + 2, nz - 1);
    rw[0] = rw[1];
    rw[1] = rw[2];
    rw[2] = ldd(RH, k2, cC);
    pw[0] = pw[1];
    pw[1] = pw[2];
    pw[2] = ldd(PI, k2, cC);
    pWw[0] = pWw[1];
    pWw[1] = pWw[2];
    pWw[2] = ldd(PI, k2, cW);
    pSw[0] = pSw[1];
    pSw[1] = pSw[2];
    pSw[2] = ldd(PI, k2, cS);
    const int ko = clampk(k + 2, nz);
    ow[0] = ow[1];
    ow[1] = ow[2];
    ow[2] = ldo(ko, cC);
    oW[0] = oW[1];
    oW[1] = ldo(ko, cW);
    oS[0] = oS[1];
    oS[1] = ldo(ko, cS);
  }
  st32(out.w, static_cast<size_t>(nz) * nxy, cC, static_cast<O>(ld32(s0.w, static_cast<size_t>(nz) * nxy, cC)));
  if (bad) atomicOr(flag, 1);
}

// ---- HEVI factorisation table ----------------------------------------------
// The tridiagonal matrix of the vertically implicit w system depends only on
// the reference state, the column depth and dtau, so it is factorised once
// per (stage, column): beta_m = a_m/pivot_m, 1/pivot_m and c'_m, stored
// [m][compact column].  One thread per compact column.
template <typename T>
__global__ void __launch_bounds__(64) k_hevi_factor(Geo<T> g, T dtau, T an, T *__restrict__ tA, T *__restrict__ tINV,
                                                    T *__restrict__ tCP) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.rk) return;
  const int nz = g.nz, n = nz - 1, rk = g.rk;
  const T dtD = dtau * g.rDc[c];
  const T an2 = an * an;
  auto ck = [&](int k) { return g.kap[k * rk + c] * dtD * g.rdzeta[k]; };
  auto r0w = [&](int k) { return T(0.5) * (g.rho0[(k - 1) * rk + c] + g.rho0[k * rk + c]); };
  T cp_prev = T(0);
  for (int m = 0; m < n; ++m) {
    const int k = m + 1;
    const T th0w = T(0.5) * (g.theta0[(k - 1) * rk + c] + g.theta0[k * rk + c]);
    const T gk = dtD * T(kCp) * th0w * g.rdzw[k];
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
template <typename T, typename I, typename O, int NW, int MAXCH>
__global__ void __launch_bounds__(64 * NW) k_substep(Geo<T> g, const I *__restrict__ u, const I *__restrict__ v,
                                                     const I *__restrict__ w, const I *__restrict__ rho,
                                                     const I *__restrict__ pi, const T *__restrict__ tA,
                                                     const T *__restrict__ tINV, const T *__restrict__ tCP, T dtau,
                                                     T an, T ao, O *__restrict__ uo, O *__restrict__ vo,
                                                     O *__restrict__ wo, O *__restrict__ rho_o, O *__restrict__ pi_o,
                                                     int check, int *__restrict__ flag) {
  extern __shared__ unsigned char smem_raw[];
  const int nlds = (g.nz - 1) * 64;
  T *s_td = reinterpret_cast<T *>(smem_raw);  // [nz-1][64] RHS -> d' -> w
  T *s_be = s_td + nlds;                      // [nz-1][64] beta
  T *s_cp = s_be + nlds;                      // [nz-1][64] c'
  __shared__ T s_mc[NW][64], s_fh[NW][64], s_ph[NW][64];
  const int lane = threadIdx.x & 63;
  // wave index: uniform within the wavefront, so the level bases below live in
  // scalar registers and loads use the scalar-base + 32-bit offset form
  const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nx = g.nx, ny = g.ny, nz = g.nz, rk = g.rk;
  const size_t nxy = static_cast<size_t>(nx) * ny;
  const int i_raw = blockIdx.x * 64 + lane;
  const bool active = i_raw < nx;
  const int i = active ? i_raw : nx - 1;
  const int j = blockIdx.y;
  const int ch = (nz + NW - 1) / NW;
  const int k0 = wv * ch;
  const int nlev = max(0, min(nz, k0 + ch) - k0);
  const int iW = wrapi(i - 1, nx), iE = wrapi(i + 1, nx), jS = wrapi(j - 1, ny), jN = wrapi(j + 1, ny);
  const uint32_t unx = static_cast<uint32_t>(nx);
  const uint32_t cC = static_cast<uint32_t>(j) * unx + i, cW = static_cast<uint32_t>(j) * unx + iW,
                 cE = static_cast<uint32_t>(j) * unx + iE, cS = static_cast<uint32_t>(jS) * unx + i,
                 cN = static_cast<uint32_t>(jN) * unx + i;
  const uint32_t gC = gcol(g, i, j), gW = gcol(g, iW, j), gE = gcol(g, iE, j), gS = gcol(g, i, jS), gN = gcol(g, i, jN);
  const T rD = g.rDc[gC];
  const T dtD = dtau * rD;
  const T Du_w = g.Du[gC], Du_e = g.Du[gE], Dv_s = g.Dv[gC], Dv_n = g.Dv[gN];
  const T sx_w = g.sx[gC], sx_e = g.sx[gE], sy_s = g.sy[gC], sy_n = g.sy[gN];
  const T mx_w = g.metx[gC], mx_e = g.metx[gE], my_s = g.mety[gC], my_n = g.mety[gN];
  const T cpdt = dtau;

  T E[MAXCH], Q[MAXCH], Wt[MAXCH], Fh[MAXCH], Ph[MAXCH];
  T mc_first = T(0), mc_last = T(0), mc0 = T(0), aw_first = T(0), r0w_first = T(0);

  // ---------------- phase A ----------------
  if (nlev > 0) {
    auto ld5 = [&](int k, T *p) {
      const size_t b = static_cast<size_t>(k) * nxy;
      p[0] = T(ld32(pi, b, cC));
      p[1] = T(ld32(pi, b, cW));
      p[2] = T(ld32(pi, b, cE));
      p[3] = T(ld32(pi, b, cS));
      p[4] = T(ld32(pi, b, cN));
    };
    // pi' of the five columns at the current and next level; the vertical
    // gradient of the interface below is carried from the previous level.
    T p0[5], pp[5], gl[5];
    ld5(k0, p0);
    T pmC = p0[0];
    if (k0 >= 1) {
      T pm[5];
      ld5(k0 - 1, pm);
      pmC = pm[0];
#pragma unroll
      for (int q = 0; q < 5; ++q) gl[q] = (p0[q] - pm[q]) * g.rdzw[k0];
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) gl[q] = T(0);
    }
    T t0C_prev = (k0 >= 1) ? ld32(g.theta0, static_cast<size_t>(k0 - 1) * rk, gC) : T(0);
    T r0C_prev = (k0 >= 1) ? ld32(g.rho0, static_cast<size_t>(k0 - 1) * rk, gC) : T(0);
    T mc_prev = T(0);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      if (c < nlev) {
        const int k = k0 + c;
        const size_t b = static_cast<size_t>(k) * nxy;
        if (k + 1 < nz) ld5(k + 1, pp); else for (int q = 0; q < 5; ++q) pp[q] = p0[q];
        const T rz_hi = g.rdzw[k + 1];
        T Gc[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          const T ghi = (k + 1 <= nz - 1) ? (pp[q] - p0[q]) * rz_hi : T(0);
          Gc[q] = T(0.5) * (gl[q] + ghi);
          gl[q] = ghi;
        }
        const T omz = T(1) - g.zc[k];
        const T px_w = (p0[0] - p0[1]) * g.rdx - (omz * mx_w) * (T(0.5) * (Gc[1] + Gc[0]));
        const T px_e = (p0[2] - p0[0]) * g.rdx - (omz * mx_e) * (T(0.5) * (Gc[0] + Gc[2]));
        const T py_s = (p0[0] - p0[3]) * g.rdy - (omz * my_s) * (T(0.5) * (Gc[3] + Gc[0]));
        const T py_n = (p0[4] - p0[0]) * g.rdy - (omz * my_n) * (T(0.5) * (Gc[0] + Gc[4]));
        const int rb = k * rk;
        const T t0C = ld32(g.theta0, static_cast<size_t>(rb), gC), t0W = ld32(g.theta0, static_cast<size_t>(rb), gW), t0E = ld32(g.theta0, static_cast<size_t>(rb), gE), t0S = ld32(g.theta0, static_cast<size_t>(rb), gS),
                t0N = ld32(g.theta0, static_cast<size_t>(rb), gN);
        const T un_w = T(ld32(u, b, cC)) - cpdt * (T(kCp) * (T(0.5) * (t0W + t0C)) * px_w);
        const T un_e = T(ld32(u, b, cE)) - cpdt * (T(kCp) * (T(0.5) * (t0C + t0E)) * px_e);
        const T vn_s = T(ld32(v, b, cC)) - cpdt * (T(kCp) * (T(0.5) * (t0S + t0C)) * py_s);
        const T vn_n = T(ld32(v, b, cN)) - cpdt * (T(kCp) * (T(0.5) * (t0C + t0N)) * py_n);
        if (active) {
          st32(uo, b, cC, static_cast<O>(un_w));
          st32(vo, b, cC, static_cast<O>(vn_s));
        }
        const T mc = T(0.5) * (un_w * sx_w + un_e * sx_e) + T(0.5) * (vn_s * sy_s + vn_n * sy_n);
        const T r0C = ld32(g.rho0, static_cast<size_t>(rb), gC), r0W = ld32(g.rho0, static_cast<size_t>(rb), gW), r0E = ld32(g.rho0, static_cast<size_t>(rb), gE), r0S = ld32(g.rho0, static_cast<size_t>(rb), gS),
                r0N = ld32(g.rho0, static_cast<size_t>(rb), gN);
        const T hd = (Du_e * (T(0.5) * (r0C + r0E)) * un_e - Du_w * (T(0.5) * (r0W + r0C)) * un_w) * g.rdx +
                     (Dv_n * (T(0.5) * (r0C + r0N)) * vn_n - Dv_s * (T(0.5) * (r0S + r0C)) * vn_s) * g.rdy;
        const T cq = ld32(g.kap, static_cast<size_t>(rb), gC) * dtD;
        E[c] = T(ld32(rho, b, cC)) - dtD * hd;
        Q[c] = p0[0] - cq * hd;
        const T wk = T(ld32(w, b, cC));
        Fh[c] = T(0);
        Wt[c] = T(0);
        if (k >= 1) {
          const T th0w = T(0.5) * (t0C_prev + t0C);
          const T gk = dtD * T(kCp) * th0w * g.rdzw[k];
          Wt[c] = wk - gk * ao * (p0[0] - pmC);
          const T r0w = T(0.5) * (r0C_prev + r0C);
          if (c > 0) {
            const T Mw = (T(1) - g.zeta[k]) * (T(0.5) * (mc_prev + mc));
            Fh[c] = r0w * (ao * wk - Mw);
          } else {
            aw_first = ao * wk;
            r0w_first = r0w;
          }
        }
        if (c == 0) mc_first = mc;
        if (k == 0) mc0 = mc;
        mc_prev = mc;
        t0C_prev = t0C;
        r0C_prev = r0C;
        pmC = p0[0];
#pragma unroll
        for (int q = 0; q < 5; ++q) p0[q] = pp[q];
      }
    }
    mc_last = mc_prev;
    s_mc[wv][lane] = mc_last;
  }
  __syncthreads();
  // ---------------- phase B: boundary vertical flux, P^, RHS ----------------
  if (nlev > 0) {
    if (k0 >= 1) {
      const T Mw = (T(1) - g.zeta[k0]) * (T(0.5) * (s_mc[wv - 1][lane] + mc_first));
      Fh[0] = r0w_first * (aw_first - Mw);
    }
    s_fh[wv][lane] = Fh[0];
  }
  __syncthreads();
  if (nlev > 0) {
    const T fh_next_chunk = (k0 + nlev < nz) ? s_fh[wv + 1][lane] : T(0);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      if (c < nlev) {
        const int k = k0 + c;
        const T fhn = (c + 1 < nlev) ? Fh[(c + 1) < MAXCH ? c + 1 : c] : fh_next_chunk;
        const T ck = ld32(g.kap, static_cast<size_t>(k) * rk, gC) * dtD * g.rdzeta[k];
        Ph[c] = Q[c] - ck * (fhn - Fh[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < MAXCH; ++c)
      if (c == nlev - 1) s_ph[wv][lane] = Ph[c];
  }
  __syncthreads();
  if (nlev > 0) {
    const T ph_prev_chunk = (k0 >= 1) ? s_ph[wv - 1][lane] : T(0);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      if (c < nlev) {
        const int k = k0 + c;
        if (k >= 1) {
          const T php = (c > 0) ? Ph[c > 0 ? c - 1 : 0] : ph_prev_chunk;
          const T th0w = T(0.5) * (ld32(g.theta0, static_cast<size_t>(k - 1) * rk, gC) + ld32(g.theta0, static_cast<size_t>(k) * rk, gC));
          const T gk = dtD * T(kCp) * th0w * g.rdzw[k];
          const int m = k - 1;
          const T td = Wt[c] - gk * an * (Ph[c] - php);
          s_td[m * 64 + lane] = td * ld32(tINV, static_cast<size_t>(m) * rk, gC);
          s_be[m * 64 + lane] = ld32(tA, static_cast<size_t>(m) * rk, gC);
          s_cp[m * 64 + lane] = ld32(tCP, static_cast<size_t>(m) * rk, gC);
        }
      }
    }
  }
  __syncthreads();
  // ---------------- phase C: Thomas sweeps (wave 0) ----------------
  // RHS, beta and c' are in LDS (staged by all waves in phase B); operands
  // are fetched one block of eight levels ahead so only the dependent
  // multiply-subtract per level stays on the critical path.
  if (wv == 0) {
    constexpr int B = 8;
    const int n = nz - 1;
    const int nb = (n + B - 1) / B;
    T ba[B], ta[B], bb[B], tb2[B];
    auto fetch_fwd = [&](int blk, T *be, T *td) {
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const int m = blk * B + q;
        const int mm = (m < n ? m : n - 1) * 64 + lane;
        be[q] = s_be[mm];
        td[q] = s_td[mm];
      }
    };
    T dp = T(0);
    fetch_fwd(0, ba, ta);
    for (int blk = 0; blk < nb; ++blk) {
      if (blk + 1 < nb) fetch_fwd(blk + 1, bb, tb2);
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const int m = blk * B + q;
        if (m < n) {
          dp = (m == 0) ? ta[q] : ta[q] - ba[q] * dp;
          s_td[m * 64 + lane] = dp;
        }
      }
#pragma unroll
      for (int q = 0; q < B; ++q) {
        ba[q] = bb[q];
        ta[q] = tb2[q];
      }
    }
    auto fetch_bwd = [&](int blk, T *cp, T *d) {
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const int m = n - 2 - (blk * B + q);
        const int mm = (m >= 0 ? m : 0) * 64 + lane;
        cp[q] = s_cp[mm];
        d[q] = s_td[mm];
      }
    };
    T x = dp;
    const int nbb = (n - 1 + B - 1) / B;
    if (nbb > 0) fetch_bwd(0, ba, ta);
    for (int blk = 0; blk < nbb; ++blk) {
      if (blk + 1 < nbb) fetch_bwd(blk + 1, bb, tb2);
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const int m = n - 2 - (blk * B + q);
        if (m >= 0) {
          x = ta[q] - ba[q] * x;
          s_td[m * 64 + lane] = x;
        }
      }
#pragma unroll
      for (int q = 0; q < B; ++q) {
        ba[q] = bb[q];
        ta[q] = tb2[q];
      }
    }
  }
  __syncthreads();
  // ---------------- phase D: close rho', pi', write w ----------------
  if (nlev > 0) {
    int bad = 0;
    auto fluxF = [&](int k, T fh) {
      if (k <= 0 || k >= nz) return T(0);
      const T r0w = T(0.5) * (ld32(g.rho0, static_cast<size_t>(k - 1) * rk, gC) + ld32(g.rho0, static_cast<size_t>(k) * rk, gC));
      return fh + (an * r0w) * s_td[(k - 1) * 64 + lane];
    };
    const T fh_next_chunk = (k0 + nlev < nz) ? s_fh[wv + 1][lane] : T(0);
    T flo = fluxF(k0, Fh[0]);
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
      if (c < nlev) {
        const int k = k0 + c;
        const size_t b = static_cast<size_t>(k) * nxy;
        const T fhn = (c + 1 < nlev) ? Fh[(c + 1) < MAXCH ? c + 1 : c] : fh_next_chunk;
        const T fhi = fluxF(k + 1, fhn);
        const T vd = -dtD * ((fhi - flo) * g.rdzeta[k]);
        const T rn = E[c] + vd;
        const T pn = Q[c] + ld32(g.kap, static_cast<size_t>(k) * rk, gC) * vd;
        const T wn = (k >= 1) ? s_td[(k - 1) * 64 + lane] : mc0;
        if (active) {
          st32(rho_o, b, cC, static_cast<O>(rn));
          st32(pi_o, b, cC, static_cast<O>(pn));
          st32(wo, b, cC, static_cast<O>(wn));
        }
        bad |= !(isfinite(rn) && isfinite(pn) && isfinite(wn));
        flo = fhi;
      }
    }
    if (k0 + nlev == nz && active) {
      const size_t b = static_cast<size_t>(nz) * nxy;
      st32(wo, b, cC, static_cast<O>(T(ld32(w, b, cC))));
    }
    if (check && bad && active) atomicOr(flag, 1);
  }
}

// ---- sponge (damping.cpp:76-138 with perturbation targets 0) -------------

__global__ void __launch_bounds__(256) k_sponge(int nx, int ny, int nz, int kmin, const double *__restrict__ alpha,
                                                double dt, double *__restrict__ u, double *__restrict__ v,
                                                double *__restrict__ w, double *__restrict__ rho,
                                                double *__restrict__ th, double *__restrict__ pi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = kmin + static_cast<int>(blockIdx.z);
  if (i >= nx || k > nz) return;
  const double a = alpha[k < nz - 1 ? k : nz - 1];
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

// ---- diagnostics ------------------------------------------------------------

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;
constexpr int kEnergyComps = 5;

// Per-cell terms of energy_budget (diagnostics.cpp:255-277), compensated.
__global__ void __launch_bounds__(kRedThreads) k_energy_partial(
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

__global__ void __launch_bounds__(kRedThreads) k_red_final(const NSum *__restrict__ part, int nparts, int ncomp,
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

// Advective CFL of the state (advection.cpp:325-340) and max wind speed.
__global__ void __launch_bounds__(256) k_cfl_state(int nx, int ny, int nz, double dx, double dy, double H,
                                                   const double *__restrict__ zs, const double *__restrict__ dzeta,
                                                   const double *__restrict__ u, const double *__restrict__ v,
                                                   const double *__restrict__ w, double dt,
                                                   unsigned long long *__restrict__ res) {
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * nz;
  double worst = 0.0, wind = 0.0;
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
  }
  __shared__ double s0[256], s1[256];
  s0[threadIdx.x] = worst;
  s1[threadIdx.x] = wind;
  __syncthreads();
  for (int hh = 128; hh > 0; hh >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(hh)) {
      s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + hh]);
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + hh]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(&res[0], static_cast<unsigned long long>(__double_as_longlong(s0[0])));
    atomicMax(&res[1], static_cast<unsigned long long>(__double_as_longlong(s1[0])));
  }
}

}  // namespace vibe_mi

using namespace vibe_mi;

// ---------------------------------------------------------------------------
// domain handle
// ---------------------------------------------------------------------------

struct FieldSet {
  double *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *th = nullptr, *pi = nullptr;
};
template <typename T>
struct AcousticSet {
  T *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *pi = nullptr;
};

struct vibe_mi_domain {
  int nx = 0, ny = 0, nz = 0;
  size_t nxy = 0, ncell = 0, nwf = 0;
  double dx = 0, dy = 0, ztop = 0, eps_s = 0;
  int precision = VIBE_MI_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  size_t bytes = 0;
  // FP64 geometry / reference for diagnostics and the FP64 path
  Geo<double> g64{};
  Geo<float> g32{};
  double *zs = nullptr, *sponge = nullptr, *rho0 = nullptr, *theta0 = nullptr, *pi0 = nullptr;
  double *dzeta64 = nullptr, *zc64 = nullptr;  // diagnostics (FP64, per level)
  std::vector<double> h_sponge;
  int kmin_sponge = 0;
  FieldSet S[3];  // [0] = s0 of the step; [1], [2] stage outputs
  int cur0 = 0;   // index of s^n
  AcousticSet<double> A64[2];
  AcousticSet<float> A32[2];
  void *om = nullptr;   // Omega (T)
  void *tab = nullptr;  // HEVI factorisation tables: 4 slots x {a, 1/pivot, c'} x (nz-1) x rk (T)
  double tab_dtau[4] = {-1.0, -1.0, -1.0, -1.0};
  int nxr = 1, nyr = 1, rk = 1;
  int nw_fp32 = 4, nw_fp64 = 8;  // waves per substep workgroup (autotuned)
  bool tuned[2] = {false, false};
  int *flag = nullptr;
  NSum *part = nullptr;
  double *dres = nullptr;
  unsigned long long *dmax = nullptr;
  // timing
  std::vector<hipEvent_t> ev;
  double ms_stage = 0, ms_sub = 0;
  int n_stage = 0, n_sub = 0;
};

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
int run_omega(vibe_mi_domain *d, const Geo<T> &g, const FieldSet &s) {
  hipLaunchKernelGGL((k_omega<T, double>), grid_of(d, d->nz + 1), dim3(kBlk), 0, d->stream, g, s.u, s.v, s.w,
                     static_cast<T *>(d->om));
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int run_stage(vibe_mi_domain *d, const Geo<T> &g, const FieldSet &cur, const FieldSet &s0, double h,
              AcousticSet<T> &a, FieldSet &nxt) {
  if (int rc = run_omega<T>(d, g, cur)) return rc;
  const StageIn ci{cur.u, cur.v, cur.w, cur.rho, cur.th, cur.pi};
  const StageIn si{s0.u, s0.v, s0.w, s0.rho, s0.th, s0.pi};
  const StageOut<T> so{a.u, a.v, a.w, a.rho, a.pi, nxt.th};
  hipLaunchKernelGGL((k_stage_col<T, T>), dim3(ceil_div(d->nx, 64), ceil_div(d->ny, 4)), dim3(64, 4), 0, d->stream, g,
                     ci, si, static_cast<const T *>(d->om), h, so, d->flag);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
T *table_slot(vibe_mi_domain *d, int slot) {
  const size_t per = static_cast<size_t>(d->nz > 1 ? d->nz - 1 : 1) * d->rk;
  return static_cast<T *>(d->tab) + static_cast<size_t>(slot) * 3 * per;
}

// Factorise the HEVI matrix for dtau into table slot `slot` (cached).
template <typename T>
int ensure_table(vibe_mi_domain *d, const Geo<T> &g, int slot, double dtau) {
  if (d->tab_dtau[slot] == dtau) return 0;
  const size_t per = static_cast<size_t>(d->nz - 1) * d->rk;
  T *t = table_slot<T>(d, slot);
  const T an = static_cast<T>(0.5 * (1.0 + d->eps_s));
  hipLaunchKernelGGL((k_hevi_factor<T>), dim3(ceil_div(d->rk, 64)), dim3(64), 0, d->stream, g, static_cast<T>(dtau),
                     an, t, t + per, t + 2 * per);
  VMI_HIP(hipGetLastError());
  d->tab_dtau[slot] = dtau;
  return 0;
}

constexpr size_t kMaxLds = 160 * 1024;  // LDS per CU (and per workgroup) on gfx950

template <typename T, typename I, typename O, int NW, int MAXCH>
int launch_substep(vibe_mi_domain *d, const Geo<T> &g, const T *tab, const I *u, const I *v, const I *w,
                   const I *rho, const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, T dtau, T an, T ao,
                   int check) {
  const size_t per = static_cast<size_t>(d->nz - 1) * d->rk;
  const size_t lds = 3 * static_cast<size_t>(d->nz - 1) * 64 * sizeof(T);
  const size_t lds_static = 3 * NW * 64 * sizeof(T);
  if (lds + lds_static > kMaxLds) return VIBE_C_ERR_BAD_ARG;  // column too tall for one workgroup
  static bool attr_set = false;  // once per instantiation
  if (!attr_set) {
    VMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_substep<T, I, O, NW, MAXCH>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds - lds_static)));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_substep<T, I, O, NW, MAXCH>), dim3(ceil_div(d->nx, 64), d->ny), dim3(64 * NW), lds, d->stream,
                     g, u, v, w, rho, pi, tab, tab + per, tab + 2 * per, dtau, an, ao, uo, vo, wo, rho_o, pi_o, check,
                     d->flag);
  return 0;
}

template <typename T, typename I, typename O, int NW>
int dispatch_substep(vibe_mi_domain *d, const Geo<T> &g, const T *tab, const I *u, const I *v, const I *w,
                     const I *rho, const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, T dtau, T an, T ao, int check) {
  const int ch = (d->nz + NW - 1) / NW;
#define VMI_SUB(M) launch_substep<T, I, O, NW, M>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check)
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
               const I *rho, const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, T dtau, T an, T ao, int check) {
  if (nw == 16)
    return dispatch_substep<T, I, O, 16>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check);
  if (nw == 8)
    return dispatch_substep<T, I, O, 8>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check);
  return dispatch_substep<T, I, O, 4>(d, g, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, dtau, an, ao, check);
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
  const TuneKey key{d->nx, d->ny, d->nz, d->rk, static_cast<int>(sizeof(T))};
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
                               static_cast<T>(dtau), an, ao, 0) == 0;
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
                const I *pi, O *uo, O *vo, O *wo, O *rho_o, O *pi_o, double dtau, int check) {
  if (int rc = ensure_table<T>(d, g, slot, dtau)) return rc;
  const T an = static_cast<T>(0.5 * (1.0 + d->eps_s)), ao = static_cast<T>(0.5 * (1.0 - d->eps_s));
  const T *tab = table_slot<T>(d, slot);
  const int nw = sizeof(T) == 8 ? d->nw_fp64 : d->nw_fp32;
  return substep_nw<T, I, O>(d, g, nw, tab, u, v, w, rho, pi, uo, vo, wo, rho_o, pi_o, static_cast<T>(dtau), an, ao,
                             check);
}

// One full RK3 step.  Stage outputs go to S[(cur0+1)%3] / S[(cur0+2)%3].
template <typename T>
int step_impl(vibe_mi_domain *d, const Geo<T> &g, AcousticSet<T> *A, double dt, int nsub, bool timed) {
  const int ns = nsub < 1 ? 1 : nsub;
  const double hk[3] = {dt / 3.0, dt / 2.0, dt};
  if (int rc = autotune_substep<T>(d, g, A, dt / static_cast<double>(ns))) return rc;
  const int i0 = d->cur0, i1 = (d->cur0 + 1) % 3, i2 = (d->cur0 + 2) % 3;
  const FieldSet &s0 = d->S[i0];
  const int outs[3] = {i1, i2, i1};
  int cur = i0;
  for (int st = 0; st < 3; ++st) {
    FieldSet &nxt = d->S[outs[st]];
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
                                          out.pi, dtau, 0))
          return rc;
        p = 1 - p;
      } else {
        if (int rc = run_substep<T, T, double>(d, g, st, in.u, in.v, in.w, in.rho, in.pi, nxt.u, nxt.v, nxt.w,
                                               nxt.rho, nxt.pi, dtau, 1))
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
                       d->sponge, dt, fin.u, fin.v, fin.w, fin.rho, fin.th, fin.pi);
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
  if (d->precision == VIBE_MI_MIXED)
    if (int rc = build_geo<float>(d, ds, d->g32)) return rc;
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
  const size_t tsz = d->precision == VIBE_MI_MIXED ? sizeof(float) : sizeof(double);
  if (d->precision == VIBE_MI_MIXED) {
    for (auto &a : d->A32)
      if (int rc = alloc_acoustic<float>(d, a)) return rc;
  } else {
    for (auto &a : d->A64)
      if (int rc = alloc_acoustic<double>(d, a)) return rc;
  }
  char *p;
  if (int rc = dalloc(d, &p, d->nwf * tsz)) return rc;
  d->om = p;
  if (int rc = dalloc(d, &p, 4 * 3 * static_cast<size_t>(d->nz - 1) * d->rk * tsz)) return rc;
  d->tab = p;
  if (int rc = dalloc(d, &d->flag, 4)) return rc;
  if (int rc = dalloc(d, &d->part, static_cast<size_t>(kRedBlocks) * kEnergyComps)) return rc;
  if (int rc = dalloc(d, &d->dres, 8)) return rc;
  if (int rc = dalloc(d, &d->dmax, 4)) return rc;
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

extern "C" {

int vibe_mi_domain_create(const vibe_mi_domain_desc *ds, vibe_mi_domain **out) {
  if (ds == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  *out = nullptr;
  if (ds->nx < 4 || ds->ny < 4 || ds->nz < 3 || ds->dx <= 0 || ds->dy <= 0 || ds->ztop <= 0 ||
      ds->zeta == nullptr || ds->zs == nullptr || ds->rho0 == nullptr || ds->theta0 == nullptr ||
      ds->pi0 == nullptr || ds->sponge == nullptr ||
      (ds->precision != VIBE_MI_FP64 && ds->precision != VIBE_MI_MIXED))
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
    if (d->precision == VIBE_MI_MIXED)
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

int vibe_mi_domain_cfl(vibe_mi_domain *d, double dt, double *cfl, double *max_wind) {
  if (d == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (int rc = set_device(d)) return rc;
  const FieldSet &f = d->S[d->cur0];
  VMI_HIP(hipMemsetAsync(d->dmax, 0, 2 * sizeof(unsigned long long), d->stream));
  const unsigned blocks = std::min<size_t>(ceil_div(d->ncell, 256), 2048);
  hipLaunchKernelGGL(k_cfl_state, dim3(blocks), dim3(256), 0, d->stream, d->nx, d->ny, d->nz, d->dx, d->dy, d->ztop,
                     d->zs, d->dzeta64, f.u, f.v, f.w, dt, d->dmax);
  VMI_HIP(hipGetLastError());
  unsigned long long r[2];
  VMI_HIP(hipMemcpyAsync(r, d->dmax, sizeof(r), hipMemcpyDeviceToHost, d->stream));
  VMI_HIP(hipStreamSynchronize(d->stream));
  if (cfl) *cfl = __builtin_bit_cast(double, r[0]);
  if (max_wind) *max_wind = __builtin_bit_cast(double, r[1]);
  return 0;
}

void *vibe_mi_domain_stream(vibe_mi_domain *d) { return d ? static_cast<void *>(d->stream) : nullptr; }

size_t vibe_mi_domain_device_bytes(vibe_mi_domain *d) { return d ? d->bytes : 0; }

}  // extern "C"
