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
#include <cstring>
#include <vector>

#include "common.hpp"
#include "stencils.hpp"
#include "vibe_mi/dycore.h"

namespace vibe_mi {

template <typename T>
struct Geo {
  int nx, ny, nz;
  T dx, dy;
  const T *zeta;   // nz+1
  const T *dzeta;  // nz
  const T *zc;     // nz
  const T *dzw;    // nz+1 (1..nz-1 valid)
  const T *Dc, *Du, *Dv, *sx, *sy;  // nx*ny
  const T *rho0, *theta0, *kap, *lam;  // nz*ny*nx
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampk(int k, int hi) { return k < 0 ? 0 : (k > hi ? hi : k); }

template <typename T>
struct Acc {
  int nx, ny;
  size_t nxy;
  __device__ size_t h(int i, int j) const {
    return static_cast<size_t>(wrapi(j, ny)) * nx + wrapi(i, nx);
  }
  __device__ size_t at(int i, int j, int k) const { return static_cast<size_t>(k) * nxy + h(i, j); }
};

// ---- kinematics -------------------------------------------------------------

template <typename T, typename S>
__device__ __forceinline__ T mc_at(const Geo<T> &g, const Acc<T> &A, const S *__restrict__ u,
                                   const S *__restrict__ v, int i, int j, int k) {
  return T(0.5) * (T(u[A.at(i, j, k)]) * g.sx[A.h(i, j)] + T(u[A.at(i + 1, j, k)]) * g.sx[A.h(i + 1, j)]) +
         T(0.5) * (T(v[A.at(i, j, k)]) * g.sy[A.h(i, j)] + T(v[A.at(i, j + 1, k)]) * g.sy[A.h(i, j + 1)]);
}

template <typename T, typename S>
__global__ void __launch_bounds__(256) k_omega(Geo<T> g, const S *__restrict__ u, const S *__restrict__ v,
                                               const S *__restrict__ w, T *__restrict__ om) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  T o = T(0);
  if (k > 0 && k < g.nz)
    o = T(w[A.at(i, j, k)]) -
        (T(1) - g.zeta[k]) * (T(0.5) * (mc_at<T, S>(g, A, u, v, i, j, k - 1) + mc_at<T, S>(g, A, u, v, i, j, k)));
  om[A.at(i, j, k)] = o;
}

// ---- pressure-gradient operators ------------------------------------------

template <typename T, typename S>
__device__ __forceinline__ T gz(const Geo<T> &g, const Acc<T> &A, const S *__restrict__ pi, int i, int j, int k) {
  if (k <= 0 || k >= g.nz) return T(0);
  return (T(pi[A.at(i, j, k)]) - T(pi[A.at(i, j, k - 1)])) / g.dzw[k];
}
template <typename T, typename S>
__device__ __forceinline__ T gzc(const Geo<T> &g, const Acc<T> &A, const S *__restrict__ pi, int i, int j, int k) {
  return T(0.5) * (gz<T, S>(g, A, pi, i, j, k) + gz<T, S>(g, A, pi, i, j, k + 1));
}
template <typename T, typename S>
__device__ __forceinline__ T px_op(const Geo<T> &g, const Acc<T> &A, const S *__restrict__ pi, int i, int j, int k) {
  const T dpx = (T(pi[A.at(i, j, k)]) - T(pi[A.at(i - 1, j, k)])) / g.dx;
  const T met = (T(1) - g.zc[k]) * g.sx[A.h(i, j)] / g.Du[A.h(i, j)];
  return dpx - met * (T(0.5) * (gzc<T, S>(g, A, pi, i - 1, j, k) + gzc<T, S>(g, A, pi, i, j, k)));
}
template <typename T, typename S>
__device__ __forceinline__ T py_op(const Geo<T> &g, const Acc<T> &A, const S *__restrict__ pi, int i, int j, int k) {
  const T dpy = (T(pi[A.at(i, j, k)]) - T(pi[A.at(i, j - 1, k)])) / g.dy;
  const T met = (T(1) - g.zc[k]) * g.sy[A.h(i, j)] / g.Dv[A.h(i, j)];
  return dpy - met * (T(0.5) * (gzc<T, S>(g, A, pi, i, j - 1, k) + gzc<T, S>(g, A, pi, i, j, k)));
}

// ---- stage kernels: R and s = s0 + h R --------------------------------------
// S: state storage (double), O: output type of the acoustic fields.

template <typename T, typename O>
__global__ void __launch_bounds__(256) k_stage_u(Geo<T> g, const double *__restrict__ u, const double *__restrict__ v,
                                                 const double *__restrict__ th, const double *__restrict__ pi,
                                                 const T *__restrict__ om, const double *__restrict__ u0, double h,
                                                 O *__restrict__ uo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const int nz = g.nz;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  auto U_ = [&](int ii, int jj, int kk) { return T(u[A.at(ii, jj, clampk(kk, nz - 1))]); };
  T fxc[2], uxc[2], fyc[2], vyc[2], fzc[2], ozc[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = i - 1 + q;
    const T Ux = T(0.5) * (g.Du[A.h(c, j)] * T(u[A.at(c, j, k)]) + g.Du[A.h(c + 1, j)] * T(u[A.at(c + 1, j, k)]));
    uxc[q] = Ux;
    fxc[q] = up5<T>(Ux, U_(c - 2, j, k), U_(c - 1, j, k), U_(c, j, k), U_(c + 1, j, k), U_(c + 2, j, k), U_(c + 3, j, k));
    const int jf = j + q;
    const T Vy = T(0.5) * (g.Dv[A.h(i - 1, jf)] * T(v[A.at(i - 1, jf, k)]) + g.Dv[A.h(i, jf)] * T(v[A.at(i, jf, k)]));
    vyc[q] = Vy;
    fyc[q] = up5<T>(Vy, U_(i, jf - 3, k), U_(i, jf - 2, k), U_(i, jf - 1, k), U_(i, jf, k), U_(i, jf + 1, k),
                    U_(i, jf + 2, k));
    const int kf = k + q;
    T O_ = T(0), F = T(0);
    if (kf > 0 && kf < nz) {
      O_ = T(0.5) * (om[A.at(i - 1, j, kf)] + om[A.at(i, j, kf)]);
      F = up5<T>(O_, U_(i, j, kf - 3), U_(i, j, kf - 2), U_(i, j, kf - 1), U_(i, j, kf), U_(i, j, kf + 1),
                 U_(i, j, kf + 2));
    }
    ozc[q] = O_;
    fzc[q] = F;
  }
  const size_t c0 = A.at(i, j, k);
  const T uc = T(u[c0]);
  const T adv = -(fxc[1] - fxc[0]) / g.dx - (fyc[1] - fyc[0]) / g.dy - (fzc[1] - fzc[0]) / g.dzeta[k] +
                uc * ((uxc[1] - uxc[0]) / g.dx + (vyc[1] - vyc[0]) / g.dy + (ozc[1] - ozc[0]) / g.dzeta[k]);
  const T thu = T(0.5) * (T(th[A.at(i - 1, j, k)]) + T(th[c0]));
  const T R = adv / g.Du[A.h(i, j)] - T(kCp) * thu * px_op<T, double>(g, A, pi, i, j, k);
  uo[c0] = static_cast<O>(u0[c0] + h * static_cast<double>(R));
}

template <typename T, typename O>
__global__ void __launch_bounds__(256) k_stage_v(Geo<T> g, const double *__restrict__ u, const double *__restrict__ v,
                                                 const double *__restrict__ th, const double *__restrict__ pi,
                                                 const T *__restrict__ om, const double *__restrict__ v0, double h,
                                                 O *__restrict__ vo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const int nz = g.nz;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  auto V_ = [&](int ii, int jj, int kk) { return T(v[A.at(ii, jj, clampk(kk, nz - 1))]); };
  T fxc[2], uxc[2], fyc[2], vyc[2], fzc[2], ozc[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int ifc = i + q;
    const T Ux = T(0.5) * (g.Du[A.h(ifc, j - 1)] * T(u[A.at(ifc, j - 1, k)]) + g.Du[A.h(ifc, j)] * T(u[A.at(ifc, j, k)]));
    uxc[q] = Ux;
    fxc[q] = up5<T>(Ux, V_(ifc - 3, j, k), V_(ifc - 2, j, k), V_(ifc - 1, j, k), V_(ifc, j, k), V_(ifc + 1, j, k),
                    V_(ifc + 2, j, k));
    const int c = j - 1 + q;
    const T Vy = T(0.5) * (g.Dv[A.h(i, c)] * T(v[A.at(i, c, k)]) + g.Dv[A.h(i, c + 1)] * T(v[A.at(i, c + 1, k)]));
    vyc[q] = Vy;
    fyc[q] = up5<T>(Vy, V_(i, c - 2, k), V_(i, c - 1, k), V_(i, c, k), V_(i, c + 1, k), V_(i, c + 2, k), V_(i, c + 3, k));
    const int kf = k + q;
    T O_ = T(0), F = T(0);
    if (kf > 0 && kf < nz) {
      O_ = T(0.5) * (om[A.at(i, j - 1, kf)] + om[A.at(i, j, kf)]);
      F = up5<T>(O_, V_(i, j, kf - 3), V_(i, j, kf - 2), V_(i, j, kf - 1), V_(i, j, kf), V_(i, j, kf + 1),
                 V_(i, j, kf + 2));
    }
    ozc[q] = O_;
    fzc[q] = F;
  }
  const size_t c0 = A.at(i, j, k);
  const T vc = T(v[c0]);
  const T adv = -(fxc[1] - fxc[0]) / g.dx - (fyc[1] - fyc[0]) / g.dy - (fzc[1] - fzc[0]) / g.dzeta[k] +
                vc * ((uxc[1] - uxc[0]) / g.dx + (vyc[1] - vyc[0]) / g.dy + (ozc[1] - ozc[0]) / g.dzeta[k]);
  const T thv = T(0.5) * (T(th[A.at(i, j - 1, k)]) + T(th[c0]));
  const T R = adv / g.Dv[A.h(i, j)] - T(kCp) * thv * py_op<T, double>(g, A, pi, i, j, k);
  vo[c0] = static_cast<O>(v0[c0] + h * static_cast<double>(R));
}

template <typename T, typename O>
__global__ void __launch_bounds__(256) k_stage_w(Geo<T> g, const double *__restrict__ u, const double *__restrict__ v,
                                                 const double *__restrict__ w, const double *__restrict__ th,
                                                 const double *__restrict__ pi, const T *__restrict__ om,
                                                 const double *__restrict__ w0, double h, O *__restrict__ wo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;  // k = 0..nz
  if (i >= g.nx) return;
  const int nz = g.nz;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  const size_t c0 = A.at(i, j, k);
  if (k == 0 || k == nz) {
    wo[c0] = static_cast<O>(w0[c0]);
    return;
  }
  auto W_ = [&](int ii, int jj, int kk) { return T(w[A.at(ii, jj, clampk(kk, nz))]); };
  T fxc[2], uxc[2], fyc[2], vyc[2], fzc[2], ozc[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int ifc = i + q;
    const T Ux = T(0.5) * (g.Du[A.h(ifc, j)] * T(u[A.at(ifc, j, k - 1)]) + g.Du[A.h(ifc, j)] * T(u[A.at(ifc, j, k)]));
    uxc[q] = Ux;
    fxc[q] = up5<T>(Ux, W_(ifc - 3, j, k), W_(ifc - 2, j, k), W_(ifc - 1, j, k), W_(ifc, j, k), W_(ifc + 1, j, k),
                    W_(ifc + 2, j, k));
    const int jf = j + q;
    const T Vy = T(0.5) * (g.Dv[A.h(i, jf)] * T(v[A.at(i, jf, k - 1)]) + g.Dv[A.h(i, jf)] * T(v[A.at(i, jf, k)]));
    vyc[q] = Vy;
    fyc[q] = up5<T>(Vy, W_(i, jf - 3, k), W_(i, jf - 2, k), W_(i, jf - 1, k), W_(i, jf, k), W_(i, jf + 1, k),
                    W_(i, jf + 2, k));
    const int c = k - 1 + q;
    const T Oc = T(0.5) * (om[A.at(i, j, c)] + om[A.at(i, j, c + 1)]);
    ozc[q] = Oc;
    fzc[q] = up5<T>(Oc, W_(i, j, c - 2), W_(i, j, c - 1), W_(i, j, c), W_(i, j, c + 1), W_(i, j, c + 2), W_(i, j, c + 3));
  }
  const T wc = T(w[c0]);
  const T adv = -(fxc[1] - fxc[0]) / g.dx - (fyc[1] - fyc[0]) / g.dy - (fzc[1] - fzc[0]) / g.dzw[k] +
                wc * ((uxc[1] - uxc[0]) / g.dx + (vyc[1] - vyc[0]) / g.dy + (ozc[1] - ozc[0]) / g.dzw[k]);
  const T D = g.Dc[A.h(i, j)];
  const size_t cb = A.at(i, j, k - 1);
  const T thw = T(0.5) * (T(th[cb]) + T(th[c0]));
  const T th0w = T(0.5) * (g.theta0[cb] + g.theta0[c0]);
  const T pz = ((T(pi[c0]) - T(pi[cb])) / g.dzw[k]) / D;
  const T R = adv / D - T(kCp) * thw * pz + T(kG) * thw / th0w;
  wo[c0] = static_cast<O>(w0[c0] + h * static_cast<double>(R));
}

template <typename T>
__device__ __forceinline__ T gam_at(const Geo<T> &g, const Acc<T> &A, int i, int j, int k) {
  const int kk = k < 1 ? 1 : (k > g.nz - 1 ? g.nz - 1 : k);
  return (g.theta0[A.at(i, j, kk)] - g.theta0[A.at(i, j, kk - 1)]) / (g.Dc[A.h(i, j)] * g.dzw[kk]);
}

template <typename T, typename O>
__global__ void __launch_bounds__(256) k_stage_cell(Geo<T> g, const double *__restrict__ u, const double *__restrict__ v,
                                                    const double *__restrict__ w, const double *__restrict__ rho,
                                                    const double *__restrict__ th, const double *__restrict__ pi,
                                                    const T *__restrict__ om, const double *__restrict__ rho0s,
                                                    const double *__restrict__ th0s, const double *__restrict__ pi0s,
                                                    double h, O *__restrict__ rho_o, double *__restrict__ th_o,
                                                    O *__restrict__ pi_o, int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const int nz = g.nz;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  auto T_ = [&](int ii, int jj, int kk) { return T(th[A.at(ii, jj, clampk(kk, nz - 1))]); };
  const T D = g.Dc[A.h(i, j)];
  T frx[2], fry[2], frz[2], ftx[2], fty[2], ftz[2], mfx[2], mfy[2], mfz[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int f = i + q;
    const T Ux = T(u[A.at(f, j, k)]);
    mfx[q] = g.Du[A.h(f, j)] * Ux;
    frx[q] = g.Du[A.h(f, j)] * (T(0.5) * (T(rho[A.at(f - 1, j, k)]) + T(rho[A.at(f, j, k)]))) * Ux;
    ftx[q] = g.Du[A.h(f, j)] * llf<T>(Ux, weno_left<T>(T_(f - 3, j, k), T_(f - 2, j, k), T_(f - 1, j, k), T_(f, j, k), T_(f + 1, j, k)),
                                      weno_right<T>(T_(f - 2, j, k), T_(f - 1, j, k), T_(f, j, k), T_(f + 1, j, k), T_(f + 2, j, k)));
    const int gg = j + q;
    const T Vy = T(v[A.at(i, gg, k)]);
    mfy[q] = g.Dv[A.h(i, gg)] * Vy;
    fry[q] = g.Dv[A.h(i, gg)] * (T(0.5) * (T(rho[A.at(i, gg - 1, k)]) + T(rho[A.at(i, gg, k)]))) * Vy;
    fty[q] = g.Dv[A.h(i, gg)] * llf<T>(Vy, weno_left<T>(T_(i, gg - 3, k), T_(i, gg - 2, k), T_(i, gg - 1, k), T_(i, gg, k), T_(i, gg + 1, k)),
                                       weno_right<T>(T_(i, gg - 2, k), T_(i, gg - 1, k), T_(i, gg, k), T_(i, gg + 1, k), T_(i, gg + 2, k)));
    const int hh = k + q;
    T O_ = T(0), FR = T(0), FT = T(0);
    if (hh > 0 && hh < nz) {
      O_ = om[A.at(i, j, hh)];
      FR = (T(0.5) * (T(rho[A.at(i, j, hh - 1)]) + T(rho[A.at(i, j, hh)]))) * O_;
      FT = llf<T>(O_, weno_left<T>(T_(i, j, hh - 3), T_(i, j, hh - 2), T_(i, j, hh - 1), T_(i, j, hh), T_(i, j, hh + 1)),
                  weno_right<T>(T_(i, j, hh - 2), T_(i, j, hh - 1), T_(i, j, hh), T_(i, j, hh + 1), T_(i, j, hh + 2)));
    }
    mfz[q] = O_;
    frz[q] = FR;
    ftz[q] = FT;
  }
  const size_t c = A.at(i, j, k);
  const T rr = -((frx[1] - frx[0]) / g.dx + (fry[1] - fry[0]) / g.dy + (frz[1] - frz[0]) / g.dzeta[k]) / D;
  const T div = (mfx[1] - mfx[0]) / g.dx + (mfy[1] - mfy[0]) / g.dy + (mfz[1] - mfz[0]) / g.dzeta[k];
  const T fd = (ftx[1] - ftx[0]) / g.dx + (fty[1] - fty[0]) / g.dy + (ftz[1] - ftz[0]) / g.dzeta[k];
  const T rt = -(fd - T(th[c]) * div) / D -
               T(0.5) * (T(w[c]) * gam_at<T>(g, A, i, j, k) + T(w[A.at(i, j, k + 1)]) * gam_at<T>(g, A, i, j, k + 1));
  const T rp = g.kap[c] * rr + g.lam[c] * rt;
  const double rn = rho0s[c] + h * static_cast<double>(rr);
  const double tn = th0s[c] + h * static_cast<double>(rt);
  const double pn = pi0s[c] + h * static_cast<double>(rp);
  rho_o[c] = static_cast<O>(rn);
  th_o[c] = tn;
  pi_o[c] = static_cast<O>(pn);
  if (!isfinite(tn)) atomicOr(flag, 1);
}

// ---- acoustic substep -------------------------------------------------------

template <typename T, typename I, typename O>
__global__ void __launch_bounds__(256) k_sub_h(Geo<T> g, const I *__restrict__ pi, const I *__restrict__ u,
                                               const I *__restrict__ v, T dtau, O *__restrict__ uo,
                                               O *__restrict__ vo, int check, int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  const size_t c = A.at(i, j, k);
  const T th0u = T(0.5) * (g.theta0[A.at(i - 1, j, k)] + g.theta0[c]);
  const T th0v = T(0.5) * (g.theta0[A.at(i, j - 1, k)] + g.theta0[c]);
  const T un = T(u[c]) - dtau * (T(kCp) * th0u * px_op<T, I>(g, A, pi, i, j, k));
  const T vn = T(v[c]) - dtau * (T(kCp) * th0v * py_op<T, I>(g, A, pi, i, j, k));
  uo[c] = static_cast<O>(un);
  vo[c] = static_cast<O>(vn);
  if (check && !(isfinite(un) && isfinite(vn))) atomicOr(flag, 1);
}

// Per-column HEVI solve.  scr holds 5 work arrays of nz*nx*ny values laid
// out [array][k][column] so that each k-step of a wavefront is coalesced.
template <typename T, typename I, typename O>
__global__ void __launch_bounds__(256) k_sub_v(Geo<T> g, const O *__restrict__ un, const O *__restrict__ vn,
                                               const I *__restrict__ w, const I *__restrict__ rho,
                                               const I *__restrict__ pi, T dtau, T an, T ao, O *__restrict__ wo,
                                               O *__restrict__ rho_o, O *__restrict__ pi_o, T *__restrict__ scr,
                                               int check, int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= g.nx) return;
  const int nz = g.nz;
  const Acc<T> A{g.nx, g.ny, static_cast<size_t>(g.nx) * g.ny};
  const size_t nxy = A.nxy, ncell = nxy * nz;
  const size_t col = A.h(i, j);
  T *MCC = scr, *HD = scr + ncell, *PH = scr + 2 * ncell, *CP = scr + 3 * ncell, *DP = scr + 4 * ncell;
  const T D = g.Dc[col];
  const T an2 = an * an;
  const T dtD = dtau / D;
  auto ck_at = [&](int k) { return dtau * g.kap[A.at(i, j, k)] / (D * g.dzeta[k]); };
  auto r0w_at = [&](int k) { return T(0.5) * (g.rho0[A.at(i, j, k - 1)] + g.rho0[A.at(i, j, k)]); };
  auto mw_at = [&](int k) {
    return (T(1) - g.zeta[k]) * (T(0.5) * (MCC[static_cast<size_t>(k - 1) * nxy + col] + MCC[static_cast<size_t>(k) * nxy + col]));
  };
  auto gk_at = [&](int k) {
    const T th0w = T(0.5) * (g.theta0[A.at(i, j, k - 1)] + g.theta0[A.at(i, j, k)]);
    return dtau * T(kCp) * th0w / (D * g.dzw[k]);
  };
  auto fh_at = [&](int k) {
    if (k <= 0 || k >= nz) return T(0);
    return r0w_at(k) * (ao * T(w[A.at(i, j, k)]) - mw_at(k));
  };
  // pass 1: terrain slope flux and horizontal rho0 mass-flux divergence
  for (int k = 0; k < nz; ++k) {
    const size_t cc = A.at(i, j, k);
    const T mc = T(0.5) * (T(un[cc]) * g.sx[col] + T(un[A.at(i + 1, j, k)]) * g.sx[A.h(i + 1, j)]) +
                 T(0.5) * (T(vn[cc]) * g.sy[col] + T(vn[A.at(i, j + 1, k)]) * g.sy[A.h(i, j + 1)]);
    const T r0e = T(0.5) * (g.rho0[cc] + g.rho0[A.at(i + 1, j, k)]);
    const T r0wst = T(0.5) * (g.rho0[A.at(i - 1, j, k)] + g.rho0[cc]);
    const T r0n = T(0.5) * (g.rho0[cc] + g.rho0[A.at(i, j + 1, k)]);
    const T r0s = T(0.5) * (g.rho0[A.at(i, j - 1, k)] + g.rho0[cc]);
    const T hd = (g.Du[A.h(i + 1, j)] * r0e * T(un[A.at(i + 1, j, k)]) - g.Du[col] * r0wst * T(un[cc])) / g.dx +
                 (g.Dv[A.h(i, j + 1)] * r0n * T(vn[A.at(i, j + 1, k)]) - g.Dv[col] * r0s * T(vn[cc])) / g.dy;
    MCC[static_cast<size_t>(k) * nxy + col] = mc;
    HD[static_cast<size_t>(k) * nxy + col] = hd;
  }
  // pass 2: explicit pi' estimate P^ (vertical fluxes with the old w)
  {
    T fh_lo = T(0);
    for (int k = 0; k < nz; ++k) {
      const T fh_hi = fh_at(k + 1);
      const size_t cc = A.at(i, j, k);
      PH[static_cast<size_t>(k) * nxy + col] =
          T(pi[cc]) - g.kap[cc] * dtD * HD[static_cast<size_t>(k) * nxy + col] - ck_at(k) * (fh_hi - fh_lo);
      fh_lo = fh_hi;
    }
  }
  // pass 3: Thomas forward sweep over interfaces 1..nz-1 (row m = k-1)
  const int n = nz - 1;
  T cprev = T(0), dprev = T(0);
  for (int m = 0; m < n; ++m) {
    const int k = m + 1;
    const T ga = gk_at(k) * an2;
    const T ckk = ck_at(k), ckm = ck_at(k - 1);
    const T am = (m > 0) ? -ga * (ckm * r0w_at(k - 1)) : T(0);
    const T bm = T(1) + ga * ((ckk + ckm) * r0w_at(k));
    const T cm = (m < n - 1) ? -ga * (ckk * r0w_at(k + 1)) : T(0);
    const T gkk = gk_at(k);
    const T td = T(w[A.at(i, j, k)]) - gkk * ao * (T(pi[A.at(i, j, k)]) - T(pi[A.at(i, j, k - 1)])) -
                 gkk * an * (PH[static_cast<size_t>(k) * nxy + col] - PH[static_cast<size_t>(k - 1) * nxy + col]);
    T denom, cp, dp;
    if (m == 0) {
      denom = bm;
      cp = cm / denom;
      dp = td / denom;
    } else {
      denom = bm - am * cprev;
      cp = cm / denom;
      dp = (td - am * dprev) / denom;
    }
    CP[static_cast<size_t>(m) * nxy + col] = cp;
    DP[static_cast<size_t>(m) * nxy + col] = dp;
    cprev = cp;
    dprev = dp;
  }
  // pass 4: back substitution fused with the rho' / pi' update, top down
  int bad = 0;
  T wup = T(0);                 // w_new at the interface above (k+1)
  T fup = T(0);                 // vertical flux F at interface k+1
  const size_t ctop = A.at(i, j, nz);
  wo[ctop] = static_cast<O>(T(w[ctop]));
  for (int k = nz - 1; k >= 0; --k) {
    T fk = T(0);
    if (k >= 1) {
      const int m = k - 1;
      T wk = DP[static_cast<size_t>(m) * nxy + col];
      if (m < n - 1) wk = wk - CP[static_cast<size_t>(m) * nxy + col] * wup;
      fk = r0w_at(k) * (an * wk + ao * T(w[A.at(i, j, k)]) - mw_at(k));
      wo[A.at(i, j, k)] = static_cast<O>(wk);
      bad |= !isfinite(wk);
      wup = wk;
    } else {
      wo[A.at(i, j, 0)] = static_cast<O>(MCC[col]);
    }
    const size_t cc = A.at(i, j, k);
    const T drho = -dtD * (HD[static_cast<size_t>(k) * nxy + col] + (fup - fk) / g.dzeta[k]);
    const T rn = T(rho[cc]) + drho;
    const T pn = T(pi[cc]) + g.kap[cc] * drho;
    rho_o[cc] = static_cast<O>(rn);
    pi_o[cc] = static_cast<O>(pn);
    bad |= !(isfinite(rn) && isfinite(pn));
    fup = fk;
  }
  if (check && bad) atomicOr(flag, 1);
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
  std::vector<double> h_sponge;
  int kmin_sponge = 0;
  FieldSet S[3];  // [0] = s0 of the step; [1], [2] stage outputs
  int cur0 = 0;   // index of s^n
  AcousticSet<double> A64[2];
  AcousticSet<float> A32[2];
  void *om = nullptr;   // Omega (T)
  void *scr = nullptr;  // 5 column work arrays (T)
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

template <typename T>
int build_geo(vibe_mi_domain *d, const vibe_mi_domain_desc *ds, Geo<T> &g) {
  const int nx = ds->nx, ny = ds->ny, nz = ds->nz;
  const size_t nxy = d->nxy, nc = d->ncell;
  std::vector<double> zeta(ds->zeta, ds->zeta + nz + 1), dzeta(nz), zc(nz), dzw(nz + 1, 0.0);
  for (int k = 0; k < nz; ++k) {
    dzeta[k] = zeta[k + 1] - zeta[k];
    zc[k] = 0.5 * (zeta[k] + zeta[k + 1]);
  }
  for (int k = 1; k < nz; ++k) dzw[k] = zc[k] - zc[k - 1];
  std::vector<double> Dc(nxy), Du(nxy), Dv(nxy), sx(nxy), sy(nxy);
  auto H = [&](int i, int j) { return static_cast<size_t>((j + ny) % ny) * nx + (i + nx) % nx; };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double z = ds->zs[H(i, j)], zw = ds->zs[H(i - 1, j)], zsth = ds->zs[H(i, j - 1)];
      Dc[H(i, j)] = ds->ztop - z;
      Du[H(i, j)] = ds->ztop - 0.5 * (zw + z);
      Dv[H(i, j)] = ds->ztop - 0.5 * (zsth + z);
      sx[H(i, j)] = (z - zw) / ds->dx;
      sy[H(i, j)] = (z - zsth) / ds->dy;
    }
  std::vector<double> kap(nc), lam(nc);
  for (size_t c = 0; c < nc; ++c) {
    kap[c] = (kRd / kCv) * ds->pi0[c] / ds->rho0[c];
    lam[c] = (kRd / kCv) * ds->pi0[c] / ds->theta0[c];
  }
  auto cvt = [](const std::vector<double> &a) { return std::vector<T>(a.begin(), a.end()); };
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.dx = static_cast<T>(ds->dx);
  g.dy = static_cast<T>(ds->dy);
  T *p;
  if (int rc = upload(d, &p, cvt(zeta))) return rc;
  g.zeta = p;
  if (int rc = upload(d, &p, cvt(dzeta))) return rc;
  g.dzeta = p;
  if (int rc = upload(d, &p, cvt(zc))) return rc;
  g.zc = p;
  if (int rc = upload(d, &p, cvt(dzw))) return rc;
  g.dzw = p;
  if (int rc = upload(d, &p, cvt(Dc))) return rc;
  g.Dc = p;
  if (int rc = upload(d, &p, cvt(Du))) return rc;
  g.Du = p;
  if (int rc = upload(d, &p, cvt(Dv))) return rc;
  g.Dv = p;
  if (int rc = upload(d, &p, cvt(sx))) return rc;
  g.sx = p;
  if (int rc = upload(d, &p, cvt(sy))) return rc;
  g.sy = p;
  if (int rc = upload(d, &p, std::vector<T>(ds->rho0, ds->rho0 + nc))) return rc;
  g.rho0 = p;
  if (int rc = upload(d, &p, std::vector<T>(ds->theta0, ds->theta0 + nc))) return rc;
  g.theta0 = p;
  if (int rc = upload(d, &p, cvt(kap))) return rc;
  g.kap = p;
  if (int rc = upload(d, &p, cvt(lam))) return rc;
  g.lam = p;
  return 0;
}

inline dim3 grid_of(const vibe_mi_domain *d, int nk) {
  return dim3(ceil_div(d->nx, 64), d->ny, nk);
}
constexpr unsigned kBlk = 64;

int timed_begin(vibe_mi_domain *d, bool timed, int slot) {
  if (timed) VMI_HIP(hipEventRecord(d->ev[slot], d->stream));
  return 0;
}

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
  const T *om = static_cast<const T *>(d->om);
  hipLaunchKernelGGL((k_stage_u<T, T>), grid_of(d, d->nz), dim3(kBlk), 0, d->stream, g, cur.u, cur.v, cur.th, cur.pi,
                     om, s0.u, h, a.u);
  hipLaunchKernelGGL((k_stage_v<T, T>), grid_of(d, d->nz), dim3(kBlk), 0, d->stream, g, cur.u, cur.v, cur.th, cur.pi,
                     om, s0.v, h, a.v);
  hipLaunchKernelGGL((k_stage_w<T, T>), grid_of(d, d->nz + 1), dim3(kBlk), 0, d->stream, g, cur.u, cur.v, cur.w,
                     cur.th, cur.pi, om, s0.w, h, a.w);
  hipLaunchKernelGGL((k_stage_cell<T, T>), grid_of(d, d->nz), dim3(kBlk), 0, d->stream, g, cur.u, cur.v, cur.w,
                     cur.rho, cur.th, cur.pi, om, s0.rho, s0.th, s0.pi, h, a.rho, nxt.th, a.pi, d->flag);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T, typename I, typename O>
int run_substep(vibe_mi_domain *d, const Geo<T> &g, const I *u, const I *v, const I *w, const I *rho, const I *pi,
                O *uo, O *vo, O *wo, O *rho_o, O *pi_o, double dtau, int check) {
  const T an = static_cast<T>(0.5 * (1.0 + d->eps_s)), ao = static_cast<T>(0.5 * (1.0 - d->eps_s));
  hipLaunchKernelGGL((k_sub_h<T, I, O>), grid_of(d, d->nz), dim3(kBlk), 0, d->stream, g, pi, u, v,
                     static_cast<T>(dtau), uo, vo, check, d->flag);
  hipLaunchKernelGGL((k_sub_v<T, I, O>), dim3(ceil_div(d->nx, 64), d->ny, 1), dim3(kBlk), 0, d->stream, g, uo, vo, w,
                     rho, pi, static_cast<T>(dtau), an, ao, wo, rho_o, pi_o, static_cast<T *>(d->scr), check, d->flag);
  VMI_HIP(hipGetLastError());
  return 0;
}

// One full RK3 step.  Stage outputs go to S[(cur0+1)%3] / S[(cur0+2)%3].
template <typename T>
int step_impl(vibe_mi_domain *d, const Geo<T> &g, AcousticSet<T> *A, double dt, int nsub, bool timed) {
  const int ns = nsub < 1 ? 1 : nsub;
  const double hk[3] = {dt / 3.0, dt / 2.0, dt};
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
        if (int rc = run_substep<T, T, T>(d, g, in.u, in.v, in.w, in.rho, in.pi, out.u, out.v, out.w, out.rho, out.pi,
                                          dtau, 0))
          return rc;
        p = 1 - p;
      } else {
        if (int rc = run_substep<T, T, double>(d, g, in.u, in.v, in.w, in.rho, in.pi, nxt.u, nxt.v, nxt.w, nxt.rho,
                                               nxt.pi, dtau, 1))
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
  if (int rc = build_geo<double>(d, ds, d->g64)) return rc;
  if (d->precision == VIBE_MI_MIXED)
    if (int rc = build_geo<float>(d, ds, d->g32)) return rc;
  const size_t nc = d->ncell;
  if (int rc = upload(d, &d->zs, std::vector<double>(ds->zs, ds->zs + d->nxy))) return rc;
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
  if (int rc = dalloc(d, &p, 5 * nc * tsz)) return rc;
  d->scr = p;
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
  if (int rc = run_substep<double, double, double>(d, d->g64, s.u, s.v, s.w, s.rho, s.pi, o.u, o.v, o.w, o.rho, o.pi,
                                                   dtau, 1))
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
                     d->dy, d->ztop, d->zs, d->g64.dzeta, d->g64.zc, d->rho0, d->theta0, d->pi0, f.u, f.v, f.w, f.rho,
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
                     d->zs, d->g64.dzeta, f.u, f.v, f.w, dt, d->dmax);
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
