// semi_implicit.hip -- the reference's semi-implicit integrator step on
// gfx950 (C ABI in include/vibe_mi/semi_implicit.h).
//
// The step keeps the reference's order and arithmetic on its literal
// equations: the slow tendencies of literal.hip, the right-hand side of the
// Crank-Nicolson pressure equation (one lane per cell, the reference's
// expression order, -ffp-contract=off), the device Helmholtz solver of
// helmholtz.hip, the forward-Euler update of the scalars, the optional
// Robert-Asselin-Williams filter and the sponge.  Every kernel is one pass
// over HBM with one lane per point; the Helmholtz solve dominates the step.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <vector>

#include "common.hpp"
#include "helmholtz_dev.hpp"
#include "literal_dev.hpp"
#include "vibe_mi/semi_implicit.h"

namespace vibe_mi {
namespace {

struct SiGrid {
  int nx, ny, nz;
  double dx, dy, ztop;
  const double *zeta, *zs, *r0, *t0, *p0;
};

__device__ inline size_t cidx(const SiGrid &g, int i, int j, int k) {
  return (static_cast<size_t>(k) * g.ny + j) * g.nx + i;
}
__device__ inline size_t uidx(const SiGrid &g, int i, int j, int k) {
  return (static_cast<size_t>(k) * g.ny + j) * (g.nx + 1) + i;
}
__device__ inline size_t vidx(const SiGrid &g, int i, int j, int k) {
  return (static_cast<size_t>(k) * (g.ny + 1) + j) * g.nx + i;
}
// std::max / std::min of the reference: (a < b) ? b : a, (b < a) ? b : a
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ inline int imin(int a, int b) { return (b < a) ? b : a; }
__device__ inline int imax(int a, int b) { return (a < b) ? b : a; }

// dz = max(dzeta(k) * (z_top - terrain(i, j)), 1)
__device__ inline double dz_at(const SiGrid &g, int i, int j, int k) {
  const double zs = g.zs == nullptr ? 0.0 : g.zs[static_cast<size_t>(j) * g.nx + i];
  const double scale = g.ztop - zs;
  return dmax((g.zeta[k + 1] - g.zeta[k]) * scale, 1.0);
}

// div(rho0 u) of one cell (the flux_div lambda of assemble_rhs,
// semi_implicit.cpp:77-87): face rho0 the mean of the two cells with the
// outer neighbour clamped
__device__ inline double flux_div(const SiGrid &g, const double *u, const double *v, const double *w, int i, int j,
                                  int k, double dz) {
  const double *r = g.r0;
  const double fx = u[uidx(g, i + 1, j, k)] * 0.5 * (r[cidx(g, i, j, k)] + r[cidx(g, imin(i + 1, g.nx - 1), j, k)]) -
                    u[uidx(g, i, j, k)] * 0.5 * (r[cidx(g, imax(i - 1, 0), j, k)] + r[cidx(g, i, j, k)]);
  const double fy = v[vidx(g, i, j + 1, k)] * 0.5 * (r[cidx(g, i, j, k)] + r[cidx(g, i, imin(j + 1, g.ny - 1), k)]) -
                    v[vidx(g, i, j, k)] * 0.5 * (r[cidx(g, i, imax(j - 1, 0), k)] + r[cidx(g, i, j, k)]);
  const double fz = w[cidx(g, i, j, k + 1)] * 0.5 * (r[cidx(g, i, j, k)] + r[cidx(g, i, j, imin(k + 1, g.nz - 1))]) -
                    w[cidx(g, i, j, k)] * 0.5 * (r[cidx(g, i, j, imax(k - 1, 0))] + r[cidx(g, i, j, k)]);
  return fx / g.dx + fy / g.dy + fz / dz;
}

// assemble_rhs (semi_implicit.cpp:58-108): a, b = 1 and the right-hand side
// per cell
__global__ void __launch_bounds__(64) k_si_assemble(SiGrid g, const double *__restrict__ su,
                                                    const double *__restrict__ sv, const double *__restrict__ sw,
                                                    const double *__restrict__ pi, const double *__restrict__ nu,
                                                    const double *__restrict__ nv, const double *__restrict__ nw,
                                                    double dt, double *__restrict__ acoef, double *__restrict__ bcoef,
                                                    double *__restrict__ rhs) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const size_t c = cidx(g, i, j, k);
  const double half_dt = 0.5 * dt;
  const double a0 = half_dt * half_dt * (kRd / kCv) * g.p0[c] * kCp * g.t0[c];
  acoef[c] = a0;
  bcoef[c] = 1.0;
  const double dz = dz_at(g, i, j, k);
  const double div_un = flux_div(g, su, sv, sw, i, j, k, dz);
  const double div_nn = flux_div(g, nu, nv, nw, i, j, k, dz);
  const double G = div_un + half_dt * div_nn;
  const double cc = (kRd / kCv) * (g.p0[c] / dmax(g.r0[c], 1e-8));
  const double p = pi[c];
  const double div_a_grad =
      (a0 * (pi[cidx(g, imin(i + 1, g.nx - 1), j, k)] - p) - a0 * (p - pi[cidx(g, imax(i - 1, 0), j, k)])) /
          (g.dx * g.dx) +
      (a0 * (pi[cidx(g, i, imin(j + 1, g.ny - 1), k)] - p) - a0 * (p - pi[cidx(g, i, imax(j - 1, 0), k)])) /
          (g.dy * g.dy);
  rhs[c] = p - dt * cc * G + div_a_grad;
}

// t += rate over n points (Field::add_scaled(1, rate): 1 * rate is exact)
__global__ void __launch_bounds__(256) k_si_add(size_t n, const double *__restrict__ rate, double *__restrict__ t) {
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256)
    t[q] += rate[q];
}

struct Euler {
  double *s[8];
  const double *t[8];
  int count;
};
// theta, qv, rho, qc..qg += dt N (semi_implicit.cpp:199-205)
__global__ void __launch_bounds__(256) k_si_euler(size_t n, Euler e, double dt) {
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n;
       q += static_cast<size_t>(gridDim.x) * 256)
    for (int f = 0; f < e.count; ++f) e.s[f][q] += dt * e.t[f][q];
}

// pi.axpy(1 - alpha, pi, alpha): pi = (1 - alpha) pi + alpha pi
// (semi_implicit.cpp:208-211, Field::axpy field.hpp:101-105)
__global__ void __launch_bounds__(256) k_si_filter(size_t n, double a1, double a2, double *__restrict__ pi) {
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256)
    pi[q] = a1 * pi[q] + a2 * pi[q];
}

// Damping::apply_sponge (damping.cpp:76-138) on the full state: u, v, w
// faces relaxed to zero, rho, theta, pi to the reference state, levels
// kmin..nz-1 (w faces kmin..nz, face nz with the rate of level nz-1)
__global__ void __launch_bounds__(64) k_si_sponge(SiGrid g, const double *__restrict__ rate, int kmin, double dt,
                                                  double *__restrict__ u, double *__restrict__ v,
                                                  double *__restrict__ w, double *__restrict__ rho,
                                                  double *__restrict__ theta, double *__restrict__ pi) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y, k = kmin + static_cast<int>(blockIdx.z);
  if (i > g.nx || j > g.ny || k > g.nz) return;
  const double a = rate[imin(k, g.nz - 1)];
  if (!(a > 0.0)) return;
  const double wgt = 1.0 / (1.0 + dt * a);
  if (k < g.nz && j < g.ny) u[uidx(g, i, j, k)] = wgt * u[uidx(g, i, j, k)];
  if (i == g.nx) return;
  if (k < g.nz) v[vidx(g, i, j, k)] = wgt * v[vidx(g, i, j, k)];
  if (j == g.ny) return;
  w[cidx(g, i, j, k)] = wgt * w[cidx(g, i, j, k)];
  if (k == g.nz) return;
  const size_t c = cidx(g, i, j, k);
  rho[c] = wgt * (rho[c] + dt * a * g.r0[c]);
  theta[c] = wgt * (theta[c] + dt * a * g.t0[c]);
  pi[c] = wgt * (pi[c] + dt * a * g.p0[c]);
}

// *flag = 1 if any of x[0..n) is not finite
__global__ void __launch_bounds__(256) k_si_nonfinite(size_t n, const double *__restrict__ x, int *__restrict__ flag) {
  bool bad = false;
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256)
    bad |= !isfinite(x[q]);
  if (bad) *flag = 1;
}
__global__ void k_si_clear(int *flag, unsigned long long *mx) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *flag = 0;
    *mx = 0ull;
  }
}

// Advection::cfl_number(s, 1) (advection.cpp:325-340): the largest
// |u|/dx + |v|/dy + |w|/dz of the cell-centred winds; non-negative doubles
// order like their bit patterns, so the maximum is an integer atomicMax
__global__ void __launch_bounds__(256) k_si_cfl(SiGrid g, const double *__restrict__ u, const double *__restrict__ v,
                                                const double *__restrict__ w, unsigned long long *__restrict__ mx) {
  const size_t ncell = static_cast<size_t>(g.nx) * g.ny * g.nz;
  double best = 0.0;
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < ncell;
       q += static_cast<size_t>(gridDim.x) * 256) {
    const int i = static_cast<int>(q % g.nx), j = static_cast<int>((q / g.nx) % g.ny),
              k = static_cast<int>(q / (static_cast<size_t>(g.nx) * g.ny));
    const double dz = dz_at(g, i, j, k);
    const double uu = 0.5 * (u[uidx(g, i, j, k)] + u[uidx(g, i + 1, j, k)]);
    const double vv = 0.5 * (v[vidx(g, i, j, k)] + v[vidx(g, i, j + 1, k)]);
    const double ww = 0.5 * (w[cidx(g, i, j, k)] + w[cidx(g, i, j, k + 1)]);
    const double val = 1.0 * (fabs(uu) / g.dx + fabs(vv) / g.dy + fabs(ww) / dz);
    best = dmax(best, val);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(h)) sh[threadIdx.x] = dmax(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(mx, static_cast<unsigned long long>(__double_as_longlong(sh[0])));
}

}  // namespace
}  // namespace vibe_mi

using namespace vibe_mi;

struct vibe_mi_si {
  vibe_mi_literal *lit = nullptr;
  vibe_mi_helmholtz *helm = nullptr;
  LiteralView v{};
  SiGrid g{};
  int solver = VIBE_MI_HELM_BICGSTAB, max_iter = 0, kmin = 0;
  double tol = 0.0, filter = 0.0;
  // dt whose coefficients the Helmholtz handle holds (a depends on dt alone:
  // pi0 and theta0 are fixed), so the solver is refactored only when dt changes
  double coef_dt = -1.0;
  bool first_step = true;
  std::vector<void *> allocs;
  double *acoef = nullptr, *bcoef = nullptr, *rhs = nullptr, *pinew = nullptr, *rate = nullptr;
  double *phys[4] = {};  // theta, qv, u, v rates (NULL: zero)
  int *flag = nullptr;
  unsigned long long *mx = nullptr;
  vibe_mi_si_stats stats{};
};

namespace {

template <typename P>
int si_alloc(vibe_mi_si *h, P **out, size_t count) {
  void *p = nullptr;
  VMI_HIP(hipMalloc(&p, (count > 0 ? count : 1) * sizeof(P)));
  h->allocs.push_back(p);
  *out = static_cast<P *>(p);
  return 0;
}

size_t ncell(const vibe_mi_si *h) { return h->v.n[3]; }
unsigned vgrid(size_t n) { return static_cast<unsigned>(std::min<size_t>(ceil_div(n, 256), 4096)); }

// N = tendencies(s) + physics rates
int tendencies(vibe_mi_si *h) {
  if (int rc = literal_tendencies_device(h->lit)) return rc;
  static const int target[4] = {4, 6, 0, 1};
  for (int q = 0; q < 4; ++q)
    if (h->phys[q] != nullptr) {
      const size_t n = h->v.n[target[q]];
      hipLaunchKernelGGL(k_si_add, dim3(vgrid(n)), dim3(256), 0, h->v.stream, n, static_cast<const double *>(h->phys[q]),
                         h->v.t[target[q]]);
    }
  VMI_HIP(hipGetLastError());
  return 0;
}

int assemble(vibe_mi_si *h, double dt) {
  const LiteralView &v = h->v;
  hipLaunchKernelGGL(k_si_assemble, dim3(ceil_div(v.nx, 64), v.ny, v.nz), dim3(64), 0, v.stream, h->g,
                     static_cast<const double *>(v.s[0]), static_cast<const double *>(v.s[1]),
                     static_cast<const double *>(v.s[2]), static_cast<const double *>(v.s[5]),
                     static_cast<const double *>(v.t[0]), static_cast<const double *>(v.t[1]),
                     static_cast<const double *>(v.t[2]), dt, h->acoef, h->bcoef, h->rhs);
  VMI_HIP(hipGetLastError());
  return 0;
}

int one_step(vibe_mi_si *h, double dt) {
  const LiteralView &v = h->v;
  hipStream_t st = v.stream;
  if (int rc = tendencies(h)) return rc;
  if (int rc = assemble(h, dt)) return rc;
  VMI_HIP(hipStreamSynchronize(st));
  if (dt != h->coef_dt) {
    h->coef_dt = -1.0;
    if (int rc = helmholtz_set_coefficients_device(h->helm, h->acoef, h->bcoef)) return rc;
    h->coef_dt = dt;
  }
  int its = 0;
  double rel = 0.0;
  if (int rc = helmholtz_solve_device(h->helm, h->rhs, h->pinew, h->solver, h->tol, h->max_iter, &its, &rel))
    return rc;
  h->stats.helmholtz_iterations += its;
  h->stats.helmholtz_residual = rel;
  Euler e{};
  for (int q : {4, 6, 3, 7, 8, 9, 10, 11}) {
    e.s[e.count] = v.s[q];
    e.t[e.count] = v.t[q];
    ++e.count;
  }
  const size_t nc = ncell(h);
  hipLaunchKernelGGL(k_si_euler, dim3(vgrid(nc)), dim3(256), 0, st, nc, e, dt);
  if (h->filter > 0.0 && !h->first_step)
    hipLaunchKernelGGL(k_si_filter, dim3(vgrid(nc)), dim3(256), 0, st, nc, 1.0 - h->filter, h->filter, v.s[5]);
  if (h->rate != nullptr && h->kmin < v.nz)
    hipLaunchKernelGGL(k_si_sponge, dim3(ceil_div(v.nx + 1, 64), v.ny + 1, v.nz + 1 - h->kmin), dim3(64), 0, st, h->g,
                       static_cast<const double *>(h->rate), h->kmin, dt, v.s[0], v.s[1], v.s[2], v.s[3], v.s[4],
                       v.s[5]);
  VMI_HIP(hipGetLastError());
  h->first_step = false;
  h->stats.steps += 1;
  hipLaunchKernelGGL(k_si_clear, dim3(1), dim3(64), 0, st, h->flag, h->mx);
  for (int q = 0; q < 12; ++q)
    hipLaunchKernelGGL(k_si_nonfinite, dim3(vgrid(v.n[q])), dim3(256), 0, st, v.n[q],
                       static_cast<const double *>(v.s[q]), h->flag);
  VMI_HIP(hipGetLastError());
  int bad = 0;
  VMI_HIP(hipMemcpyAsync(&bad, h->flag, sizeof(bad), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  if (bad) {
    h->stats.diverged = 1;
    return VIBE_MI_ERR_NONFINITE;
  }
  return 0;
}

}  // namespace

extern "C" {

int vibe_mi_si_create(const vibe_mi_si_desc *d, vibe_mi_si **out) {
  if (d == nullptr || out == nullptr || !(d->tol > 0.0) || d->max_iter < 0 || !(d->time_filter >= 0.0) ||
      (d->solver != VIBE_MI_HELM_BICGSTAB && d->solver != VIBE_MI_HELM_REFINE && d->solver != VIBE_MI_HELM_MULTIGRID))
    return VIBE_C_ERR_BAD_ARG;
  const vibe_mi_literal_desc &e = d->eq;
  if (e.nx <= 0 || e.ny <= 0 || e.nz <= 0 || e.ny >= 65535 || e.nz >= 65535) return VIBE_C_ERR_BAD_ARG;
  auto *h = new (std::nothrow) vibe_mi_si;
  if (h == nullptr) return VIBE_MI_ERR_HIP;
  auto fail = [&](int rc) {
    vibe_mi_si_destroy(h);
    return rc;
  };
  if (int rc = vibe_mi_literal_create(&e, &h->lit)) return fail(rc);
  const size_t nxy = static_cast<size_t>(e.nx) * e.ny;
  std::vector<double> zs0;
  if (e.zs == nullptr) zs0.assign(nxy, 0.0);
  vibe_mi_helmholtz_desc hd{};
  hd.nx = e.nx;
  hd.ny = e.ny;
  hd.nz = e.nz;
  hd.dx = e.dx;
  hd.dy = e.dy;
  hd.ztop = e.z_top;
  hd.zeta = e.zeta;
  hd.zs = e.zs != nullptr ? e.zs : zs0.data();
  if (int rc = vibe_mi_helmholtz_create(&hd, &h->helm)) return fail(rc);
  h->v = literal_view(h->lit);
  h->g = SiGrid{e.nx, e.ny, e.nz, e.dx, e.dy, e.z_top, h->v.zeta, h->v.zs, h->v.rho0, h->v.theta0, h->v.pi0};
  h->solver = d->solver;
  h->tol = d->tol;
  h->max_iter = d->max_iter;
  h->filter = d->time_filter;
  const size_t nc = h->v.n[3];
  for (double **p : {&h->acoef, &h->bcoef, &h->rhs, &h->pinew})
    if (int rc = si_alloc(h, p, nc)) return fail(rc);
  if (int rc = si_alloc(h, &h->flag, 1)) return fail(rc);
  if (int rc = si_alloc(h, &h->mx, 1)) return fail(rc);
  h->kmin = e.nz;
  if (d->sponge != nullptr) {
    for (int k = 0; k < e.nz; ++k)
      if (d->sponge[k] > 0.0) {
        h->kmin = k;
        break;
      }
    if (int rc = si_alloc(h, &h->rate, static_cast<size_t>(e.nz))) return fail(rc);
    if (hipMemcpy(h->rate, d->sponge, e.nz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(VIBE_MI_ERR_HIP);
  }
  if (hipMemset(h->pinew, 0, nc * sizeof(double)) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  *out = h;
  return 0;
}

int vibe_mi_si_destroy(vibe_mi_si *h) {
  if (h == nullptr) return 0;
  if (h->v.stream != nullptr) (void)hipStreamSynchronize(h->v.stream);
  for (void *p : h->allocs) (void)hipFree(p);
  for (double *p : h->phys)
    if (p != nullptr) (void)hipFree(p);
  if (h->helm != nullptr) vibe_mi_helmholtz_destroy(h->helm);
  if (h->lit != nullptr) vibe_mi_literal_destroy(h->lit);
  delete h;
  return 0;
}

int vibe_mi_si_set_state(vibe_mi_si *h, const double *const s[12]) {
  if (h == nullptr || s == nullptr) return VIBE_C_ERR_BAD_ARG;
  for (int q = 0; q < 12; ++q)
    if (s[q] == nullptr) return VIBE_C_ERR_BAD_ARG;
  for (int q = 0; q < 12; ++q)
    VMI_HIP(hipMemcpyAsync(h->v.s[q], s[q], h->v.n[q] * sizeof(double), hipMemcpyHostToDevice, h->v.stream));
  VMI_HIP(hipStreamSynchronize(h->v.stream));
  return 0;
}

int vibe_mi_si_get_state(vibe_mi_si *h, double *const s[12]) {
  if (h == nullptr || s == nullptr) return VIBE_C_ERR_BAD_ARG;
  for (int q = 0; q < 12; ++q)
    if (s[q] == nullptr) return VIBE_C_ERR_BAD_ARG;
  for (int q = 0; q < 12; ++q)
    VMI_HIP(hipMemcpyAsync(s[q], h->v.s[q], h->v.n[q] * sizeof(double), hipMemcpyDeviceToHost, h->v.stream));
  VMI_HIP(hipStreamSynchronize(h->v.stream));
  return 0;
}

int vibe_mi_si_set_physics_tendency(vibe_mi_si *h, const double *theta, const double *qv, const double *u,
                                    const double *v) {
  if (h == nullptr) return VIBE_C_ERR_BAD_ARG;
  const double *src[4] = {theta, qv, u, v};
  static const int target[4] = {4, 6, 0, 1};
  VMI_HIP(hipStreamSynchronize(h->v.stream));
  for (int q = 0; q < 4; ++q) {
    if (src[q] == nullptr) {
      if (h->phys[q] != nullptr) (void)hipFree(h->phys[q]);
      h->phys[q] = nullptr;
      continue;
    }
    const size_t n = h->v.n[target[q]];
    if (h->phys[q] == nullptr) VMI_HIP(hipMalloc(&h->phys[q], n * sizeof(double)));
    VMI_HIP(hipMemcpy(h->phys[q], src[q], n * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

int vibe_mi_si_set_time_filter(vibe_mi_si *h, double alpha) {
  if (h == nullptr || !(alpha >= 0.0)) return VIBE_C_ERR_BAD_ARG;
  h->filter = alpha;
  return 0;
}

int vibe_mi_si_step(vibe_mi_si *h, double dt, int nsteps) {
  if (h == nullptr || !(dt > 0.0) || nsteps < 0) return VIBE_C_ERR_BAD_ARG;
  const auto t0 = std::chrono::steady_clock::now();
  for (int n = 0; n < nsteps; ++n)
    if (int rc = one_step(h, dt)) return rc;
  h->stats.wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int vibe_mi_si_get_stats(vibe_mi_si *h, vibe_mi_si_stats *out) {
  if (h == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  *out = h->stats;
  return 0;
}

int vibe_mi_si_reset_stats(vibe_mi_si *h) {
  if (h == nullptr) return VIBE_C_ERR_BAD_ARG;
  h->stats = vibe_mi_si_stats{};
  return 0;
}

int vibe_mi_si_max_stable_dt(vibe_mi_si *h, double *dt_out) {
  if (h == nullptr || dt_out == nullptr) return VIBE_C_ERR_BAD_ARG;
  hipStream_t st = h->v.stream;
  hipLaunchKernelGGL(k_si_clear, dim3(1), dim3(64), 0, st, h->flag, h->mx);
  hipLaunchKernelGGL(k_si_cfl, dim3(vgrid(ncell(h))), dim3(256), 0, st, h->g, static_cast<const double *>(h->v.s[0]),
                     static_cast<const double *>(h->v.s[1]), static_cast<const double *>(h->v.s[2]), h->mx);
  VMI_HIP(hipGetLastError());
  unsigned long long bits = 0;
  VMI_HIP(hipMemcpyAsync(&bits, h->mx, sizeof(bits), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  const double cfl = __builtin_bit_cast(double, bits);
  *dt_out = cfl <= 0.0 ? 600.0 : 0.6 / cfl;
  return 0;
}

int vibe_mi_si_rhs(vibe_mi_si *h, double dt, double *a, double *rhs) {
  if (h == nullptr || !(dt > 0.0)) return VIBE_C_ERR_BAD_ARG;
  if (int rc = tendencies(h)) return rc;
  if (int rc = assemble(h, dt)) return rc;
  const size_t nc = ncell(h);
  if (a != nullptr) VMI_HIP(hipMemcpyAsync(a, h->acoef, nc * sizeof(double), hipMemcpyDeviceToHost, h->v.stream));
  if (rhs != nullptr) VMI_HIP(hipMemcpyAsync(rhs, h->rhs, nc * sizeof(double), hipMemcpyDeviceToHost, h->v.stream));
  VMI_HIP(hipStreamSynchronize(h->v.stream));
  return 0;
}

int vibe_mi_si_pressure(vibe_mi_si *h, double *pi_new) {
  if (h == nullptr || pi_new == nullptr) return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipMemcpy(pi_new, h->pinew, ncell(h) * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
