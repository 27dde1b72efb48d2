// literal.hip -- the reference's literal discrete equations on gfx950
// (vibe_mi/literal.h): Equations::tendencies, acoustic_tendencies and
// max_sound_speed with the reference's conventions and operation order.
//
// One lane per output point; each field family (u faces, v faces, w faces,
// cells) is one launch.  Built with -ffp-contract=off and IEEE division and
// square root, so every value is the reference's bits.  Not a hot path: the
// kernels read their neighbours straight from HBM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "literal_dev.hpp"
#include "vibe_mi/literal.h"

namespace vibe_mi {
namespace {

constexpr double kMoistB = 0.608;  // Rv/Rd - 1 (equations.cpp:40)

struct Lit {
  int nx, ny, nz;
  double dx, dy, ztop;
  const double *zeta, *zs, *r0, *t0, *p0, *fj;  // zs NULL: flat; fj: f at rows 0..ny
  int order, cor, dord, moist;
  double dcoef;
};

__device__ inline int cl_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ inline size_t ic(const Lit &g, int i, int j, int k) {
  return (static_cast<size_t>(k) * g.ny + j) * g.nx + i;
}
__device__ inline size_t iu(const Lit &g, int i, int j, int k) {
  return (static_cast<size_t>(k) * g.ny + j) * (g.nx + 1) + i;
}
__device__ inline size_t iv(const Lit &g, int i, int j, int k) {
  return (static_cast<size_t>(k) * (g.ny + 1) + j) * g.nx + i;
}
// Field::clamp_at of a cell field
__device__ inline double cat(const Lit &g, const double *f, int i, int j, int k) {
  return f[ic(g, cl_(i, 0, g.nx - 1), cl_(j, 0, g.ny - 1), cl_(k, 0, g.nz - 1))];
}
// Grid::dzeta / Grid::terrain: clamped indices
__device__ inline double dzeta(const Lit &g, int k) {
  const int kk = cl_(k, 0, g.nz - 1);
  return g.zeta[kk + 1] - g.zeta[kk];
}
__device__ inline double terr(const Lit &g, int i, int j) {
  return g.zs == nullptr ? 0.0 : g.zs[static_cast<size_t>(cl_(j, 0, g.ny - 1)) * g.nx + cl_(i, 0, g.nx - 1)];
}

__device__ inline double grad_x(const Lit &g, const double *q, int i, int j, int k, double r) {
  if (g.order == 4)
    return (cat(g, q, i - 2, j, k) - 8.0 * cat(g, q, i - 1, j, k) + 8.0 * cat(g, q, i + 1, j, k) -
            cat(g, q, i + 2, j, k)) *
           (r / 12.0);
  if (g.order == 6)
    return (-cat(g, q, i - 3, j, k) + 9.0 * cat(g, q, i - 2, j, k) - 45.0 * cat(g, q, i - 1, j, k) +
            45.0 * cat(g, q, i + 1, j, k) - 9.0 * cat(g, q, i + 2, j, k) + cat(g, q, i + 3, j, k)) *
           (r / 60.0);
  return (cat(g, q, i + 1, j, k) - cat(g, q, i - 1, j, k)) * (0.5 * r);
}
__device__ inline double grad_y(const Lit &g, const double *q, int i, int j, int k, double r) {
  if (g.order == 4)
    return (cat(g, q, i, j - 2, k) - 8.0 * cat(g, q, i, j - 1, k) + 8.0 * cat(g, q, i, j + 1, k) -
            cat(g, q, i, j + 2, k)) *
           (r / 12.0);
  if (g.order == 6)
    return (-cat(g, q, i, j - 3, k) + 9.0 * cat(g, q, i, j - 2, k) - 45.0 * cat(g, q, i, j - 1, k) +
            45.0 * cat(g, q, i, j + 1, k) - 9.0 * cat(g, q, i, j + 2, k) + cat(g, q, i, j + 3, k)) *
           (r / 60.0);
  return (cat(g, q, i, j + 1, k) - cat(g, q, i, j - 1, k)) * (0.5 * r);
}

// Advection::scalar_nonconservative at one cell
__device__ inline double adv_nc(const Lit &g, const double *q, const double *u, const double *v, const double *w,
                                int i, int j, int k) {
  const double dz = dzeta(g, k) * (g.ztop - terr(g, i, j));
  const double ua = 0.5 * (u[iu(g, i, j, k)] + u[iu(g, i + 1, j, k)]);
  const double va = 0.5 * (v[iv(g, i, j, k)] + v[iv(g, i, j + 1, k)]);
  const double wa = 0.5 * (w[ic(g, i, j, k)] + w[ic(g, i, j, k + 1)]);
  return -(ua * grad_x(g, q, i, j, k, 1.0 / g.dx) + va * grad_y(g, q, i, j, k, 1.0 / g.dy) +
           wa * (cat(g, q, i, j, k + 1) - cat(g, q, i, j, k - 1)) * (0.5 / dz));
}

__device__ inline double lap2(const Lit &g, const double *f, int i, int j, int k) {
  return (cat(g, f, i + 1, j, k) - 2.0 * f[ic(g, i, j, k)] + cat(g, f, i - 1, j, k)) / (g.dx * g.dx) +
         (cat(g, f, i, j + 1, k) - 2.0 * f[ic(g, i, j, k)] + cat(g, f, i, j - 1, k)) / (g.dy * g.dy);
}

struct Fields {
  const double *s[12];
  double *t[12];
};

// u faces i = 0..nx: momentum advection, Coriolis, pressure gradient
__global__ void __launch_bounds__(256) k_lit_u(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i > g.nx) return;
  const double *u = F.s[0], *v = F.s[1], *w = F.s[2], *th = F.s[4], *pi = F.s[5];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int ip = min(i, nx - 1), im = max(i - 1, 0);
  const double dz = dzeta(g, min(k, nz - 1)) * (g.ztop - terr(g, ip, j));
  const double uc = 0.5 * (u[iu(g, im, j, k)] + u[iu(g, ip, j, k)]);
  const double ux = (u[iu(g, ip, j, k)] - u[iu(g, im, j, k)]) / g.dx;
  const double uy = (u[iu(g, ip, min(j + 1, ny - 1), k)] - u[iu(g, ip, max(j - 1, 0), k)]) / (2.0 * g.dy);
  const double uz = (u[iu(g, ip, j, min(k + 1, nz - 1))] - u[iu(g, ip, j, max(k - 1, 0))]) / (2.0 * dz);
  const double vc = 0.25 * (v[iv(g, im, max(j - 1, 0), k)] + v[iv(g, ip, max(j - 1, 0), k)] + v[iv(g, im, j, k)] +
                            v[iv(g, ip, j, k)]);
  const double wc = 0.5 * (w[ic(g, ip, j, k)] + w[ic(g, ip, j, min(k + 1, nz - 1))]);
  double acc = 0.0;
  acc -= uc * ux + vc * uy + wc * uz;
  if (g.cor != 0) {
    const double vu = 0.25 * (v[iv(g, im, max(j - 1, 0), k)] + v[iv(g, min(i, nx - 1), max(j - 1, 0), k)] +
                              v[iv(g, im, j, k)] + v[iv(g, min(i, nx - 1), j, k)]);
    acc += g.fj[j] * vu;
  }
  const double thf = 0.5 * (th[ic(g, im, j, k)] + th[ic(g, ip, j, k)]);
  acc -= kCp * thf * (pi[ic(g, ip, j, k)] - pi[ic(g, im, j, k)]) / g.dx;
  F.t[0][iu(g, i, j, k)] = acc;
}

// v faces j = 0..ny
__global__ void __launch_bounds__(256) k_lit_v(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const double *u = F.s[0], *v = F.s[1], *w = F.s[2], *th = F.s[4], *pi = F.s[5];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int jp = min(j, ny - 1), jm = max(j - 1, 0);
  const double dz = dzeta(g, min(k, nz - 1)) * (g.ztop - terr(g, i, jp));
  const double vc = 0.5 * (v[iv(g, i, jm, k)] + v[iv(g, i, jp, k)]);
  const double vy = (v[iv(g, i, jp, k)] - v[iv(g, i, jm, k)]) / g.dy;
  const double vx = (v[iv(g, min(i + 1, nx - 1), jp, k)] - v[iv(g, max(i - 1, 0), jp, k)]) / (2.0 * g.dx);
  const double vz = (v[iv(g, i, jp, min(k + 1, nz - 1))] - v[iv(g, i, jp, max(k - 1, 0))]) / (2.0 * dz);
  const double uc = 0.25 * (u[iu(g, max(i - 1, 0), jm, k)] + u[iu(g, i, jm, k)] + u[iu(g, max(i - 1, 0), jp, k)] +
                            u[iu(g, i, jp, k)]);
  const double wc = 0.5 * (w[ic(g, i, jp, k)] + w[ic(g, i, jp, min(k + 1, nz - 1))]);
  double acc = 0.0;
  acc -= uc * vx + vc * vy + wc * vz;
  if (g.cor != 0) {
    const double uv = 0.25 * (u[iu(g, max(i - 1, 0), max(j - 1, 0), k)] + u[iu(g, i, max(j - 1, 0), k)] +
                              u[iu(g, max(i - 1, 0), min(j, ny - 1), k)] + u[iu(g, i, min(j, ny - 1), k)]);
    acc -= g.fj[j] * uv;
  }
  const double thf = 0.5 * (th[ic(g, i, jm, k)] + th[ic(g, i, jp, k)]);
  acc -= kCp * thf * (pi[ic(g, i, jp, k)] - pi[ic(g, i, jm, k)]) / g.dy;
  F.t[1][iv(g, i, j, k)] = acc;
}

// w faces k = 0..nz: advection, vertical pressure gradient, buoyancy
__global__ void __launch_bounds__(256) k_lit_w(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const double *u = F.s[0], *v = F.s[1], *w = F.s[2], *th = F.s[4], *pi = F.s[5];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int kc = min(max(k, 0), nz - 1), kt = min(k, nz), kb = max(k - 1, 0), km = kb;
  const double scale = g.ztop - terr(g, i, j);
  const double dza = dzeta(g, kc) * scale;
  const double wc = 0.5 * (w[ic(g, i, j, kc)] + w[ic(g, i, j, kt)]);
  const double wz = (w[ic(g, i, j, kt)] - w[ic(g, i, j, kb)]) / dza;
  const double wx = (w[ic(g, min(i + 1, nx - 1), j, kt)] - w[ic(g, max(i - 1, 0), j, kt)]) / (2.0 * g.dx);
  const double wy = (w[ic(g, i, min(j + 1, ny - 1), kt)] - w[ic(g, i, max(j - 1, 0), kt)]) / (2.0 * g.dy);
  const double uc = 0.5 * (u[iu(g, i, j, kc)] + u[iu(g, i + 1, j, kc)]);
  const double vc = 0.5 * (v[iv(g, i, j, kc)] + v[iv(g, i, j + 1, kc)]);
  double acc = 0.0;
  acc -= uc * wx + vc * wy + wc * wz;
  const double dz = dza > 1.0 ? dza : 1.0;
  const double thf = 0.5 * (th[ic(g, i, j, km)] + th[ic(g, i, j, kc)]);
  acc -= kCp * thf * (pi[ic(g, i, j, kc)] - pi[ic(g, i, j, km)]) / dz;
  const size_t c = ic(g, i, j, kc);
  const double th0 = g.t0[c];
  const double thp = th[c] - th0;
  const double qcond = F.s[7][c] + F.s[8][c] + F.s[9][c] + F.s[10][c] + F.s[11][c];
  acc += kG * (thp / (th0 > 1.0 ? th0 : 1.0) + kMoistB * F.s[6][c] - qcond);
  F.t[2][ic(g, i, j, k)] = acc;
}

// cells: mass continuity (rho(nx, j) and rho(i, ny) are halo reads: zero),
// theta and water advection, positivity, Del^2 divergence damping of pi
__global__ void __launch_bounds__(256) k_lit_cells(Lit g, Fields F, double *lap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const double *u = F.s[0], *v = F.s[1], *w = F.s[2], *rho = F.s[3], *pi = F.s[5];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t c = ic(g, i, j, k);
  const double dzv = dzeta(g, k) * (g.ztop - terr(g, i, j));
  const double dz = dzv > 1.0 ? dzv : 1.0;
  const double re = i + 1 < nx ? rho[ic(g, i + 1, j, k)] : 0.0;
  const double rn = j + 1 < ny ? rho[ic(g, i, j + 1, k)] : 0.0;
  const double fx = u[iu(g, i + 1, j, k)] * 0.5 * (rho[c] + re) -
                    u[iu(g, i, j, k)] * 0.5 * (rho[ic(g, max(i - 1, 0), j, k)] + rho[c]);
  const double fy = v[iv(g, i, j + 1, k)] * 0.5 * (rho[c] + rn) -
                    v[iv(g, i, j, k)] * 0.5 * (rho[ic(g, i, max(j - 1, 0), k)] + rho[c]);
  const double fz = w[ic(g, i, j, k + 1)] * 0.5 * (rho[c] + rho[ic(g, i, j, min(k + 1, nz - 1))]) -
                    w[c] * 0.5 * (rho[ic(g, i, j, max(k - 1, 0))] + rho[c]);
  F.t[3][c] = -(fx / g.dx + fy / g.dy + fz / dz);
  F.t[4][c] = adv_nc(g, F.s[4], u, v, w, i, j, k);
  for (int q = 6; q < 12; ++q) {
    double dq = (q == 6 || g.moist) ? adv_nc(g, F.s[q], u, v, w, i, j, k) : 0.0;
    if (F.s[q][c] <= 0.0 && dq < 0.0) dq = 0.0;
    F.t[q][c] = dq;
  }
  double dpi = 0.0;
  if (g.dcoef > 0.0) {
    if (g.dord <= 2)
      dpi -= g.dcoef * lap2(g, pi, i, j, k);
    else
      lap[c] = lap2(g, pi, i, j, k);
  }
  F.t[5][c] = dpi;
}

// Del^4 damping: Del^2 of the Del^2 held in lap
__global__ void __launch_bounds__(256) k_lit_dpi4(Lit g, const double *lap, double *dpi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const size_t c = ic(g, i, j, k);
  double acc = 0.0;
  acc -= g.dcoef * lap2(g, lap, i, j, k);
  dpi[c] = acc;
}

// acoustic tendencies (equations.cpp:152-209)
__global__ void __launch_bounds__(256) k_ac_u(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i > g.nx) return;
  const int ip = min(i, g.nx - 1), im = max(i - 1, 0);
  const double th0 = 0.5 * (g.t0[ic(g, im, j, k)] + g.t0[ic(g, ip, j, k)]);
  F.t[0][iu(g, i, j, k)] = -kCp * th0 * (F.s[5][ic(g, ip, j, k)] - F.s[5][ic(g, im, j, k)]) / g.dx;
}
__global__ void __launch_bounds__(256) k_ac_v(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const int jp = min(j, g.ny - 1), jm = max(j - 1, 0);
  const double th0 = 0.5 * (g.t0[ic(g, i, jm, k)] + g.t0[ic(g, i, jp, k)]);
  F.t[1][iv(g, i, j, k)] = -kCp * th0 * (F.s[5][ic(g, i, jp, k)] - F.s[5][ic(g, i, jm, k)]) / g.dy;
}
__global__ void __launch_bounds__(256) k_ac_w(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const int kc = min(max(k, 0), g.nz - 1), km = max(k - 1, 0);
  const double dzv = dzeta(g, kc) * (g.ztop - terr(g, i, j));
  const double dz = dzv > 1.0 ? dzv : 1.0;
  const double th0 = 0.5 * (g.t0[ic(g, i, j, km)] + g.t0[ic(g, i, j, kc)]);
  F.t[2][ic(g, i, j, k)] = -kCp * th0 * (F.s[5][ic(g, i, j, kc)] - F.s[5][ic(g, i, j, km)]) / dz;
}
__global__ void __launch_bounds__(256) k_ac_cells(Lit g, Fields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const double *u = F.s[0], *v = F.s[1], *w = F.s[2], *r0 = g.r0;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t c = ic(g, i, j, k);
  const double dzv = dzeta(g, k) * (g.ztop - terr(g, i, j));
  const double dz = dzv > 1.0 ? dzv : 1.0;
  const double fx = u[iu(g, i + 1, j, k)] * 0.5 * (r0[c] + r0[ic(g, min(i + 1, nx - 1), j, k)]) -
                    u[iu(g, i, j, k)] * 0.5 * (r0[ic(g, max(i - 1, 0), j, k)] + r0[c]);
  const double fy = v[iv(g, i, j + 1, k)] * 0.5 * (r0[c] + r0[ic(g, i, min(j + 1, ny - 1), k)]) -
                    v[iv(g, i, j, k)] * 0.5 * (r0[ic(g, i, max(j - 1, 0), k)] + r0[c]);
  const double fz = w[ic(g, i, j, k + 1)] * 0.5 * (r0[c] + r0[ic(g, i, j, min(k + 1, nz - 1))]) -
                    w[c] * 0.5 * (r0[ic(g, i, j, max(k - 1, 0))] + r0[c]);
  const double drho = -(fx / g.dx + fy / g.dy + fz / dz);
  F.t[3][c] = drho;
  const double rr = r0[c] > 1e-8 ? r0[c] : 1e-8;
  F.t[5][c] = (kRd / kCv) * (g.p0[c] / rr) * drho;
}

// Equations::max_sound_speed: maximum of non-negative doubles through their
// bit patterns
__global__ void __launch_bounds__(256) k_lit_cs(const double *th, const double *pi, size_t n,
                                                unsigned long long *res) {
  double best = 0.0;
  for (size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
       c += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double gamma = kCp / kCv;
    const double t = th[c] * pi[c];
    const double cs = sqrt(gamma * kRd * (t > 1.0 ? t : 1.0));
    if (cs > best) best = cs;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(h)) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(res, static_cast<unsigned long long>(__double_as_longlong(sh[0])));
}

__global__ void k_zero_ull(unsigned long long *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

}  // namespace
}  // namespace vibe_mi

using namespace vibe_mi;

struct vibe_mi_literal {
  Lit g{};
  hipStream_t stream = nullptr;
  double *geo = nullptr;  // zeta | zs | rho0 | theta0 | pi0 | f rows
  double *fields = nullptr;  // 12 state arrays, 12 tendencies, the Del^2 temporary
  unsigned long long *res = nullptr;
  size_t n[12] = {};
  double *s[12] = {}, *t[12] = {}, *lap = nullptr;
};

namespace {

int upload_state(vibe_mi_literal *h, const double *const s[12]) {
  for (int q = 0; q < 12; ++q) {
    if (s[q] == nullptr) return VIBE_C_ERR_BAD_ARG;
    VMI_HIP(hipMemcpyAsync(h->s[q], s[q], h->n[q] * sizeof(double), hipMemcpyHostToDevice, h->stream));
  }
  return 0;
}

int download(vibe_mi_literal *h, double *const t[12]) {
  for (int q = 0; q < 12; ++q) {
    if (t[q] == nullptr) return VIBE_C_ERR_BAD_ARG;
    VMI_HIP(hipMemcpyAsync(t[q], h->t[q], h->n[q] * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  VMI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

Fields fields_of(vibe_mi_literal *h) {
  Fields F;
  for (int q = 0; q < 12; ++q) {
    F.s[q] = h->s[q];
    F.t[q] = h->t[q];
  }
  return F;
}

}  // namespace

namespace vibe_mi {

LiteralView literal_view(vibe_mi_literal *h) {
  LiteralView v{};
  const Lit &g = h->g;
  v.nx = g.nx;
  v.ny = g.ny;
  v.nz = g.nz;
  v.dx = g.dx;
  v.dy = g.dy;
  v.ztop = g.ztop;
  v.zeta = g.zeta;
  v.zs = g.zs;
  v.rho0 = g.r0;
  v.theta0 = g.t0;
  v.pi0 = g.p0;
  for (int q = 0; q < 12; ++q) {
    v.s[q] = h->s[q];
    v.t[q] = h->t[q];
    v.n[q] = h->n[q];
  }
  v.stream = h->stream;
  return v;
}

int literal_tendencies_device(vibe_mi_literal *h) {
  const Lit &g = h->g;
  const Fields F = fields_of(h);
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_lit_u, dim3(ceil_div(g.nx + 1, 64), g.ny, g.nz), dim3(64), 0, st, g, F);
  hipLaunchKernelGGL(k_lit_v, dim3(ceil_div(g.nx, 64), g.ny + 1, g.nz), dim3(64), 0, st, g, F);
  hipLaunchKernelGGL(k_lit_w, dim3(ceil_div(g.nx, 64), g.ny, g.nz + 1), dim3(64), 0, st, g, F);
  hipLaunchKernelGGL(k_lit_cells, dim3(ceil_div(g.nx, 64), g.ny, g.nz), dim3(64), 0, st, g, F, h->lap);
  if (g.dcoef > 0.0 && g.dord > 2)
    hipLaunchKernelGGL(k_lit_dpi4, dim3(ceil_div(g.nx, 64), g.ny, g.nz), dim3(64), 0, st, g,
                       static_cast<const double *>(h->lap), h->t[5]);
  VMI_HIP(hipGetLastError());
  return 0;
}

}  // namespace vibe_mi

extern "C" {

int vibe_mi_literal_create(const vibe_mi_literal_desc *d, vibe_mi_literal **out) {
  if (d == nullptr || out == nullptr || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || d->zeta == nullptr ||
      d->rho0 == nullptr || d->theta0 == nullptr || d->pi0 == nullptr || d->dx <= 0 || d->dy <= 0 ||
      (d->order != 2 && d->order != 4 && d->order != 6) || d->coriolis < 0 || d->coriolis > 2)
    return VIBE_C_ERR_BAD_ARG;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return VIBE_MI_ERR_NO_DEVICE;
  auto *h = new (std::nothrow) vibe_mi_literal;
  if (h == nullptr) return VIBE_MI_ERR_HIP;
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const size_t nxy = static_cast<size_t>(nx) * ny, nc = nxy * nz;
  // Coriolis parameter per row (coriolis.cpp:22-38; single rank, js = 0),
  // a set-up constant formed on the host with the reference's expression
  std::vector<double> f(static_cast<size_t>(ny) + 1, 0.0);
  const double kPiC = 3.14159265358979323846, kOmega = 7.2921159e-5, kEarth = 6.371229e6;
  const double f0 = 2.0 * kOmega * std::sin(d->latitude * (kPiC / 180.0));
  const double beta = 2.0 * kOmega * std::cos(d->latitude * (kPiC / 180.0)) / kEarth;
  for (int j = 0; j <= ny; ++j) {
    if (d->coriolis == 2) {
      const double y = (0.0 + static_cast<double>(j) + 0.5) * d->dy;
      const double y0 = 0.5 * static_cast<double>(ny) * d->dy;
      f[j] = f0 + beta * (y - y0);
    } else {
      f[j] = f0;
    }
  }
  const size_t ngeo = (nz + 1) + nxy + 3 * nc + (ny + 1);
  auto fail = [&](int rc) {
    vibe_mi_literal_destroy(h);
    return rc;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  if (hipMalloc(&h->geo, ngeo * sizeof(double)) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  const size_t sizes[12] = {(nx + 1) * static_cast<size_t>(ny) * nz, nx * static_cast<size_t>(ny + 1) * nz,
                            nxy * (nz + 1), nc, nc, nc, nc, nc, nc, nc, nc, nc};
  size_t tot = nc;
  for (int q = 0; q < 12; ++q) {
    h->n[q] = sizes[q];
    tot += 2 * sizes[q];
  }
  if (hipMalloc(&h->fields, tot * sizeof(double)) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  if (hipMalloc(&h->res, 256) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  double *p = h->fields;
  for (int q = 0; q < 12; ++q) {
    h->s[q] = p;
    p += sizes[q];
    h->t[q] = p;
    p += sizes[q];
  }
  h->lap = p;
  double *gz = h->geo, *gs = gz + (nz + 1), *gr = gs + nxy, *gt = gr + nc, *gp = gt + nc, *gf = gp + nc;
  bool ok = hipMemcpy(gz, d->zeta, (nz + 1) * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  if (d->zs != nullptr) ok = ok && hipMemcpy(gs, d->zs, nxy * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(gr, d->rho0, nc * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(gt, d->theta0, nc * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(gp, d->pi0, nc * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(gf, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return fail(VIBE_MI_ERR_HIP);
  h->g = Lit{nx, ny, nz, d->dx, d->dy, d->z_top, gz, d->zs ? gs : nullptr, gr, gt, gp, gf,
             d->order, d->coriolis, d->div_order, d->moist ? 1 : 0, d->div_coeff};
  *out = h;
  return 0;
}

int vibe_mi_literal_destroy(vibe_mi_literal *h) {
  if (h == nullptr) return 0;
  if (h->stream != nullptr) (void)hipStreamSynchronize(h->stream);
  if (h->geo != nullptr) (void)hipFree(h->geo);
  if (h->fields != nullptr) (void)hipFree(h->fields);
  if (h->res != nullptr) (void)hipFree(h->res);
  if (h->stream != nullptr) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int vibe_mi_literal_tendencies(vibe_mi_literal *h, const double *const s[12], double *const t[12]) {
  if (h == nullptr || s == nullptr || t == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (int rc = upload_state(h, s)) return rc;
  if (int rc = vibe_mi::literal_tendencies_device(h)) return rc;
  return download(h, t);
}

int vibe_mi_literal_acoustic(vibe_mi_literal *h, const double *const s[12], double *const t[12]) {
  if (h == nullptr || s == nullptr || t == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (int rc = upload_state(h, s)) return rc;
  const Lit &g = h->g;
  const Fields F = fields_of(h);
  hipStream_t st = h->stream;
  for (int q : {4, 6, 7, 8, 9, 10, 11})
    VMI_HIP(hipMemsetAsync(h->t[q], 0, h->n[q] * sizeof(double), st));
  hipLaunchKernelGGL(k_ac_u, dim3(ceil_div(g.nx + 1, 64), g.ny, g.nz), dim3(64), 0, st, g, F);
  hipLaunchKernelGGL(k_ac_v, dim3(ceil_div(g.nx, 64), g.ny + 1, g.nz), dim3(64), 0, st, g, F);
  hipLaunchKernelGGL(k_ac_w, dim3(ceil_div(g.nx, 64), g.ny, g.nz + 1), dim3(64), 0, st, g, F);
  hipLaunchKernelGGL(k_ac_cells, dim3(ceil_div(g.nx, 64), g.ny, g.nz), dim3(64), 0, st, g, F);
  VMI_HIP(hipGetLastError());
  return download(h, t);
}

int vibe_mi_literal_max_sound_speed(vibe_mi_literal *h, const double *const s[12], double *out) {
  if (h == nullptr || s == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (int rc = upload_state(h, s)) return rc;
  const size_t n = h->n[3];
  hipLaunchKernelGGL(k_zero_ull, dim3(1), dim3(64), 0, h->stream, h->res);
  hipLaunchKernelGGL(k_lit_cs, dim3(std::min<size_t>(ceil_div(n, 256), 1024)), dim3(256), 0, h->stream, h->s[4],
                     h->s[5], n, h->res);
  VMI_HIP(hipGetLastError());
  unsigned long long r = 0;
  VMI_HIP(hipMemcpyAsync(&r, h->res, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  VMI_HIP(hipStreamSynchronize(h->stream));
  *out = __builtin_bit_cast(double, r);
  return 0;
}

}  // extern "C"
