// helmholtz.hip -- 3-D Helmholtz solver of the semi-implicit path on the GPU
// (C ABI in include/vibe_mi/helmholtz.h).
//
// Operator (src/time/helmholtz.cpp:173-201 of the reference): one thread
// per column marching up the levels, 7-point stencil with face coefficients
// the mean of the two cell values of a, neighbours clamped at every
// boundary, dz from the column's depth.  The FP64 instance performs the reference's operations in the
// reference's order, so with -ffp-contract=off it rounds like a CPU
// evaluation of the same expression.
//
// Solver: BiCGSTAB (van der Vorst) right-preconditioned by exact solves of
// the vertical lines.  Thin layers make the vertical couplings (a/dz^2) far
// larger than the horizontal ones, so the per-column tridiagonal part of A
// (vertical couplings plus the horizontal diagonal) captures the stiffness:
// one thread per column runs the factored Thomas sweep.  All iteration
// scalars (rho, alpha, omega, beta and the dot products) live in device
// memory: dot products reduce per block into fixed slots and a one-thread
// kernel combines them in a fixed order (run-to-run reproducible), so an
// iteration is a chain of kernels on one stream; the host reads the residual
// every few iterations.  The mixed solver runs the same BiCGSTAB in FP32 on
// the FP64 residual equation (iterative refinement); the two-grid solver
// starts the FP64 BiCGSTAB from the prolonged solution of the 2:1 coarser
// problem (restricted right-hand side, coefficients and terrain).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "helmholtz_dev.hpp"
#include "vibe_mi/helmholtz.h"

namespace vibe_mi {
namespace {

constexpr int kDotBlocks = 1024;  // partial sums per dot product (fixed: reproducible)
constexpr int kCheckEvery = 4;    // iterations between host residual reads

// scalar slots (FP64) of one BiCGSTAB instance
enum { S_RHO, S_ALPHA, S_OMEGA, S_BETA, S_RV, S_TS, S_TT, S_RR, S_FLAG, S_COUNT };

template <typename T>
struct Op {
  int nx, ny, nz;
  T dx2, dy2, ztop;  // dx*dx, dy*dy
  const T *zs;       // ny*nx
  const T *dzeta;    // nz
  const T *a, *b;    // cells
};

__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// y = A x (mode 0) or y = f - A x (mode 1).  One thread per column of a
// 64 x 4 tile marching up the levels: the column's x and a at k-1, k, k+1
// stay in registers (the next level's are loaded one level ahead), the
// E/W/N/S neighbours of level k come from the same or adjacent rows of the
// tile (cache hits for all but the tile's edge rows).  Same operations in the
// same order as the reference's loop body.
//
// With part != nullptr the kernel also forms the block's partial sums of
// <d, y> and (tt) <y, y> for the BiCGSTAB scalars, part[2 * block + 0/1]
// (block = blockIdx.y * gridDim.x + blockIdx.x), so no separate pass reads
// y again.
constexpr int kApplyRows = 4;
__device__ __forceinline__ void block_partials(double s0, double s1, double *__restrict__ part, int blk) {
  __shared__ double sh[2][256];
  const int t = static_cast<int>(threadIdx.y * blockDim.x + threadIdx.x);
  sh[0][t] = s0;
  sh[1][t] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sh[0][t] += sh[0][t + w];
      sh[1][t] += sh[1][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[2 * blk] = sh[0][0];
    part[2 * blk + 1] = sh[1][0];
  }
}
template <typename T>
__global__ void __launch_bounds__(64 * kApplyRows) k_apply(Op<T> o, const T *__restrict__ x, const T *__restrict__ f,
                                                           T *__restrict__ y, int mode, const T *__restrict__ d,
                                                           int tt, double *__restrict__ part) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * kApplyRows + threadIdx.y;
  double s0 = 0.0, s1 = 0.0;
  if (i < o.nx && j < o.ny) {
  const size_t nxy = static_cast<size_t>(o.nx) * o.ny;
  const size_t col = static_cast<size_t>(j) * o.nx + i;
  const size_t cE = static_cast<size_t>(j) * o.nx + clampi(i + 1, o.nx);
  const size_t cW = static_cast<size_t>(j) * o.nx + clampi(i - 1, o.nx);
  const size_t cN = static_cast<size_t>(clampi(j + 1, o.ny)) * o.nx + i;
  const size_t cS = static_cast<size_t>(clampi(j - 1, o.ny)) * o.nx + i;
  const T depth = o.ztop - o.zs[col];
  const int k1 = o.nz > 1 ? 1 : 0;
  T xm = x[col], xc = xm, xp = x[static_cast<size_t>(k1) * nxy + col];
  T am = o.a[col], ac = am, ap = o.a[static_cast<size_t>(k1) * nxy + col];
  // level k's horizontal neighbours and b, loaded during level k-1
  T q[9];
  auto ld_nb = [&](int k, T *v) {
    const size_t lv = static_cast<size_t>(k) * nxy;
    v[0] = x[lv + cE];
    v[1] = x[lv + cW];
    v[2] = x[lv + cN];
    v[3] = x[lv + cS];
    v[4] = o.a[lv + cE];
    v[5] = o.a[lv + cW];
    v[6] = o.a[lv + cN];
    v[7] = o.a[lv + cS];
    v[8] = o.b[lv + col];
  };
  ld_nb(0, q);
  for (int k = 0; k < o.nz; ++k) {
    const size_t lv = static_cast<size_t>(k) * nxy;
    const size_t ln = static_cast<size_t>(k + 2 < o.nz ? k + 2 : o.nz - 1) * nxy;
    const T xn = x[ln + col], an_ = o.a[ln + col];  // level k+2, one level ahead
    T qn[9];
    ld_nb(k + 1 < o.nz ? k + 1 : k, qn);  // level k+1's neighbours (the last level reloads itself)
    const T xE = q[0], xW = q[1], xN = q[2], xS = q[3];
    const T aE = q[4], aW = q[5], aN = q[6], aS = q[7];
    const T bc = q[8];
    T dz = o.dzeta[k] * depth;
    dz = dz > T(1) ? dz : T(1);
    const T ae = T(0.5) * (ac + aE);
    const T aw = T(0.5) * (ac + aW);
    const T an = T(0.5) * (ac + aN);
    const T as = T(0.5) * (ac + aS);
    const T atp = T(0.5) * (ac + ap);
    const T ab = T(0.5) * (ac + am);
    const T fx = (ae * (xE - xc) - aw * (xc - xW)) / o.dx2;
    const T fy = (an * (xN - xc) - as * (xc - xS)) / o.dy2;
    const T fz = (atp * (xp - xc) - ab * (xc - xm)) / (dz * dz);
    const T ax = -fx - fy - fz + bc * xc;
    const T yv = mode ? f[lv + col] - ax : ax;
    y[lv + col] = yv;
    if (part != nullptr) {
      s0 += static_cast<double>(d[lv + col]) * static_cast<double>(yv);
      if (tt) s1 += static_cast<double>(yv) * static_cast<double>(yv);
    }
    xm = xc;
    xc = xp;
    xp = xn;
    am = ac;
    ac = ap;
    ap = an_;
#pragma unroll
    for (int e = 0; e < 9; ++e) q[e] = qn[e];
  }
  }
  if (part != nullptr) block_partials(s0, s1, part, static_cast<int>(blockIdx.y * gridDim.x + blockIdx.x));
}

// Factorisation of the line preconditioner, one thread per column: rows
// lo_k x_{k-1} + d_k x_k + up_k x_{k+1} with the vertical couplings of A and
// its horizontal diagonal (zero-flux boundaries drop the outer couplings).
// Stores lo_k, c'_k = up_k / den_k and 1 / den_k (Thomas).
__global__ void __launch_bounds__(256) k_line_factor(Op<double> o, double *lo, double *cp, double *inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= o.nx) return;
  const size_t nxy = static_cast<size_t>(o.nx) * o.ny;
  const size_t col = static_cast<size_t>(j) * o.nx + i;
  double cprev = 0.0;
  for (int k = 0; k < o.nz; ++k) {
    const size_t c = static_cast<size_t>(k) * nxy + col;
    auto A = [&](int ii, int jj, int kk) {
      return o.a[static_cast<size_t>(kk) * nxy + static_cast<size_t>(jj) * o.nx + ii];
    };
    const double ac = o.a[c];
    double dz = o.dzeta[k] * (o.ztop - o.zs[col]);
    dz = dz > 1.0 ? dz : 1.0;
    double dh = 0.0;
    if (i + 1 < o.nx) dh += 0.5 * (ac + A(i + 1, j, k)) / o.dx2;
    if (i > 0) dh += 0.5 * (ac + A(i - 1, j, k)) / o.dx2;
    if (j + 1 < o.ny) dh += 0.5 * (ac + A(i, j + 1, k)) / o.dy2;
    if (j > 0) dh += 0.5 * (ac + A(i, j - 1, k)) / o.dy2;
    const double up = k + 1 < o.nz ? -0.5 * (ac + A(i, j, k + 1)) / (dz * dz) : 0.0;
    const double low = k > 0 ? -0.5 * (ac + A(i, j, k - 1)) / (dz * dz) : 0.0;
    const double d = -up - low + dh + o.b[c];
    const double den = d - low * cprev;
    const double iv = 1.0 / den;
    lo[c] = low;
    cp[c] = up * iv;
    inv[c] = iv;
    cprev = up * iv;
  }
}

template <typename T>
struct Line {
  const T *lo, *cp, *inv;
};

// z = M^-1 r (forward and back substitution per column)
template <typename T>
__global__ void __launch_bounds__(256) k_line_solve(int nx, int ny, int nz, Line<T> m, const T *__restrict__ r,
                                                    T *__restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx) return;
  const size_t nxy = static_cast<size_t>(nx) * ny, col = static_cast<size_t>(j) * nx + i;
  T prev = T(0);
  for (int k = 0; k < nz; ++k) {
    const size_t c = static_cast<size_t>(k) * nxy + col;
    prev = (r[c] - m.lo[c] * prev) * m.inv[c];
    z[c] = prev;
  }
  T next = z[static_cast<size_t>(nz - 1) * nxy + col];
  for (int k = nz - 2; k >= 0; --k) {
    const size_t c = static_cast<size_t>(k) * nxy + col;
    next = z[c] - m.cp[c] * next;
    z[c] = next;
  }
}

// partial sums (FP64) of up to two dot products: part[2 * block + q]
template <typename T>
__global__ void __launch_bounds__(256) k_dot(size_t n, const T *__restrict__ x0, const T *__restrict__ y0,
                                             const T *__restrict__ x1, const T *__restrict__ y1,
                                             double *__restrict__ part) {
  __shared__ double sh[2][256];
  double s0 = 0.0, s1 = 0.0;
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256) {
    s0 += static_cast<double>(x0[q]) * static_cast<double>(y0[q]);
    if (x1 != nullptr) s1 += static_cast<double>(x1[q]) * static_cast<double>(y1[q]);
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// Sums the partials in a fixed order and updates the scalars.  op:
//   0  rho_new = <r^, r>: beta = (rho_new / rho)(alpha / omega), rho = rho_new
//   1  alpha = rho / <r^, v>
//   2  omega = <t, s> / <t, t>
//   3  rr = <r, r>
//   4  rr = <r, r> and, with the second partials, op 0 on <r^, r>
__global__ void __launch_bounds__(256) k_scalar(const double *__restrict__ part, int nblk, double *__restrict__ S,
                                                int op) {
  __shared__ double sh[2][256];
  double s0 = 0.0, s1 = 0.0;
  for (int q = threadIdx.x; q < nblk; q += 256) {
    s0 += part[2 * q];
    s1 += part[2 * q + 1];
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double d0 = sh[0][0], d1 = sh[1][0];
  if (op == 0) {
    if (d0 == 0.0 || S[S_OMEGA] == 0.0) S[S_FLAG] = 1.0;
    S[S_BETA] = S[S_FLAG] != 0.0 ? 0.0 : (d0 / S[S_RHO]) * (S[S_ALPHA] / S[S_OMEGA]);
    S[S_RHO] = d0;
  } else if (op == 1) {
    S[S_RV] = d0;
    if (d0 == 0.0) S[S_FLAG] = 1.0;
    S[S_ALPHA] = d0 != 0.0 ? S[S_RHO] / d0 : 0.0;
  } else if (op == 2) {
    S[S_TS] = d0;
    S[S_TT] = d1;
    S[S_OMEGA] = d1 != 0.0 ? d0 / d1 : 0.0;
  } else if (op == 3) {
    S[S_RR] = d0;
  } else {
    S[S_RR] = d0;
    if (d1 == 0.0 || S[S_OMEGA] == 0.0) S[S_FLAG] = 1.0;
    S[S_BETA] = S[S_FLAG] != 0.0 ? 0.0 : (d1 / S[S_RHO]) * (S[S_ALPHA] / S[S_OMEGA]);
    S[S_RHO] = d1;
  }
}

// p = r + beta (p - omega v)
template <typename T>
__global__ void __launch_bounds__(256) k_upd_p(size_t n, const double *__restrict__ S, const T *__restrict__ r,
                                               const T *__restrict__ v, T *__restrict__ p) {
  const T beta = static_cast<T>(S[S_BETA]), omega = static_cast<T>(S[S_OMEGA]);
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256)
    p[q] = r[q] + beta * (p[q] - omega * v[q]);
}
// s = r - alpha v
template <typename T>
__global__ void __launch_bounds__(256) k_upd_s(size_t n, const double *__restrict__ S, const T *__restrict__ r,
                                               const T *__restrict__ v, T *__restrict__ s) {
  const T alpha = static_cast<T>(S[S_ALPHA]);
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256)
    s[q] = r[q] - alpha * v[q];
}
// x += alpha y + omega z;  r = s - omega t; the block's partial sums of
// <r, r> and <r^, r> (the next iteration's rho) into part
template <typename T>
__global__ void __launch_bounds__(256) k_upd_xr(size_t n, const double *__restrict__ S, const T *__restrict__ y,
                                                const T *__restrict__ z, const T *__restrict__ s,
                                                const T *__restrict__ t, T *__restrict__ x, T *__restrict__ r,
                                                const T *__restrict__ rh, double *__restrict__ part) {
  const T alpha = static_cast<T>(S[S_ALPHA]), omega = static_cast<T>(S[S_OMEGA]);
  double s0 = 0.0, s1 = 0.0;
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256) {
    x[q] = x[q] + alpha * y[q] + omega * z[q];
    const T rn = s[q] - omega * t[q];
    r[q] = rn;
    s0 += static_cast<double>(rn) * static_cast<double>(rn);
    s1 += static_cast<double>(rh[q]) * static_cast<double>(rn);
  }
  block_partials(s0, s1, part, static_cast<int>(blockIdx.x));
}
// dst = scale * src (with conversion), or dst += scale * src (acc)
template <typename D, typename S_>
__global__ void __launch_bounds__(256) k_axpy(size_t n, double scale, const S_ *__restrict__ src, D *__restrict__ dst,
                                              int acc) {
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256) {
    const double v = scale * static_cast<double>(src[q]);
    dst[q] = acc ? static_cast<D>(static_cast<double>(dst[q]) + v) : static_cast<D>(v);
  }
}
template <typename T>
__global__ void __launch_bounds__(256) k_fill(size_t n, T v, T *__restrict__ dst) {
  for (size_t q = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<size_t>(gridDim.x) * 256)
    dst[q] = v;
}
// multigrid_restrict (helmholtz.cpp:131-145 of the reference): 2:1 full
// weighting in x and y at every level, fine indices clamped to the fine
// field (Field::clamp_at); the coarse field's own extents set the loop.
__global__ void __launch_bounds__(256) k_mg_restrict(const double *__restrict__ f, int fnx, int fny, int fnz,
                                                     double *__restrict__ c, int cnx, int cny) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= cnx) return;
  const double *fk = f + static_cast<size_t>(clampi(k, fnz)) * fny * fnx;
  auto at = [&](int ii, int jj) { return fk[static_cast<size_t>(clampi(jj, fny)) * fnx + clampi(ii, fnx)]; };
  const int fi = 2 * i, fj = 2 * j;
  double acc = 0.25 * at(fi, fj);
  acc += 0.125 * (at(fi - 1, fj) + at(fi + 1, fj) + at(fi, fj - 1) + at(fi, fj + 1));
  acc += 0.0625 * (at(fi - 1, fj - 1) + at(fi + 1, fj - 1) + at(fi - 1, fj + 1) + at(fi + 1, fj + 1));
  c[(static_cast<size_t>(k) * cny + j) * cnx + i] = acc;
}
// multigrid_prolong (helmholtz.cpp:147-164): bilinear interpolation, fine
// point i at coarse coordinate i/2, coarse indices clamped.
__global__ void __launch_bounds__(256) k_mg_prolong(const double *__restrict__ c, int cnx, int cny, int cnz,
                                                    double *__restrict__ f, int fnx, int fny) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= fnx) return;
  const double *ck = c + static_cast<size_t>(clampi(k, cnz)) * cny * cnx;
  auto at = [&](int ii, int jj) { return ck[static_cast<size_t>(clampi(jj, cny)) * cnx + clampi(ii, cnx)]; };
  auto lerp = [](double a, double b, double w) { return a + w * (b - a); };
  const double xs = 0.5 * static_cast<double>(i), ys = 0.5 * static_cast<double>(j);
  const int i0 = static_cast<int>(floor(xs)), j0 = static_cast<int>(floor(ys));
  const double wx = xs - static_cast<double>(i0), wy = ys - static_cast<double>(j0);
  f[(static_cast<size_t>(k) * fny + j) * fnx + i] =
      lerp(lerp(at(i0, j0), at(i0 + 1, j0), wx), lerp(at(i0, j0 + 1), at(i0 + 1, j0 + 1), wx), wy);
}

__global__ void k_init_scalars(double *S) {
  if (threadIdx.x == 0) {
    for (int q = 0; q < S_COUNT; ++q) S[q] = 0.0;
    S[S_RHO] = S[S_ALPHA] = S[S_OMEGA] = 1.0;
  }
}

// one precision's working set
template <typename T>
struct Work {
  T *a = nullptr, *b = nullptr, *zs = nullptr, *dzeta = nullptr;
  T *lo = nullptr, *cp = nullptr, *inv = nullptr;
  T *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *y = nullptr, *z = nullptr,
    *t = nullptr;
  double *S = nullptr;
};

}  // namespace
}  // namespace vibe_mi

using namespace vibe_mi;

struct vibe_mi_helmholtz {
  int nx = 0, ny = 0, nz = 0;
  size_t n = 0;
  double dx = 0, dy = 0, ztop = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  Work<double> w64;
  Work<float> w32;
  double *f64 = nullptr;  // right-hand side
  double *part = nullptr;  // dot partials
  bool coeffs = false;
  unsigned vgrid = 1;      // grid of the vector kernels
  // two-grid start of VIBE_MI_HELM_MULTIGRID: the 2:1 coarser operator
  // (coefficients and terrain restricted), built at its first use after
  // set_coefficients
  vibe_mi_helmholtz *coarse = nullptr;
  bool coarse_ready = false;
};

namespace vibe_mi {
namespace {

template <typename P>
int halloc(vibe_mi_helmholtz *h, P **out, size_t count) {
  void *p = nullptr;
  VMI_HIP(hipMalloc(&p, (count > 0 ? count : 1) * sizeof(P)));
  h->allocs.push_back(p);
  *out = static_cast<P *>(p);
  return 0;
}

template <typename T>
int alloc_work(vibe_mi_helmholtz *h, Work<T> &w) {
  const size_t n = h->n, nxy = static_cast<size_t>(h->nx) * h->ny;
  for (T **p : {&w.a, &w.b, &w.lo, &w.cp, &w.inv, &w.x, &w.r, &w.rh, &w.p, &w.v, &w.s, &w.y, &w.z, &w.t})
    if (int rc = halloc(h, p, n)) return rc;
  if (int rc = halloc(h, &w.zs, nxy)) return rc;
  if (int rc = halloc(h, &w.dzeta, static_cast<size_t>(h->nz))) return rc;
  if (int rc = halloc(h, &w.S, static_cast<size_t>(S_COUNT))) return rc;
  return 0;
}

template <typename T>
Op<T> op_of(const vibe_mi_helmholtz *h, const Work<T> &w) {
  return Op<T>{h->nx, h->ny, h->nz, static_cast<T>(h->dx * h->dx), static_cast<T>(h->dy * h->dy),
               static_cast<T>(h->ztop), w.zs, w.dzeta, w.a, w.b};
}

// blocks of the operator's grid (the partial sums it forms with part)
unsigned apply_blocks(const vibe_mi_helmholtz *h) { return ceil_div(h->nx, 64) * ceil_div(h->ny, kApplyRows); }

// y = A x (mode 0) or f - A x (mode 1); with part, also the partial sums of
// <d, y> (and <y, y> with tt) for k_scalar over apply_blocks(h) blocks
template <typename T>
int apply(vibe_mi_helmholtz *h, const Work<T> &w, const T *x, const T *f, T *y, int mode, const T *d = nullptr,
          int tt = 0, double *part = nullptr) {
  hipLaunchKernelGGL((k_apply<T>), dim3(ceil_div(h->nx, 64), ceil_div(h->ny, kApplyRows)), dim3(64, kApplyRows), 0,
                     h->stream, op_of(h, w), x, f, y, mode, d, tt, part);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int dot(vibe_mi_helmholtz *h, double *S, const T *x0, const T *y0, const T *x1, const T *y1, int op) {
  hipLaunchKernelGGL((k_dot<T>), dim3(kDotBlocks), dim3(256), 0, h->stream, h->n, x0, y0, x1, y1, h->part);
  hipLaunchKernelGGL(k_scalar, dim3(1), dim3(256), 0, h->stream, h->part, kDotBlocks, S, op);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int line(vibe_mi_helmholtz *h, const Work<T> &w, const T *r, T *z) {
  hipLaunchKernelGGL((k_line_solve<T>), dim3(ceil_div(h->nx, 256), h->ny), dim3(256), 0, h->stream, h->nx, h->ny,
                     h->nz, Line<T>{w.lo, w.cp, w.inv}, r, z);
  VMI_HIP(hipGetLastError());
  return 0;
}

// restrict / prolong between two cell fields on `stream`
int mg_restrict(const double *f, int fnx, int fny, int fnz, double *c, int cnx, int cny, int cnz, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_restrict, dim3(ceil_div(cnx, 256), cny, cnz), dim3(256), 0, st, f, fnx, fny, fnz, c, cnx,
                     cny);
  VMI_HIP(hipGetLastError());
  return 0;
}
int mg_prolong(const double *c, int cnx, int cny, int cnz, double *f, int fnx, int fny, int fnz, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_prolong, dim3(ceil_div(fnx, 256), fny, fnz), dim3(256), 0, st, c, cnx, cny, cnz, f, fnx,
                     fny);
  VMI_HIP(hipGetLastError());
  return 0;
}

// BiCGSTAB on A x = rhs (rhs in w.r on entry) until
// sqrt(<r,r>) <= tol * sqrt(<rhs,rhs>) or max_iter iterations; x starts
// from zero, or from w.x when from_x (r = rhs - A x first); returns the
// iterations taken in *its.
template <typename T>
int bicgstab(vibe_mi_helmholtz *h, Work<T> &w, double tol, int max_iter, int *its, bool from_x = false) {
  const size_t n = h->n;
  const unsigned g = h->vgrid;
  hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, h->stream, w.S);
  if (!from_x) hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(256), 0, h->stream, n, T(0), w.x);
  hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(256), 0, h->stream, n, T(0), w.p);
  hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(256), 0, h->stream, n, T(0), w.v);
  if (int rc = dot(h, w.S, w.r, w.r, static_cast<const T *>(nullptr), static_cast<const T *>(nullptr), 3)) return rc;
  double S0[S_COUNT];
  VMI_HIP(hipMemcpyAsync(S0, w.S, sizeof(S0), hipMemcpyDeviceToHost, h->stream));
  VMI_HIP(hipStreamSynchronize(h->stream));
  const double rr0 = S0[S_RR];
  *its = 0;
  if (!(rr0 > 0.0)) {  // zero right-hand side: x = 0
    if (from_x) hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(256), 0, h->stream, n, T(0), w.x);
    VMI_HIP(hipGetLastError());
    return 0;
  }
  if (from_x) {
    if (int rc = apply(h, w, static_cast<const T *>(w.x), static_cast<const T *>(w.r), w.t, 1)) return rc;
    VMI_HIP(hipMemcpyAsync(w.r, w.t, n * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
    if (int rc = dot(h, w.S, w.r, w.r, static_cast<const T *>(nullptr), static_cast<const T *>(nullptr), 3)) return rc;
  }
  VMI_HIP(hipMemcpyAsync(w.rh, w.r, n * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
  const double stop = tol * tol * rr0;
  int it = 0;
  while (it < max_iter) {
    const int burst = std::min(kCheckEvery, max_iter - it);
    const int nab = static_cast<int>(apply_blocks(h));
    for (int q = 0; q < burst; ++q) {
      // rho = <r^, r>: a pass of its own on the first iteration, afterwards
      // formed by the previous iteration's x/r update (op 4)
      if (it + q == 0)
        if (int rc = dot(h, w.S, w.rh, w.r, static_cast<const T *>(nullptr), static_cast<const T *>(nullptr), 0))
          return rc;
      hipLaunchKernelGGL((k_upd_p<T>), dim3(g), dim3(256), 0, h->stream, n, w.S, w.r, w.v, w.p);
      if (int rc = line(h, w, w.p, w.y)) return rc;
      // v = A y with <r^, v>
      if (int rc = apply(h, w, w.y, static_cast<const T *>(nullptr), w.v, 0, static_cast<const T *>(w.rh), 0, h->part))
        return rc;
      hipLaunchKernelGGL(k_scalar, dim3(1), dim3(256), 0, h->stream, h->part, nab, w.S, 1);
      hipLaunchKernelGGL((k_upd_s<T>), dim3(g), dim3(256), 0, h->stream, n, w.S, w.r, w.v, w.s);
      if (int rc = line(h, w, w.s, w.z)) return rc;
      // t = A z with <s, t> and <t, t>
      if (int rc = apply(h, w, w.z, static_cast<const T *>(nullptr), w.t, 0, static_cast<const T *>(w.s), 1, h->part))
        return rc;
      hipLaunchKernelGGL(k_scalar, dim3(1), dim3(256), 0, h->stream, h->part, nab, w.S, 2);
      // x, r with <r, r> and the next rho
      hipLaunchKernelGGL((k_upd_xr<T>), dim3(g), dim3(256), 0, h->stream, n, w.S, w.y, w.z, w.s, w.t, w.x, w.r,
                         static_cast<const T *>(w.rh), h->part);
      hipLaunchKernelGGL(k_scalar, dim3(1), dim3(256), 0, h->stream, h->part, static_cast<int>(g), w.S, 4);
      VMI_HIP(hipGetLastError());
    }
    it += burst;
    VMI_HIP(hipMemcpyAsync(S0, w.S, sizeof(S0), hipMemcpyDeviceToHost, h->stream));
    VMI_HIP(hipStreamSynchronize(h->stream));
    if (S0[S_RR] <= stop || S0[S_FLAG] != 0.0 || !std::isfinite(S0[S_RR])) break;
  }
  *its = it;
  return 0;
}

// FP64 relative residual ||f - A x|| / ||f|| of w64.x (r64 <- f - A x)
int true_residual(vibe_mi_helmholtz *h, double *rel, double *rnorm) {
  Work<double> &w = h->w64;
  if (int rc = apply(h, w, w.x, static_cast<const double *>(h->f64), w.r, 1)) return rc;
  if (int rc = dot(h, w.S, w.r, w.r, static_cast<const double *>(h->f64), static_cast<const double *>(h->f64), 2))
    return rc;
  double S0[S_COUNT];
  VMI_HIP(hipMemcpyAsync(S0, w.S, sizeof(S0), hipMemcpyDeviceToHost, h->stream));
  VMI_HIP(hipStreamSynchronize(h->stream));
  const double rr = S0[S_TS], ff = S0[S_TT];
  *rnorm = std::sqrt(rr);
  *rel = ff > 0.0 ? std::sqrt(rr / ff) : std::sqrt(rr);
  return 0;
}

int create_handle(const vibe_mi_helmholtz_desc *ds, bool fp32, vibe_mi_helmholtz **out);

// factors the line preconditioner of w64's a, b and (fp32) copies the
// coefficients and the factorisation to w32; synchronous
int factor(vibe_mi_helmholtz *h) {
  Work<double> &w = h->w64;
  hipLaunchKernelGGL(k_line_factor, dim3(ceil_div(h->nx, 256), h->ny), dim3(256), 0, h->stream, op_of(h, w), w.lo,
                     w.cp, w.inv);
  VMI_HIP(hipGetLastError());
  if (h->w32.a != nullptr) {
    const unsigned g = h->vgrid;
    const size_t n = h->n;
    hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(256), 0, h->stream, n, 1.0, w.a, h->w32.a, 0);
    hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(256), 0, h->stream, n, 1.0, w.b, h->w32.b, 0);
    hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(256), 0, h->stream, n, 1.0, w.lo, h->w32.lo, 0);
    hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(256), 0, h->stream, n, 1.0, w.cp, h->w32.cp, 0);
    hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(256), 0, h->stream, n, 1.0, w.inv, h->w32.inv, 0);
    VMI_HIP(hipGetLastError());
  }
  VMI_HIP(hipStreamSynchronize(h->stream));
  h->coeffs = true;
  h->coarse_ready = false;
  return 0;
}

// the coarse operator of the two-grid start: (nx+1)/2 x (ny+1)/2 x nz cells
// of spacing 2 dx, 2 dy (coarse cell i at fine cell 2i), the same levels,
// terrain, a and b restricted from the fine ones
int build_coarse(vibe_mi_helmholtz *h) {
  if (h->coarse == nullptr) {
    vibe_mi_helmholtz_desc cd{};
    cd.nx = (h->nx + 1) / 2;
    cd.ny = (h->ny + 1) / 2;
    cd.nz = h->nz;
    cd.dx = 2.0 * h->dx;
    cd.dy = 2.0 * h->dy;
    cd.ztop = h->ztop;
    if (int rc = create_handle(&cd, false, &h->coarse)) return rc;
  }
  vibe_mi_helmholtz *c = h->coarse;
  VMI_HIP(hipStreamSynchronize(h->stream));
  VMI_HIP(hipMemcpyAsync(c->w64.dzeta, h->w64.dzeta, h->nz * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (int rc = mg_restrict(h->w64.zs, h->nx, h->ny, 1, c->w64.zs, c->nx, c->ny, 1, c->stream)) return rc;
  if (int rc = mg_restrict(h->w64.a, h->nx, h->ny, h->nz, c->w64.a, c->nx, c->ny, c->nz, c->stream)) return rc;
  if (int rc = mg_restrict(h->w64.b, h->nx, h->ny, h->nz, c->w64.b, c->nx, c->ny, c->nz, c->stream)) return rc;
  if (int rc = factor(c)) return rc;
  h->coarse_ready = true;
  return 0;
}

// Solves A x = f64 into w64.x; iterations and the final FP64 relative
// residual in *total / *rel.
int solve_core(vibe_mi_helmholtz *h, int kind, double tol, int max_iter, int *total_out, double *rel_out) {
  const size_t n = h->n;
  const unsigned g = h->vgrid;
  Work<double> &w = h->w64;
  int total = 0;
  double rel = 0.0, rn = 0.0;
  if (kind == VIBE_MI_HELM_BICGSTAB) {
    VMI_HIP(hipMemcpyAsync(w.r, h->f64, n * 8, hipMemcpyDeviceToDevice, h->stream));
    if (int rc = bicgstab(h, w, tol, max_iter, &total)) return rc;
    if (int rc = true_residual(h, &rel, &rn)) return rc;
  } else if (kind == VIBE_MI_HELM_MULTIGRID) {
    // two-grid start: restrict f, solve the coarse problem to the looser of
    // tol and 1e-3, prolong it as the initial guess of the fine solve
    if (!h->coarse_ready)
      if (int rc = build_coarse(h)) return rc;
    vibe_mi_helmholtz *c = h->coarse;
    VMI_HIP(hipStreamSynchronize(h->stream));
    if (int rc = mg_restrict(h->f64, h->nx, h->ny, h->nz, c->w64.r, c->nx, c->ny, c->nz, c->stream)) return rc;
    int its_c = 0;
    if (int rc = bicgstab(c, c->w64, std::max(tol, 1e-3), max_iter, &its_c)) return rc;
    if (int rc = mg_prolong(c->w64.x, c->nx, c->ny, c->nz, w.x, h->nx, h->ny, h->nz, c->stream)) return rc;
    VMI_HIP(hipStreamSynchronize(c->stream));
    VMI_HIP(hipMemcpyAsync(w.r, h->f64, n * 8, hipMemcpyDeviceToDevice, h->stream));
    if (int rc = bicgstab(h, w, tol, max_iter, &total, true)) return rc;
    if (int rc = true_residual(h, &rel, &rn)) return rc;
  } else {
    // iterative refinement: FP32 solves of A dx = r (r scaled to unit norm),
    // FP64 residuals and accumulation
    Work<float> &v = h->w32;
    hipLaunchKernelGGL((k_fill<double>), dim3(g), dim3(256), 0, h->stream, n, 0.0, w.x);
    VMI_HIP(hipGetLastError());
    const double tol_in = tol > 1e-5 ? tol : 1e-5;
    for (int outer = 0; outer < 30; ++outer) {
      if (int rc = true_residual(h, &rel, &rn)) return rc;
      if (rel < tol || total >= max_iter || !(rn > 0.0) || !std::isfinite(rel)) break;
      hipLaunchKernelGGL((k_axpy<float, double>), dim3(g), dim3(256), 0, h->stream, n, 1.0 / rn, w.r, v.r, 0);
      VMI_HIP(hipGetLastError());
      int its = 0;
      if (int rc = bicgstab(h, v, tol_in, max_iter - total, &its)) return rc;
      total += its;
      hipLaunchKernelGGL((k_axpy<double, float>), dim3(g), dim3(256), 0, h->stream, n, rn, v.x, w.x, 1);
      VMI_HIP(hipGetLastError());
      if (its == 0) break;
    }
    if (int rc = true_residual(h, &rel, &rn)) return rc;
  }
  *total_out = total;
  *rel_out = rel;
  return 0;
}

int create_handle(const vibe_mi_helmholtz_desc *ds, bool fp32, vibe_mi_helmholtz **out) {
  auto *h = new vibe_mi_helmholtz();
  h->nx = ds->nx;
  h->ny = ds->ny;
  h->nz = ds->nz;
  h->n = static_cast<size_t>(ds->nx) * ds->ny * ds->nz;
  h->dx = ds->dx;
  h->dy = ds->dy;
  h->ztop = ds->ztop;
  auto fail = [&](int rc) {
    vibe_mi_helmholtz_destroy(h);
    return rc;
  };
  if (hipGetDevice(&h->device) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  if (int rc = alloc_work(h, h->w64)) return fail(rc);
  if (fp32)
    if (int rc = alloc_work(h, h->w32)) return fail(rc);
  if (int rc = halloc(h, &h->f64, h->n)) return fail(rc);
  h->vgrid = static_cast<unsigned>(std::min<size_t>(ceil_div(h->n, 256), 8192));
  const size_t nparts = std::max<size_t>({static_cast<size_t>(kDotBlocks), apply_blocks(h), h->vgrid});
  if (int rc = halloc(h, &h->part, 2 * nparts)) return fail(rc);
  if (ds->zeta != nullptr && ds->zs != nullptr) {
    const size_t nxy = static_cast<size_t>(h->nx) * h->ny;
    std::vector<double> dz(h->nz);
    for (int k = 0; k < h->nz; ++k) dz[k] = ds->zeta[k + 1] - ds->zeta[k];
    std::vector<float> dz32(dz.begin(), dz.end()), zs32(ds->zs, ds->zs + nxy);
    if (hipMemcpy(h->w64.dzeta, dz.data(), dz.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->w64.zs, ds->zs, nxy * 8, hipMemcpyHostToDevice) != hipSuccess)
      return fail(VIBE_MI_ERR_HIP);
    if (fp32 && (hipMemcpy(h->w32.dzeta, dz32.data(), dz32.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                 hipMemcpy(h->w32.zs, zs32.data(), nxy * 4, hipMemcpyHostToDevice) != hipSuccess))
      return fail(VIBE_MI_ERR_HIP);
  }
  *out = h;
  return 0;
}

bool valid_kind(int kind) {
  return kind == VIBE_MI_HELM_BICGSTAB || kind == VIBE_MI_HELM_REFINE || kind == VIBE_MI_HELM_MULTIGRID;
}

}  // namespace

// Device-array forms for in-library callers (helmholtz_dev.hpp).
int helmholtz_set_coefficients_device(vibe_mi_helmholtz *h, const double *a, const double *b) {
  if (h == nullptr || a == nullptr || b == nullptr) return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(h->device));
  VMI_HIP(hipMemcpyAsync(h->w64.a, a, h->n * 8, hipMemcpyDeviceToDevice, h->stream));
  VMI_HIP(hipMemcpyAsync(h->w64.b, b, h->n * 8, hipMemcpyDeviceToDevice, h->stream));
  return factor(h);
}

int helmholtz_solve_device(vibe_mi_helmholtz *h, const double *f, double *x, int kind, double tol, int max_iter,
                           int *iterations, double *rel_res) {
  if (h == nullptr || f == nullptr || x == nullptr || !h->coeffs || !(tol > 0.0) || max_iter < 0 || !valid_kind(kind))
    return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(h->device));
  VMI_HIP(hipMemcpyAsync(h->f64, f, h->n * 8, hipMemcpyDeviceToDevice, h->stream));
  int total = 0;
  double rel = 0.0;
  if (int rc = solve_core(h, kind, tol, max_iter, &total, &rel)) return rc;
  VMI_HIP(hipMemcpyAsync(x, h->w64.x, h->n * 8, hipMemcpyDeviceToDevice, h->stream));
  VMI_HIP(hipStreamSynchronize(h->stream));
  if (iterations) *iterations = total;
  if (rel_res) *rel_res = rel;
  return 0;
}

}  // namespace vibe_mi

extern "C" {

int vibe_mi_helmholtz_create(const vibe_mi_helmholtz_desc *ds, vibe_mi_helmholtz **out) {
  if (ds == nullptr || out == nullptr || ds->nx <= 0 || ds->ny <= 0 || ds->nz <= 0 || !(ds->dx > 0) ||
      !(ds->dy > 0) || ds->zeta == nullptr || ds->zs == nullptr || ds->ny > 65535 || ds->nz > 65535)
    return VIBE_C_ERR_BAD_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return VIBE_MI_ERR_NO_DEVICE;
  return create_handle(ds, true, out);
}

int vibe_mi_helmholtz_destroy(vibe_mi_helmholtz *h) {
  if (h == nullptr) return 0;
  if (h->coarse != nullptr) vibe_mi_helmholtz_destroy(h->coarse);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int vibe_mi_helmholtz_set_coefficients(vibe_mi_helmholtz *h, const double *a, const double *b) {
  if (h == nullptr || a == nullptr || b == nullptr) return VIBE_C_ERR_BAD_ARG;
  for (size_t q = 0; q < h->n; ++q)
    if (!(b[q] > 0.0) || !(a[q] >= 0.0)) return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(h->device));
  const size_t n = h->n;
  VMI_HIP(hipMemcpy(h->w64.a, a, n * 8, hipMemcpyHostToDevice));
  VMI_HIP(hipMemcpy(h->w64.b, b, n * 8, hipMemcpyHostToDevice));
  return factor(h);
}

int vibe_mi_helmholtz_apply(vibe_mi_helmholtz *h, const double *x, double *y) {
  if (h == nullptr || x == nullptr || y == nullptr || !h->coeffs) return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(h->device));
  Work<double> &w = h->w64;
  VMI_HIP(hipMemcpyAsync(w.x, x, h->n * 8, hipMemcpyHostToDevice, h->stream));
  if (int rc = apply(h, w, static_cast<const double *>(w.x), static_cast<const double *>(nullptr), w.v, 0)) return rc;
  VMI_HIP(hipMemcpyAsync(y, w.v, h->n * 8, hipMemcpyDeviceToHost, h->stream));
  VMI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int vibe_mi_helmholtz_residual(vibe_mi_helmholtz *h, const double *f, const double *x, double *res_l2) {
  if (h == nullptr || f == nullptr || x == nullptr || res_l2 == nullptr || !h->coeffs) return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(h->device));
  VMI_HIP(hipMemcpyAsync(h->f64, f, h->n * 8, hipMemcpyHostToDevice, h->stream));
  VMI_HIP(hipMemcpyAsync(h->w64.x, x, h->n * 8, hipMemcpyHostToDevice, h->stream));
  double rel = 0.0, rn = 0.0;
  if (int rc = true_residual(h, &rel, &rn)) return rc;
  *res_l2 = rn;
  return 0;
}

int vibe_mi_helmholtz_solve(vibe_mi_helmholtz *h, const double *f, double *x, int kind, double tol, int max_iter,
                            int *iterations, double *rel_res) {
  if (h == nullptr || f == nullptr || x == nullptr || !h->coeffs || !(tol > 0.0) || max_iter < 0 || !valid_kind(kind))
    return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(h->device));
  VMI_HIP(hipMemcpyAsync(h->f64, f, h->n * 8, hipMemcpyHostToDevice, h->stream));
  int total = 0;
  double rel = 0.0;
  if (int rc = solve_core(h, kind, tol, max_iter, &total, &rel)) return rc;
  VMI_HIP(hipMemcpyAsync(x, h->w64.x, h->n * 8, hipMemcpyDeviceToHost, h->stream));
  VMI_HIP(hipStreamSynchronize(h->stream));
  if (iterations) *iterations = total;
  if (rel_res) *rel_res = rel;
  return 0;
}

int vibe_mi_multigrid_restrict_f64(const double *fine, int fnx, int fny, int fnz, double *coarse, int cnx, int cny,
                                   int cnz, void *stream) {
  if (fine == nullptr || coarse == nullptr || fnx <= 0 || fny <= 0 || fnz <= 0 || cnx <= 0 || cny <= 0 || cnz <= 0 ||
      cny > 65535 || cnz > 65535)
    return VIBE_C_ERR_BAD_ARG;
  return mg_restrict(fine, fnx, fny, fnz, coarse, cnx, cny, cnz, static_cast<hipStream_t>(stream));
}

int vibe_mi_multigrid_prolong_f64(const double *coarse, int cnx, int cny, int cnz, double *fine, int fnx, int fny,
                                  int fnz, void *stream) {
  if (fine == nullptr || coarse == nullptr || fnx <= 0 || fny <= 0 || fnz <= 0 || cnx <= 0 || cny <= 0 || cnz <= 0 ||
      fny > 65535 || fnz > 65535)
    return VIBE_C_ERR_BAD_ARG;
  return mg_prolong(coarse, cnx, cny, cnz, fine, fnx, fny, fnz, static_cast<hipStream_t>(stream));
}

}  // extern "C"
