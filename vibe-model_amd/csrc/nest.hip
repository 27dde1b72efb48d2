// nest.hip -- nest boundaries on the device: bilinear parent -> child
// interpolation, Davies relaxation of the child's lateral zone and the
// child -> parent feedback (src/grid/nest.cpp:34-128 roles), for domains on
// the same GPU.
//
// One thread per child point (i, j, k) loops over the six prognostics (w has
// one level more).  The relaxation touches only the zone columns (a thin
// frame), so it is a small launch next to a step; interior threads exit at
// once.
// Arithmetic follows c_kernels.c:108-144 (lerp form a + w (b - a), the same
// association) in FP64 on the FP64 states.
#include <cmath>

#include "domain.hpp"
#include "vibe_mi/nest.h"

namespace vibe_mi {
namespace {

constexpr int kMaxZone = 16;

struct NestArgs {
  const double *pn[6];  // parent, current state
  const double *po[6];  // parent, state before its last step
  double *c[6];         // child, current state
  int pnx, pny, cnx, cny, nz;
  int r, i0, j0, zone;
  double alpha[kMaxZone];  // Davies weights [zone]
  double frac;          // time weight of the new parent state
  int full;             // 1: overwrite the whole child with the parent state
};

__device__ __forceinline__ int wrap_mod(int i, int n) {
  const int m = i % n;
  return m < 0 ? m + n : m;
}

// Bilinear value of level-plane f (nx x ny, periodic) at index point (xp, yp).
__device__ __forceinline__ double bilinear(const double *__restrict__ f, int nx, int ny, double xp, double yp) {
  const double fx = floor(xp), fy = floor(yp);
  const int ia = static_cast<int>(fx), ja = static_cast<int>(fy);
  const double wx = xp - fx, wy = yp - fy;
  const int i0 = wrap_mod(ia, nx), i1 = wrap_mod(ia + 1, nx);
  const size_t r0 = static_cast<size_t>(wrap_mod(ja, ny)) * nx, r1 = static_cast<size_t>(wrap_mod(ja + 1, ny)) * nx;
  const double q00 = f[r0 + i0], q10 = f[r0 + i1], q01 = f[r1 + i0], q11 = f[r1 + i1];
  return (q00 + wx * (q10 - q00)) + wy * ((q01 + wx * (q11 - q01)) - (q00 + wx * (q10 - q00)));
}

__global__ void __launch_bounds__(256) k_nest(NestArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= a.cnx) return;
  const int d = min(min(i, a.cnx - 1 - i), min(j, a.cny - 1 - j));
  double al = 1.0;
  if (!a.full) {
    if (d >= a.zone) return;
    al = a.alpha[d];
  }
  const double rinv = 1.0 / static_cast<double>(a.r);
  const double xc = (static_cast<double>(a.i0 + i) + 0.5) * rinv - 0.5;
  const double yc = (static_cast<double>(a.j0 + j) + 0.5) * rinv - 0.5;
  const double xu = static_cast<double>(a.i0 + i) * rinv;  // u: west face
  const double yv = static_cast<double>(a.j0 + j) * rinv;  // v: south face
  const size_t pxy = static_cast<size_t>(a.pnx) * a.pny, cxy = static_cast<size_t>(a.cnx) * a.cny;
  const size_t cc = static_cast<size_t>(j) * a.cnx + i;
  for (int f = 0; f < 6; ++f) {
    const double xp = f == 0 ? xu : xc;
    const double yp = f == 1 ? yv : yc;
    const int nk = f == 2 ? a.nz + 1 : a.nz;
    if (k < nk) {
      const double pnew = bilinear(a.pn[f] + k * pxy, a.pnx, a.pny, xp, yp);
      double p = pnew;
      if (!a.full && a.frac != 1.0) {
        const double pold = bilinear(a.po[f] + k * pxy, a.pnx, a.pny, xp, yp);
        p = pold + a.frac * (pnew - pold);
      }
      double *cf = a.c[f] + k * cxy + cc;
      *cf = a.full ? p : (1.0 - al) * *cf + al * p;
    }
  }
}

// Two-way feedback: one thread per covered parent cell (I0 + i, J0 + j) and
// level.  Cell fields and w take the mean of the r x r child cells over the
// parent cell (accumulated jj-outer, ii-inner, times 1/r^2:
// nest.cpp:85-108); u (west face) the mean of the r child faces on the
// parent face, v likewise.
struct RestrictArgs {
  const double *c[6];  // child, current state
  double *p[6];        // parent, current state
  int pnx, pny, cnx, cny, nz, r, I0, J0, iext, jext;
};

__global__ void __launch_bounds__(64) k_nest_restrict(RestrictArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= a.iext) return;
  const size_t pxy = static_cast<size_t>(a.pnx) * a.pny, cxy = static_cast<size_t>(a.cnx) * a.cny;
  const int r = a.r;
  const double w2 = 1.0 / static_cast<double>(r * r), w1 = 1.0 / static_cast<double>(r);
  const size_t pc = static_cast<size_t>(wrap_mod(a.J0 + j, a.pny)) * a.pnx + wrap_mod(a.I0 + i, a.pnx);
  for (int f = 0; f < 6; ++f) {
    if (k >= (f == 2 ? a.nz + 1 : a.nz)) continue;
    const double *__restrict__ src = a.c[f] + static_cast<size_t>(k) * cxy;
    double acc = 0.0, wt = w2;
    if (f == 0) {
      for (int jj = 0; jj < r; ++jj) acc += src[static_cast<size_t>(j * r + jj) * a.cnx + i * r];
      wt = w1;
    } else if (f == 1) {
      for (int ii = 0; ii < r; ++ii) acc += src[static_cast<size_t>(j * r) * a.cnx + i * r + ii];
      wt = w1;
    } else {
      for (int jj = 0; jj < r; ++jj)
        for (int ii = 0; ii < r; ++ii) acc += src[static_cast<size_t>(j * r + jj) * a.cnx + i * r + ii];
    }
    a.p[f][static_cast<size_t>(k) * pxy + pc] = acc * wt;
  }
}

}  // namespace

int nest_restrict_enqueue(const vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd,
                          hipStream_t stream) {
  if (child == nullptr || parent == nullptr || nd == nullptr || child == parent) return VIBE_C_ERR_BAD_ARG;
  if (nd->ratio < 2 || child->nz != parent->nz || nd->i0 % nd->ratio != 0 || nd->j0 % nd->ratio != 0 ||
      child->nx < nd->ratio || child->ny < nd->ratio)
    return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(child->device));
  const FieldSet &cs = child->S[child->cur0];
  const FieldSet &ps = parent->S[parent->cur0];
  RestrictArgs a{};
  const double *cf[6] = {cs.u, cs.v, cs.w, cs.rho, cs.th, cs.pi};
  double *pf[6] = {ps.u, ps.v, ps.w, ps.rho, ps.th, ps.pi};
  for (int f = 0; f < 6; ++f) {
    a.c[f] = cf[f];
    a.p[f] = pf[f];
  }
  a.pnx = parent->nx;
  a.pny = parent->ny;
  a.cnx = child->nx;
  a.cny = child->ny;
  a.nz = child->nz;
  a.r = nd->ratio;
  a.I0 = nd->i0 / nd->ratio;
  a.J0 = nd->j0 / nd->ratio;
  a.iext = child->nx / nd->ratio;
  a.jext = child->ny / nd->ratio;
  hipLaunchKernelGGL(k_nest_restrict, dim3(ceil_div(a.iext, 64), a.jext, a.nz + 1), dim3(64), 0, stream, a);
  VMI_HIP(hipGetLastError());
  return 0;
}

int nest_enqueue(vibe_mi_domain *child, const vibe_mi_domain *parent, const vibe_mi_nest_desc *nd, double frac,
                 int full, hipStream_t stream) {
  if (child == nullptr || parent == nullptr || nd == nullptr || child == parent) return VIBE_C_ERR_BAD_ARG;
  if (nd->ratio < 1 || nd->zone < 1 || nd->zone > kMaxZone || child->nz != parent->nz ||
      !(frac >= 0.0 && frac <= 1.0))
    return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(child->device));
  const FieldSet &pn = parent->S[parent->cur0];
  const FieldSet &po = parent->S[(parent->cur0 + 2) % 3];
  const FieldSet &cs = child->S[child->cur0];
  NestArgs a{};
  if (int rc = vibe_mi_davies_profile(nd->zone, nd->alpha_max, a.alpha)) return rc;
  const double *pnf[6] = {pn.u, pn.v, pn.w, pn.rho, pn.th, pn.pi};
  const double *pof[6] = {po.u, po.v, po.w, po.rho, po.th, po.pi};
  double *cf[6] = {cs.u, cs.v, cs.w, cs.rho, cs.th, cs.pi};
  for (int f = 0; f < 6; ++f) {
    a.pn[f] = pnf[f];
    a.po[f] = pof[f];
    a.c[f] = cf[f];
  }
  a.pnx = parent->nx;
  a.pny = parent->ny;
  a.cnx = child->nx;
  a.cny = child->ny;
  a.nz = child->nz;
  a.r = nd->ratio;
  a.i0 = nd->i0;
  a.j0 = nd->j0;
  a.zone = nd->zone;
  a.frac = frac;
  a.full = full;
  hipLaunchKernelGGL(k_nest, dim3(ceil_div(child->nx, 256), child->ny, child->nz + 1), dim3(256), 0, stream, a);
  VMI_HIP(hipGetLastError());
  return 0;
}

namespace {

int nest_launch(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd, double frac, int full) {
  if (child == nullptr || parent == nullptr) return VIBE_C_ERR_BAD_ARG;
  VMI_HIP(hipSetDevice(child->device));
  VMI_HIP(hipStreamSynchronize(parent->stream));
  if (int rc = nest_enqueue(child, parent, nd, frac, full, child->stream)) return rc;
  VMI_HIP(hipStreamSynchronize(child->stream));
  return 0;
}

}  // namespace
}  // namespace vibe_mi

using namespace vibe_mi;

extern "C" {

int vibe_mi_davies_profile(int zone, double alpha_max, double *alpha) {
  if (zone < 1 || alpha == nullptr) return VIBE_C_ERR_BAD_ARG;
  const double pi = 3.14159265358979323846;
  for (int n = 0; n < zone; ++n)
    alpha[n] = alpha_max * 0.5 * (1.0 - std::cos(pi * static_cast<double>(zone - n) / static_cast<double>(zone)));
  return 0;
}

int vibe_mi_nest_interpolate(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd) {
  const int rc = nest_launch(child, parent, nd, 1.0, 1);
  if (rc) set_last_error(rc);
  return rc;
}

int vibe_mi_nest_relax(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd, double frac) {
  const int rc = nest_launch(child, parent, nd, frac, 0);
  if (rc) set_last_error(rc);
  return rc;
}

int vibe_mi_nest_restrict(vibe_mi_domain *child, vibe_mi_domain *parent, const vibe_mi_nest_desc *nd) {
  if (child == nullptr || parent == nullptr) return VIBE_C_ERR_BAD_ARG;
  int rc = 0;
  if (hipSetDevice(parent->device) != hipSuccess || hipStreamSynchronize(child->stream) != hipSuccess) {
    rc = VIBE_MI_ERR_HIP;
  } else if ((rc = nest_restrict_enqueue(child, parent, nd, parent->stream)) == 0) {
    if (hipStreamSynchronize(parent->stream) != hipSuccess) rc = VIBE_MI_ERR_HIP;
  }
  if (rc) set_last_error(rc);
  return rc;
}

}  // extern "C"
