// nest.hip -- one-way nest lateral boundary on the device: bilinear
// parent -> child interpolation and Davies relaxation of the child's lateral
// zone (src/grid/nest.cpp:34-128 roles), for domains on the same GPU.
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

}  // namespace

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

}  // extern "C"
