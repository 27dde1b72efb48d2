// refstate.cpp -- conversion between the reference's prognostic State
// (include/vibe/dyn/state.hpp:53-56: 12 species, each a halo'd FP64 Field,
// include/vibe/grid/field.hpp:34-96) and this library's dense perturbation
// state (vibe_mi/dycore.h).  Host code: the reference-side adapter of an
// Integrator::step calls it before and after the device step.
//
// Reference field layout: nx x ny x nz interior points with h halo points on
// every side and one more face along the staggered axis (u: x, v: y, w: z),
// i fastest, offset(i, j, k) = ((k + h) nsy + (j + h)) nsx + (i + h) with
// nsx = nx + 2h (+1 for u), nsy = ny + 2h (+1 for v).  The dense state holds
// u on west faces i = 0..nx-1 and v on south faces j = 0..ny-1 (the domain is
// periodic: face nx is face 0), w on interfaces 0..nz, and rho', theta', pi'
// as deviations from the reference state.
#include <cstddef>
#include <cstdint>

#include "vibe_mi/dycore.h"

namespace {

struct RefLayout {
  int nx, ny, nz, h, sx, sy, sz;  // interior sizes, halo, stagger of the field (0/1 per axis)
  size_t nsx() const { return static_cast<size_t>(nx + 2 * h + sx); }
  size_t nsy() const { return static_cast<size_t>(ny + 2 * h + sy); }
  size_t nsz() const { return static_cast<size_t>(nz + 2 * h + sz); }
  size_t size() const { return nsx() * nsy() * nsz(); }
  size_t at(int i, int j, int k) const {
    return (static_cast<size_t>(k + h) * nsy() + static_cast<size_t>(j + h)) * nsx() + static_cast<size_t>(i + h);
  }
};

RefLayout layout_of(int nx, int ny, int nz, int h, int species) {
  return RefLayout{nx, ny, nz, h, species == 0 ? 1 : 0, species == 1 ? 1 : 0, species == 2 ? 1 : 0};
}

inline int wrap(int i, int n) {
  const int m = i % n;
  return m < 0 ? m + n : m;
}

}  // namespace

extern "C" {

size_t vibe_mi_reference_field_size(int nx, int ny, int nz, int halo, int species) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || halo < 0 || species < 0 || species >= 12) return 0;
  return layout_of(nx, ny, nz, halo, species).size();
}

int vibe_mi_state_from_reference(int nx, int ny, int nz, int halo, const double *const species[12],
                                 const double *rho0, const double *theta0, const double *pi0, double *u, double *v,
                                 double *w, double *rho_p, double *theta_p, double *pi_p) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || halo < 0 || species == nullptr || rho0 == nullptr || theta0 == nullptr ||
      pi0 == nullptr || u == nullptr || v == nullptr || w == nullptr || rho_p == nullptr || theta_p == nullptr ||
      pi_p == nullptr)
    return VIBE_C_ERR_BAD_ARG;
  for (int q = 0; q < 6; ++q)
    if (species[q] == nullptr) return VIBE_C_ERR_BAD_ARG;
  const RefLayout lu = layout_of(nx, ny, nz, halo, 0), lv = layout_of(nx, ny, nz, halo, 1),
                  lw = layout_of(nx, ny, nz, halo, 2), lc = layout_of(nx, ny, nz, halo, 3);
  const size_t nxy = static_cast<size_t>(nx) * ny;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t d = static_cast<size_t>(k) * nxy + static_cast<size_t>(j) * nx + i;
        w[d] = species[2][lw.at(i, j, k)];
        if (k == nz) continue;
        u[d] = species[0][lu.at(i, j, k)];
        v[d] = species[1][lv.at(i, j, k)];
        const size_t c = lc.at(i, j, k);
        rho_p[d] = species[3][c] - rho0[c];
        theta_p[d] = species[4][c] - theta0[c];
        pi_p[d] = species[5][c] - pi0[c];
      }
  return 0;
}

int vibe_mi_state_to_reference(int nx, int ny, int nz, int halo, const double *u, const double *v, const double *w,
                               const double *rho_p, const double *theta_p, const double *pi_p, const double *rho0,
                               const double *theta0, const double *pi0, int fill_halo, double *const species[12]) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || halo < 0 || species == nullptr || rho0 == nullptr || theta0 == nullptr ||
      pi0 == nullptr || u == nullptr || v == nullptr || w == nullptr || rho_p == nullptr || theta_p == nullptr ||
      pi_p == nullptr)
    return VIBE_C_ERR_BAD_ARG;
  for (int q = 0; q < 6; ++q)
    if (species[q] == nullptr) return VIBE_C_ERR_BAD_ARG;
  const RefLayout lu = layout_of(nx, ny, nz, halo, 0), lv = layout_of(nx, ny, nz, halo, 1),
                  lw = layout_of(nx, ny, nz, halo, 2), lc = layout_of(nx, ny, nz, halo, 3);
  const size_t nxy = static_cast<size_t>(nx) * ny;
  // every horizontal point of the written range maps to a periodic interior
  // point: the interior (and the duplicated faces u(nx), v(ny)), plus the
  // lateral halo when fill_halo is set (HaloExchange::exchange_local role,
  // src/grid/halo.cpp:58-86); vertical halo levels are left as they are
  const int hh = fill_halo ? halo : 0;
  for (int k = 0; k <= nz; ++k)
    for (int j = -hh; j < ny + hh + 1; ++j)
      for (int i = -hh; i < nx + hh + 1; ++i) {
        const int iw = wrap(i, nx), jw = wrap(j, ny);
        const size_t d = static_cast<size_t>(k) * nxy + static_cast<size_t>(jw) * nx + iw;
        const bool xin = i < nx + hh, yin = j < ny + hh;  // the cell/row range without the extra face
        if (xin && yin) species[2][lw.at(i, j, k)] = w[d];
        if (k == nz) continue;
        if (yin) species[0][lu.at(i, j, k)] = u[d];
        if (xin) species[1][lv.at(i, j, k)] = v[d];
        if (xin && yin) {
          // the reference state is defined on the interior (ReferenceState
          // fills no halo): take it at the periodic interior point
          const size_t c = lc.at(i, j, k), c0 = lc.at(iw, jw, k);
          species[3][c] = rho0[c0] + rho_p[d];
          species[4][c] = theta0[c0] + theta_p[d];
          species[5][c] = pi0[c0] + pi_p[d];
        }
      }
  return 0;
}

}  // extern "C"
