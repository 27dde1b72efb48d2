// setup.cpp -- host-side input generators for the dynamical-core step:
// vertical levels, terrain, hydrostatic reference state, sponge profile and
// idealised initial perturbations.  They build the arrays a caller uploads
// with vibe_mi_domain_create / vibe_mi_domain_set_state; nothing here runs
// inside the time step.
//
// Reference roles: make_stretched_zeta (src/grid/geometry.cpp:39-60),
// witch_of_agnesi (geometry.cpp:103-117), ReferenceState::isothermal /
// from_profile (src/dyn/reference_state.cpp:26-110), Damping's sponge table
// (src/dyn/damping.cpp:24-41), init_warm_bubble / add_random_noise /
// apply_shear / enforce_state_equation (src/driver/initial_conditions.cpp:
// 94-159).  Random numbers come from SplitMix64 (explicit, portable), not
// libstdc++'s mt19937, so inputs are reproducible on every host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "vibe_mi/dycore.h"

namespace {

constexpr double kG = 9.80665, kRd = 287.05, kCp = 1004.64, kCv = 717.63, kP0 = 100000.0;
constexpr double kPi = 3.14159265358979323846;

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

inline double zc_of(const double *zeta, int k) { return 0.5 * (zeta[k] + zeta[k + 1]); }

}  // namespace

extern "C" {

int vibe_mi_zeta_uniform(int nz, double *zeta) {
  if (nz <= 0 || zeta == nullptr) return 2;
  for (int k = 0; k <= nz; ++k) zeta[k] = static_cast<double>(k) / static_cast<double>(nz);
  zeta[nz] = 1.0;
  return 0;
}

int vibe_mi_zeta_tanh(int nz, double ztop, double dz_bottom, double dz_top, double *zeta) {
  if (nz <= 0 || zeta == nullptr || ztop <= 0.0 || dz_bottom <= 0.0 || dz_top < dz_bottom) return 2;
  const double mean = ztop / nz;
  if (mean <= dz_bottom || mean >= dz_top) return 2;
  const double width = std::max(1.0, nz / 8.0);
  auto total = [&](double k0) {
    double s = 0.0;
    for (int k = 0; k < nz; ++k)
      s += dz_bottom + (dz_top - dz_bottom) * 0.5 * (1.0 + std::tanh((k + 0.5 - k0) / width));
    return s;
  };
  // total() decreases with k0: bisect for total(k0) == ztop.
  double lo = -4.0 * nz, hi = 5.0 * nz;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (total(mid) > ztop) lo = mid; else hi = mid;
  }
  const double k0 = 0.5 * (lo + hi);
  std::vector<double> dz(nz);
  double s = 0.0;
  for (int k = 0; k < nz; ++k) {
    dz[k] = dz_bottom + (dz_top - dz_bottom) * 0.5 * (1.0 + std::tanh((k + 0.5 - k0) / width));
    s += dz[k];
  }
  zeta[0] = 0.0;
  double acc = 0.0;
  for (int k = 0; k < nz; ++k) {
    acc += dz[k];
    zeta[k + 1] = acc / s;
  }
  zeta[nz] = 1.0;
  return 0;
}

int vibe_mi_terrain_agnesi_ridge(int nx, int ny, double dx, double h, double a, double *zs) {
  if (nx <= 0 || ny <= 0 || zs == nullptr) return 2;
  const double xc = 0.5 * nx * dx;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double x = (i + 0.5) * dx - xc;
      zs[static_cast<size_t>(j) * nx + i] = h * a * a / (a * a + x * x);
    }
  return 0;
}

int vibe_mi_zeta_stretched(int nz, double first_thickness, double stretch, double *zeta) {
  if (nz <= 0 || zeta == nullptr || !(first_thickness > 0.0) || !(stretch > 0.0)) return 2;
  // thicknesses h, h s, h s^2, ... (running product, as accumulated in
  // geometry.cpp:46-51), interfaces = running sums / total
  std::vector<double> dz(nz);
  double total = 0.0, h = first_thickness;
  for (int k = 0; k < nz; ++k) {
    dz[k] = h;
    total += h;
    h *= stretch;
  }
  zeta[0] = 0.0;
  double acc = 0.0;
  for (int k = 0; k < nz; ++k) {
    acc += dz[k];
    zeta[k + 1] = acc / total;
  }
  zeta[nz] = 1.0;
  return 0;
}

int vibe_mi_terrain_agnesi_bell(int nx, int ny, double dx, double dy, double h0, double a, double xc, double yc,
                                double *zs) {
  if (nx <= 0 || ny <= 0 || zs == nullptr) return 2;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double x = static_cast<double>(i) * dx - xc, y = static_cast<double>(j) * dy - yc;
      zs[static_cast<size_t>(j) * nx + i] = h0 * a * a / (a * a + (x * x + y * y));
    }
  return 0;
}

int vibe_mi_terrain_smooth(double *zs, int nx, int ny, int passes, double max_slope) {
  if (zs == nullptr || nx <= 0 || ny <= 0 || passes < 0) return 2;
  const size_t sx = static_cast<size_t>(nx);
  std::vector<double> f(static_cast<size_t>(nx) * ny);
  for (int p = 0; p < passes; ++p) {
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        const size_t c = static_cast<size_t>(j) * sx + i;
        f[c] = 0.25 * (zs[c - 1] + zs[c + 1] + zs[c - sx] + zs[c + sx]) + 0.5 * zs[c];
      }
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        const size_t c = static_cast<size_t>(j) * sx + i;
        zs[c] = std::min(zs[c], f[c]);
      }
    if (max_slope > 0.0)
      for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
          const size_t c = static_cast<size_t>(j) * sx + i;
          const double top = std::max(std::max(zs[c - 1], zs[c + 1]), std::max(zs[c - sx], zs[c + sx]));
          zs[c] = std::min(zs[c], top + max_slope);
        }
  }
  return 0;
}

int vibe_mi_terrain_gaussian(int nx, int ny, double dx, double dy, double sigma, double L,
                             uint64_t seed, double *zs) {
  if (nx <= 0 || ny <= 0 || zs == nullptr || L <= 0.0) return 2;
  SplitMix64 rng{seed};
  const size_t n = static_cast<size_t>(nx) * ny;
  std::vector<double> a(n), b(n);
  for (size_t q = 0; q < n; ++q) {  // Box-Muller
    const double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
    a[q] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
  }
  // separable periodic Gaussian filter exp(-r^2 / L^2)
  auto kernel = [&](double h, int m) {
    const int r = std::min(m / 2, static_cast<int>(std::ceil(3.0 * L / h)));
    std::vector<double> w(2 * r + 1);
    for (int q = -r; q <= r; ++q) w[q + r] = std::exp(-(q * h) * (q * h) / (L * L));
    return w;
  };
  const auto wx = kernel(dx, nx), wy = kernel(dy, ny);
  const int rx = static_cast<int>(wx.size() / 2), ry = static_cast<int>(wy.size() / 2);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double s = 0.0;
      for (int q = -rx; q <= rx; ++q) s += wx[q + rx] * a[static_cast<size_t>(j) * nx + ((i + q) % nx + nx) % nx];
      b[static_cast<size_t>(j) * nx + i] = s;
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double s = 0.0;
      for (int q = -ry; q <= ry; ++q) s += wy[q + ry] * b[static_cast<size_t>(((j + q) % ny + ny) % ny) * nx + i];
      a[static_cast<size_t>(j) * nx + i] = s;
    }
  double mean = 0.0, var = 0.0, mn = 1e300;
  for (size_t q = 0; q < n; ++q) mean += a[q];
  mean /= static_cast<double>(n);
  for (size_t q = 0; q < n; ++q) var += (a[q] - mean) * (a[q] - mean);
  const double sd = std::sqrt(var / static_cast<double>(n));
  for (size_t q = 0; q < n; ++q) {
    zs[q] = (sd > 0.0) ? sigma * (a[q] - mean) / sd : 0.0;
    mn = std::min(mn, zs[q]);
  }
  for (size_t q = 0; q < n; ++q) zs[q] -= mn;
  return 0;
}

int vibe_mi_reference_state(int kind, double theta_sfc, double bv, double p_sfc, int nx,
                            int ny, int nz, double ztop, const double *zeta, const double *zs,
                            double *rho0, double *theta0, double *pi0) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || zeta == nullptr || zs == nullptr || rho0 == nullptr ||
      theta0 == nullptr || pi0 == nullptr || (kind != 0 && kind != 1) || (kind == 1 && bv <= 0.0))
    return 2;
  const double pis = std::pow(p_sfc / kP0, kRd / kCp);  // reference_state.cpp:46
  const double n2g = bv * bv / kG;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t col = static_cast<size_t>(j) * nx + i;
        const size_t c = static_cast<size_t>(k) * ny * nx + col;
        const double z = zs[col] + zc_of(zeta, k) * (ztop - zs[col]);
        double th, pi;
        if (kind == 0) {  // reference_state.cpp:51
          th = theta_sfc;
          pi = pis - kG * z / (kCp * theta_sfc);
        } else {  // exact hydrostatic Exner for theta = theta_s exp(N^2 z / g)
          th = theta_sfc * std::exp(n2g * z);
          pi = pis + (kG * kG / (kCp * theta_sfc * bv * bv)) * (std::exp(-n2g * z) - 1.0);
        }
        theta0[c] = th;
        pi0[c] = pi;
        rho0[c] = kP0 * std::pow(pi, kCv / kRd) / (kRd * th);  // reference_state.cpp:26-28
      }
  return 0;
}

// ReferenceState::from_profile (reference_state.cpp:61-110), column by column:
// interface heights z_k = zs + zeta_k (ztop - zs) (Grid::z_interface); theta at
// an interface is the linear interpolant of the (z, theta) profile, held at
// its end values outside it (the profile segment is found by the same forward
// search); pi is integrated upward from pi_s = (p_s / p00)^(Rd/cp) by
// pi -= g dz / (cp theta_mid), theta_mid the mean of the two interface values.
// Cell k takes theta0 = theta_mid and pi0 = the integrated value at its UPPER
// interface (the reference stores pi after the decrement, so its pi0 sits half
// a cell above the cell centre; kept for parity), rho0 from the equation of
// state.  The arithmetic is the reference's, operation for operation.
int vibe_mi_reference_state_from_profile(int nx, int ny, int nz, double ztop, const double *zeta,
                                         const double *zs, int np, const double *zp, const double *thp,
                                         double p_sfc, double *rho0, double *theta0, double *pi0) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || np < 2 || zeta == nullptr || zs == nullptr || zp == nullptr ||
      thp == nullptr || rho0 == nullptr || theta0 == nullptr || pi0 == nullptr)
    return 2;
  const size_t n = static_cast<size_t>(np);
  auto theta_at = [&](double z) {
    if (z <= zp[0]) return thp[0];
    if (z >= zp[n - 1]) return thp[n - 1];
    size_t k = 0;
    while (k + 2 < n && zp[k + 1] < z) ++k;
    const double w = (zp[k + 1] - zp[k] > 0.0) ? (z - zp[k]) / (zp[k + 1] - zp[k]) : 0.0;
    return thp[k] + w * (thp[k + 1] - thp[k]);
  };
  const double pis = std::pow(p_sfc / kP0, kRd / kCp);
  const size_t nxy = static_cast<size_t>(nx) * ny;
  std::vector<double> zi(static_cast<size_t>(nz) + 1);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t col = static_cast<size_t>(j) * nx + i;
      for (int k = 0; k <= nz; ++k) zi[k] = zs[col] + zeta[k] * (ztop - zs[col]);
      double pi = pis;
      for (int k = 0; k < nz; ++k) {
        const double th_lo = theta_at(zi[k]), th_hi = theta_at(zi[k + 1]);
        const double th_mid = 0.5 * (th_lo + th_hi);
        pi -= kG * (zi[k + 1] - zi[k]) / (kCp * th_mid);
        const size_t c = static_cast<size_t>(k) * nxy + col;
        theta0[c] = th_mid;
        pi0[c] = pi;
        rho0[c] = kP0 * std::pow(pi, kCv / kRd) / (kRd * th_mid);
      }
    }
  return 0;
}

int vibe_mi_sponge(int nz, double ztop, const double *zeta, double zs00, double start_fraction,
                   double alpha_max, double *alpha) {
  if (nz <= 0 || zeta == nullptr || alpha == nullptr) return 2;
  const double z0 = start_fraction * ztop;
  for (int k = 0; k < nz; ++k) {
    const double z = zs00 + zc_of(zeta, k) * (ztop - zs00);
    if (z <= z0) {
      alpha[k] = 0.0;
      continue;
    }
    const double t = std::clamp((z - z0) / std::max(ztop - z0, 1.0), 0.0, 1.0);
    const double s = std::sin(0.5 * kPi * t);
    alpha[k] = alpha_max * s * s;
  }
  return 0;
}

int vibe_mi_ic_bubble(int nx, int ny, int nz, double dx, double dy, double ztop,
                      const double *zeta, const double *zs, double xc, double yc, double zc,
                      double rx, double rz, double amp, double *theta) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || zeta == nullptr || zs == nullptr || theta == nullptr ||
      rx <= 0.0 || rz <= 0.0)
    return 2;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t col = static_cast<size_t>(j) * nx + i;
        const double x = (i + 0.5) * dx, y = (j + 0.5) * dy;
        const double z = zs[col] + zc_of(zeta, k) * (ztop - zs[col]);
        const double r2 = ((x - xc) / rx) * ((x - xc) / rx) + ((y - yc) / rx) * ((y - yc) / rx) +
                          ((z - zc) / rz) * ((z - zc) / rz);
        if (r2 < 1.0) {
          const double c = std::cos(0.5 * kPi * std::sqrt(r2));
          theta[static_cast<size_t>(k) * ny * nx + col] += amp * c * c;
        }
      }
  return 0;
}

int vibe_mi_ic_noise(int nx, int ny, int nz, double ztop, const double *zeta, const double *zs,
                     double amp, double zmax, uint64_t seed, double *theta) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || zeta == nullptr || zs == nullptr || theta == nullptr) return 2;
  SplitMix64 rng{seed};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t col = static_cast<size_t>(j) * nx + i;
        const double r = rng.uniform();  // drawn for every cell: layout-independent stream
        const double z = zs[col] + zc_of(zeta, k) * (ztop - zs[col]);
        if (z < zmax) theta[static_cast<size_t>(k) * ny * nx + col] += amp * (2.0 * r - 1.0);
      }
  return 0;
}

int vibe_mi_ic_wind(int nx, int ny, int nz, double ztop, const double *zeta, const double *zs,
                    double u0, double umax, double hscale, double *u) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || zeta == nullptr || zs == nullptr || u == nullptr) return 2;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t col = static_cast<size_t>(j) * nx + i;
        const double z = zs[col] + zc_of(zeta, k) * (ztop - zs[col]);
        u[static_cast<size_t>(k) * ny * nx + col] = u0 + (hscale > 0.0 ? umax * std::tanh(z / hscale) : 0.0);
      }
  return 0;
}

int vibe_mi_ic_balance_pi(size_t ncell, const double *rho0, const double *theta0,
                          const double *rho, const double *theta, double *pi) {
  if (rho0 == nullptr || theta0 == nullptr || rho == nullptr || theta == nullptr || pi == nullptr)
    return 2;
  for (size_t c = 0; c < ncell; ++c) {
    const double full = std::pow((rho0[c] + rho[c]) * kRd * (theta0[c] + theta[c]) / kP0, kRd / kCv);
    const double base = std::pow(rho0[c] * kRd * theta0[c] / kP0, kRd / kCv);
    pi[c] = full - base;
  }
  return 0;
}

}  // extern "C"
