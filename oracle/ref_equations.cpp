// ref_equations.cpp -- golden vectors of the reference's literal discrete
// equations, computed by the REFERENCE's own code: Equations::tendencies,
// acoustic_tendencies, max_sound_speed and diagnose (src/dyn/equations.cpp,
// with src/dyn/advection.cpp in the same translation unit, and the unmodified
// coriolis, damping, tendency, state, reference_state and geometry sources),
// compiled from /root/reference by oracle/ref.mk.
//
// Test infrastructure: tests/golden/make_equations_golden.py runs it and
// commits the arrays it prints; tests/test_oracle_literal.py checks the
// oracle's literal restatement (oracle/literal.c) against them and
// tests/test_gpu_literal.py the device kernels.
//
// Inputs are integer hashes, so the tests rebuild them bit for bit:
//   r(i, j, k, s) = h / 2^20 - 0.5,
//   h = ((i+7)*73856093 ^ (j+11)*19349663 ^ (k+13)*83492791 ^ (s+3)*2654435761)
//       & 0xFFFFF                                   (unsigned 64-bit)
// over every point the reference's loops read (u faces 0..nx, v faces 0..ny,
// w faces 0..nz, cells); halos stay zero, as on the reference's single-rank
// path.  Fields (s = species index of state.hpp:53-56):
//   u = 8 + 6 r, v = -3 + 5 r, w = 0.8 r,
//   rho = rho0 + 0.02 r, theta = theta0 + 3 r, pi = pi0 + 0.002 r,
//   qv = max(0, 0.01 (r + 0.2)), q{c,r,i,s,g} = 0.002 (r + 0.1).
// Grids: 8 x 6 x 5, dx 1000 m, dy 1500 m, z_top 12 km, zeta =
// make_stretched_zeta(5, 1, 1.1); "hill": witch_of_agnesi(8, 6, 1000, 1500,
// 600, 2500, 4000, 4500) with a constant-N from_profile reference state
// (N = 0.01, 300 K, p_s 1000 hPa); "flat": no terrain, isothermal 300 K.
// Configurations (config.hpp:55-92 fields):
//   c2: hill, central2, f-plane 45 deg, no divergence damping, kessler
//   c4: hill, central4, no Coriolis, 2nd-order damping 2e5, microphysics none
//   c6: flat, central6, beta plane 30 deg, 4th-order damping 3e9, kessler
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "vibe/config/config.hpp"
#include "vibe/dyn/equations.hpp"
#include "vibe/dyn/reference_state.hpp"
#include "vibe/dyn/state.hpp"
#include "vibe/dyn/tendency.hpp"
#include "vibe/grid/geometry.hpp"

using vibe::Int;
using vibe::Real;
namespace grid = vibe::grid;
namespace dyn = vibe::dyn;

namespace {

bool g_first = true;

void put(const std::string &name, const std::vector<Real> &v) {
  std::printf("%s  \"%s\": [", g_first ? "" : ",\n", name.c_str());
  g_first = false;
  for (size_t q = 0; q < v.size(); ++q) std::printf("%s%.17g", q ? ", " : "", v[q]);
  std::printf("]");
}

Real hr(Int i, Int j, Int k, Int s) {
  const std::uint64_t h = ((static_cast<std::uint64_t>(i + 7) * 73856093ull) ^
                           (static_cast<std::uint64_t>(j + 11) * 19349663ull) ^
                           (static_cast<std::uint64_t>(k + 13) * 83492791ull) ^
                           (static_cast<std::uint64_t>(s + 3) * 2654435761ull)) &
                          0xFFFFFull;
  return static_cast<Real>(h) / 1048576.0 - 0.5;
}

constexpr Int NX = 8, NY = 6, NZ = 5;

// the dense layout of the tests: faces/levels of the field's stagger,
// i fastest
std::vector<Real> dump(const grid::Field<Real> &f, Int ex, Int ey, Int ez) {
  std::vector<Real> out;
  for (Int k = 0; k < NZ + ez; ++k)
    for (Int j = 0; j < NY + ey; ++j)
      for (Int i = 0; i < NX + ex; ++i) out.push_back(f(i, j, k));
  return out;
}

void fill_state(dyn::State &s, const dyn::ReferenceState &r) {
  for (Int k = 0; k <= NZ; ++k)
    for (Int j = 0; j <= NY; ++j)
      for (Int i = 0; i <= NX; ++i) {
        if (k < NZ && j < NY) s.u()(i, j, k) = 8.0 + 6.0 * hr(i, j, k, 0);
        if (k < NZ && i < NX) s.v()(i, j, k) = -3.0 + 5.0 * hr(i, j, k, 1);
        if (i < NX && j < NY) s.w()(i, j, k) = 0.8 * hr(i, j, k, 2);
        if (i < NX && j < NY && k < NZ) {
          s.rho()(i, j, k) = r.rho0()(i, j, k) + 0.02 * hr(i, j, k, 3);
          s.theta()(i, j, k) = r.theta0()(i, j, k) + 3.0 * hr(i, j, k, 4);
          s.pi()(i, j, k) = r.pi0()(i, j, k) + 0.002 * hr(i, j, k, 5);
          s.qv()(i, j, k) = std::max(0.0, 0.01 * (hr(i, j, k, 6) + 0.2));
          for (int q = 7; q < 12; ++q)
            s.field(static_cast<dyn::Species>(q))(i, j, k) = 0.002 * (hr(i, j, k, q) + 0.1);
        }
      }
}

const char *kSpecies[12] = {"u", "v", "w", "rho", "theta", "pi", "qv", "qc", "qr", "qi", "qs", "qg"};

void run(const std::string &tag, const grid::Grid &g, const dyn::ReferenceState &r, const std::string &adv,
         const std::string &cor, Real lat, Real damping, int order, const std::string &micro) {
  vibe::config::ModelConfig cfg;
  cfg.numerics.advection = adv;
  cfg.numerics.coriolis = cor;
  cfg.numerics.fplane_latitude = lat;
  cfg.numerics.divergence_damping = damping;
  cfg.numerics.divergence_order = order;
  cfg.physics.microphysics = micro;
  const dyn::Equations eq(g, r, cfg);
  dyn::State s(g);
  fill_state(s, r);
  dyn::Tendency d(g);
  eq.tendencies(s, d);
  for (int q = 0; q < 12; ++q)
    put(tag + "_tend_" + kSpecies[q], dump(d.field(static_cast<dyn::Species>(q)), q == 0, q == 1, q == 2));
  dyn::Tendency a(g);
  eq.acoustic_tendencies(s, a);
  for (int q : {0, 1, 2, 3, 5})
    put(tag + "_acoustic_" + kSpecies[q], dump(a.field(static_cast<dyn::Species>(q)), q == 0, q == 1, q == 2));
  put(tag + "_max_sound_speed", {eq.max_sound_speed(s)});
  dyn::State t = s;
  eq.diagnose(t);
  put(tag + "_diagnose_pi", dump(t.pi(), 0, 0, 0));
}

}  // namespace

int main() {
  const Real dx = 1000.0, dy = 1500.0, ztop = 12000.0;
  grid::Geometry flat;
  flat.nx = NX;
  flat.ny = NY;
  flat.nz = NZ;
  flat.dx = dx;
  flat.dy = dy;
  flat.z_top = ztop;
  flat.zeta = grid::make_stretched_zeta(NZ, 1.0, 1.1);
  flat.name = "literal";
  grid::Geometry hill = flat;
  hill.zs = grid::witch_of_agnesi(NX, NY, dx, dy, 600.0, 2500.0, 4000.0, 4500.0);
  const auto dec = grid::Decomposition::make(NX, NY, NZ, 1, 1, 4, true, true);
  const grid::Grid gflat(flat, dec), ghill(hill, dec);
  std::vector<Real> zp, thp;
  for (Int q = 0; q <= 12; ++q) {
    zp.push_back(1000.0 * q);
    thp.push_back(300.0 * std::exp(1e-4 / 9.80665 * 1000.0 * q));
  }
  const auto rhill = dyn::ReferenceState::from_profile(ghill, zp, thp, 100000.0, 0.0);
  const auto rflat = dyn::ReferenceState::isothermal(gflat, 300.0, 100000.0, 0.0);
  std::printf("{\n");
  put("zeta", flat.zeta);
  put("zs_hill", hill.zs);
  for (const auto &[tag, r] : {std::pair<std::string, const dyn::ReferenceState *>{"hill", &rhill},
                               std::pair<std::string, const dyn::ReferenceState *>{"flat", &rflat}}) {
    put(tag + "_rho0", dump(r->rho0(), 0, 0, 0));
    put(tag + "_theta0", dump(r->theta0(), 0, 0, 0));
    put(tag + "_pi0", dump(r->pi0(), 0, 0, 0));
  }
  run("c2", ghill, rhill, "central2", "fplane", 45.0, 0.0, 2, "kessler");
  run("c4", ghill, rhill, "central4", "none", 45.0, 2e5, 2, "none");
  run("c6", gflat, rflat, "central6", "betaplane", 30.0, 3e9, 4, "kessler");
  std::printf("\n}\n");
  return 0;
}
