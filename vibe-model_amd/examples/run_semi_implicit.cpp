// run_semi_implicit.cpp -- C++20 caller of the reference's semi-implicit step
// on the GPU (include/vibe_mi/semi_implicit.hpp), the role of a reference
// driver holding a SemiImplicitIntegrator (src/driver/driver.cpp:128-133):
// build an Agnesi-ridge grid, the reference's isothermal state and a
// perturbed initial state, construct the integrator, loop step() on the
// host state with a boundary callback, report the statistics and
// max_stable_dt, and dump the final 12 species.
//
//   run_semi_implicit [--nx N] [--ny N] [--nz N] [--steps K] [--dt S]
//                     [--dump FILE]
//   (defaults 40 x 24 x 16, 4 steps of 3 s)
//
// The initial state is a deterministic hash of the indices (no libm calls,
// so any caller can rebuild it bit for bit): r(q, i, j, k) in [-0.5, 0.5),
// u = 8 + 4 r, v = -2 + 3 r, w = 0.2 r, rho = rho0 (1 + 1e-3 r),
// theta = theta0 + r, pi = pi0 (1 + 1e-4 r), qv = 5e-3 (r + 0.5), qc..qg =
// 1e-4 (r + 0.5); physics rate on theta 1e-4 r(12, i, j, k) K/s.  Numerics:
// fourth-order advection, f-plane at 40 degrees, Del^2 divergence damping
// 1e5, moisture on, sponge above 70 % of the depth (alpha 0.02), the
// two-grid Helmholtz solver to 1e-9, time filter 0.05.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "vibe_mi/dycore.h"
#include "vibe_mi/semi_implicit.hpp"

namespace {

double hash_r(int q, int i, int j, int k) {
  const std::uint64_t h = ((static_cast<std::uint64_t>(i + 7) * 73856093ull) ^
                           (static_cast<std::uint64_t>(j + 11) * 19349663ull) ^
                           (static_cast<std::uint64_t>(k + 13) * 83492791ull) ^
                           (static_cast<std::uint64_t>(q + 3) * 2654435761ull)) &
                          0xFFFFFull;
  return static_cast<double>(h) / 1048576.0 - 0.5;
}

}  // namespace

int main(int argc, char **argv) {
  int nx = 40, ny = 24, nz = 16, steps = 4;
  double dt = 3.0;
  std::string dump;
  for (int a = 1; a < argc; ++a) {
    const std::string k = argv[a];
    if (a + 1 >= argc) {
      std::fprintf(stderr, "missing value after %s\n", k.c_str());
      return 2;
    }
    const char *v = argv[++a];
    if (k == "--nx") nx = std::atoi(v);
    else if (k == "--ny") ny = std::atoi(v);
    else if (k == "--nz") nz = std::atoi(v);
    else if (k == "--steps") steps = std::atoi(v);
    else if (k == "--dt") dt = std::atof(v);
    else if (k == "--dump") dump = v;
    else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      return 2;
    }
  }
  vibe_mi::SemiImplicitSpec sp;
  sp.nx = nx;
  sp.ny = ny;
  sp.nz = nz;
  sp.dx = 1000.0;
  sp.dy = 1200.0;
  sp.ztop = 15000.0;
  const std::size_t nc = static_cast<std::size_t>(nx) * ny * nz;
  sp.zeta.resize(nz + 1);
  sp.zs.resize(static_cast<std::size_t>(nx) * ny);
  sp.rho0.resize(nc);
  sp.theta0.resize(nc);
  sp.pi0.resize(nc);
  sp.sponge.resize(nz);
  vibe_mi::check(vibe_mi_zeta_stretched(nz, 1.0, 1.05, sp.zeta.data()), "vibe_mi_zeta_stretched");
  vibe_mi::check(vibe_mi_terrain_agnesi_bell(nx, ny, sp.dx, sp.dy, 600.0, 8000.0, 20000.0, 12000.0, sp.zs.data()),
                 "vibe_mi_terrain_agnesi_bell");
  vibe_mi::check(vibe_mi_reference_state(0, 300.0, 0.0, 100000.0, nx, ny, nz, sp.ztop, sp.zeta.data(), sp.zs.data(),
                                         sp.rho0.data(), sp.theta0.data(), sp.pi0.data()),
                 "vibe_mi_reference_state");
  vibe_mi::check(vibe_mi_sponge(nz, sp.ztop, sp.zeta.data(), 0.0, 0.7, 0.02, sp.sponge.data()), "vibe_mi_sponge");
  sp.order = 4;
  sp.coriolis = 1;
  sp.latitude = 40.0;
  sp.div_coeff = 1e5;
  sp.div_order = 2;
  sp.moist = 1;
  sp.solver = VIBE_MI_HELM_MULTIGRID;
  sp.tol = 1e-9;
  sp.max_iter = 400;

  vibe_mi::SemiImplicitIntegrator integ(sp);
  integ.set_time_filter(0.05);
  vibe_mi::LiteralState s = integ.make_state();
  for (int q = 0; q < 12; ++q) {
    const int ex = nx + (q == 0), ey = ny + (q == 1), ez = nz + (q == 2);
    for (int k = 0; k < ez; ++k)
      for (int j = 0; j < ey; ++j)
        for (int i = 0; i < ex; ++i) {
          const std::size_t n = (static_cast<std::size_t>(k) * ey + j) * ex + i;
          const double r = hash_r(q, i, j, k);
          double val;
          switch (q) {
            case 0: val = 8.0 + 4.0 * r; break;
            case 1: val = -2.0 + 3.0 * r; break;
            case 2: val = 0.2 * r; break;
            case 3: val = sp.rho0[n] * (1.0 + 1e-3 * r); break;
            case 4: val = sp.theta0[n] + r; break;
            case 5: val = sp.pi0[n] * (1.0 + 1e-4 * r); break;
            case 6: val = 5e-3 * (r + 0.5); break;
            default: val = 1e-4 * (r + 0.5); break;
          }
          s.q[q][n] = val;
        }
  }
  std::vector<double> rate(nc);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) rate[(static_cast<std::size_t>(k) * ny + j) * nx + i] = 1e-4 * hash_r(12, i, j, k);
  integ.set_physics_tendency(rate, {}, {}, {});
  int calls = 0;
  integ.set_boundary_callback([&](vibe_mi::LiteralState &, const vibe_mi::StepContext &) { ++calls; });

  for (int n = 0; n < steps; ++n) integ.step(s, vibe_mi::StepContext::make(dt, 1, s.time));
  const vibe_mi::IntegratorStats &st = integ.stats();
  std::printf("%s: %lld steps, %d Helmholtz iterations, last residual %.3e, boundary calls %d, time %.1f s\n",
              integ.name(), st.steps, st.helmholtz_iterations, st.helmholtz_residual, calls, s.time);
  std::printf("max_stable_dt %.17g\n", integ.max_stable_dt(s));
  if (!dump.empty()) {
    std::FILE *f = std::fopen(dump.c_str(), "wb");
    if (f == nullptr) {
      std::fprintf(stderr, "cannot write %s\n", dump.c_str());
      return 1;
    }
    for (int q = 0; q < 12; ++q) std::fwrite(s.q[q].data(), sizeof(double), s.q[q].size(), f);
    std::fclose(f);
  }
  return 0;
}
