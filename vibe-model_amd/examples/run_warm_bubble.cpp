// run_warm_bubble.cpp -- C++20 caller of the MI355X step, the role of the
// reference's examples/embed_model.cpp:61-131: build grid, reference state
// and initial state, construct the integrator, loop step(), report the
// energy budget before and after; optionally write the final state as a
// .vibebin restart (write_restart role, src/io/field_io.cpp:207-227).
//
//   run_warm_bubble [nx ny nz steps [restart.vibebin]]   (defaults 64 64 40 150)
//   RELAX=1 adds a Newtonian relaxation of theta' through the per-stage
//   physics callback.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vibe_mi/dycore.h"
#include "vibe_mi/dycore.hpp"

int main(int argc, char **argv) {
  const int nx = argc > 1 ? std::atoi(argv[1]) : 64;
  const int ny = argc > 2 ? std::atoi(argv[2]) : 64;
  const int nz = argc > 3 ? std::atoi(argv[3]) : 40;
  const int steps = argc > 4 ? std::atoi(argv[4]) : 150;
  const double dx = 250.0, ztop = 250.0 * nz, dt = 2.0;
  const int nsub = 6;

  vibe_mi::DomainSpec spec;
  spec.nx = nx;
  spec.ny = ny;
  spec.nz = nz;
  spec.dx = spec.dy = dx;
  spec.ztop = ztop;
  const std::size_t ncell = static_cast<std::size_t>(nx) * ny * nz;
  spec.zeta.resize(nz + 1);
  spec.zs.assign(static_cast<std::size_t>(nx) * ny, 0.0);
  spec.rho0.resize(ncell);
  spec.theta0.resize(ncell);
  spec.pi0.resize(ncell);
  spec.sponge.resize(nz);
  vibe_mi::check(vibe_mi_zeta_uniform(nz, spec.zeta.data()), "zeta");
  vibe_mi::check(vibe_mi_reference_state(0, 300.0, 0.0, 1e5, nx, ny, nz, ztop, spec.zeta.data(), spec.zs.data(),
                                         spec.rho0.data(), spec.theta0.data(), spec.pi0.data()),
                 "reference state");
  vibe_mi::check(vibe_mi_sponge(nz, ztop, spec.zeta.data(), 0.0, (ztop - 3000.0) / ztop, 0.2, spec.sponge.data()),
                 "sponge");

  vibe_mi::RungeKutta3Integrator integ(spec);
  vibe_mi::State s = integ.make_state();
  vibe_mi::check(vibe_mi_ic_bubble(nx, ny, nz, dx, dx, ztop, spec.zeta.data(), spec.zs.data(), 0.5 * nx * dx,
                                   0.5 * ny * dx, 2000.0, 2000.0, 2000.0, 2.0, s.theta.data()),
                 "bubble");
  vibe_mi::check(vibe_mi_ic_balance_pi(ncell, spec.rho0.data(), spec.theta0.data(), s.rho.data(), s.theta.data(),
                                       s.pi.data()),
                 "balance");
  integ.upload(s);
  // a weak Newtonian relaxation of theta' (10-minute time scale) evaluated
  // on every stage state: the per-stage physics coupling
  const vibe_mi_physics_fn relax = [](void *user, int, double, void *, const double *, const double *,
                                      const double *, const double *, const double *theta_p, const double *,
                                      double *theta_rate, double *, double *) -> int {
    const std::size_t n = *static_cast<const std::size_t *>(user);
    for (std::size_t c = 0; c < n; ++c) theta_rate[c] = -theta_p[c] / 600.0;
    return 0;
  };
  std::size_t ncell_user = ncell;
  if (std::getenv("RELAX") != nullptr) integ.set_physics(relax, &ncell_user, true);
  const auto e0 = integ.energy_budget();
  std::printf("initial: mass %.12e kg  total energy %.12e J\n", e0.mass, e0.total);
  double t = 0.0;
  for (int n = 1; n <= steps; ++n) {
    integ.advance(vibe_mi::StepContext::make(dt, nsub, t));
    t += dt;
    if (n % 50 == 0 || n == steps) {
      integ.download(s);
      double wmax = 0.0;
      for (double v : s.w) wmax = std::fmax(wmax, v);
      std::printf("  step %4d  t = %7.1f s  w_max = %8.3f m/s\n", n, t, wmax);
    }
  }
  const auto e1 = integ.energy_budget();
  std::printf("final:   mass %.12e kg  total energy %.12e J\n", e1.mass, e1.total);
  std::printf("relative mass change %.3e, energy change %.3e\n", (e1.mass - e0.mass) / e0.mass,
              (e1.total - e0.total) / e0.total);
  if (argc > 5) {
    integ.download(s);
    vibe_mi::write_restart(argv[5], spec, s, {t, steps});
    vibe_mi::State back;
    const vibe_mi::RestartInfo info = vibe_mi::read_restart(argv[5], spec, back);
    std::printf("restart %s: t = %.1f s, step %d, theta' round trip %s\n", argv[5], info.time, info.step,
                back.theta == s.theta ? "exact" : "DIFFERS");
  }
  return 0;
}
