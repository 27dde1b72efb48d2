// run_warm_bubble.cpp -- C++20 caller of the MI355X step, the role of the
// reference's examples/embed_model.cpp:61-131: build grid, reference state
// and initial state, construct the integrator, loop step(), report the
// energy budget before and after; optionally write the final state as a
// .vibebin restart (write_restart role, src/io/field_io.cpp:207-227).
//
//   run_warm_bubble [--nx N] [--ny N] [--nz N] [--steps K] [--restart FILE]
//                   [--reference-layout] [--dump FILE]
//   (defaults 64 x 64 x 40, 150 steps: the 250 m warm bubble)
//
// --reference-layout drives the integrator the way a reference-side adapter
// of Integrator::step does: the state lives in the reference's 12-species
// halo-4 layout, and every step converts it to the perturbation state, calls
// step() and converts back (ReferenceFields, from_reference, to_reference),
// with a boundary callback counting the steps.  --dump writes the final
// perturbation state (u, v, w, rho', theta', pi', FP64, the layout of
// vibe_mi/dycore.h) to FILE.  RELAX=1 adds a Newtonian relaxation of theta'
// through the per-stage physics callback.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vibe_mi/dycore.h"
#include "vibe_mi/dycore.hpp"

int main(int argc, char **argv) {
  int nx = 64, ny = 64, nz = 40, steps = 150;
  std::string restart, dump;
  bool ref_layout = false;
  for (int a = 1; a < argc; ++a) {
    const std::string k = argv[a];
    auto next = [&]() -> const char * {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "missing value after %s\n", k.c_str());
        std::exit(2);
      }
      return argv[++a];
    };
    if (k == "--nx") nx = std::atoi(next());
    else if (k == "--ny") ny = std::atoi(next());
    else if (k == "--nz") nz = std::atoi(next());
    else if (k == "--steps") steps = std::atoi(next());
    else if (k == "--restart") restart = next();
    else if (k == "--dump") dump = next();
    else if (k == "--reference-layout") ref_layout = true;
    else {
      std::fprintf(stderr, "unknown argument %s\n", k.c_str());
      return 2;
    }
  }
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
  std::printf("max stable dt %.6f s (step %.1f s)\n", integ.max_stable_dt(), dt);
  double t = 0.0;
  if (ref_layout) {
    // the reference's State: 12 species, halo 4; the reference state in the
    // same cell layout
    vibe_mi::ReferenceFields ref(nx, ny, nz, 4);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const std::size_t d = (static_cast<std::size_t>(k) * ny + j) * nx + i, c = ref.offset(i, j, k, 3);
          ref.rho0[c] = spec.rho0[d];
          ref.theta0[c] = spec.theta0[d];
          ref.pi0[c] = spec.pi0[d];
        }
    vibe_mi::to_reference(s, ref, true);
    int boundary_calls = 0;
    integ.set_boundary_callback([&](vibe_mi::State &, const vibe_mi::StepContext &) { ++boundary_calls; });
    vibe_mi::Integrator &base = integ;
    vibe_mi::State ps;
    for (int n = 1; n <= steps; ++n) {
      vibe_mi::from_reference(ref, ps);
      base.step(ps, vibe_mi::StepContext::make(dt, nsub, t));
      vibe_mi::to_reference(ps, ref, true);
      t += dt;
    }
    vibe_mi::from_reference(ref, s);
    const vibe_mi::IntegratorStats &st = base.stats();
    std::printf("reference layout: %lld steps, %lld acoustic substeps, %d boundary calls, last step %.3f ms\n",
                st.steps, st.acoustic_substeps, boundary_calls, 1e3 * st.wall_time);
  } else {
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
    integ.download(s);
  }
  const auto e1 = integ.energy_budget();
  std::printf("final:   mass %.12e kg  total energy %.12e J\n", e1.mass, e1.total);
  std::printf("relative mass change %.3e, energy change %.3e\n", (e1.mass - e0.mass) / e0.mass,
              (e1.total - e0.total) / e0.total);
  std::printf("max stable dt %.6f s\n", integ.max_stable_dt());
  if (!dump.empty()) {
    std::FILE *f = std::fopen(dump.c_str(), "wb");
    if (f == nullptr) {
      std::fprintf(stderr, "cannot write %s\n", dump.c_str());
      return 1;
    }
    for (const auto *v : {&s.u, &s.v, &s.w, &s.rho, &s.theta, &s.pi})
      if (std::fwrite(v->data(), sizeof(double), v->size(), f) != v->size()) return 1;
    std::fclose(f);
  }
  if (!restart.empty()) {
    vibe_mi::write_restart(restart, spec, s, {t, steps});
    vibe_mi::State back;
    const vibe_mi::RestartInfo info = vibe_mi::read_restart(restart, spec, back);
    std::printf("restart %s: t = %.1f s, step %d, theta' round trip %s\n", restart.c_str(), info.time, info.step,
                back.theta == s.theta ? "exact" : "DIFFERS");
  }
  return 0;
}
