// vibe_mi/semi_implicit.hpp -- C++20 host interface over semi_implicit.h.
//
// Mirrors the roles of the reference's semi-implicit integrator:
//   vibe::timeint::SemiImplicitIntegrator  include/vibe/time/semi_implicit.hpp:48-92
//     (step, name, stats, reset_stats, max_stable_dt, set_time_filter)
//   Integrator::set_boundary_callback / set_physics
//                                          include/vibe/time/integrator.hpp:74-77
//   vibe::dyn::State's 12 species as FULL values
//                                          include/vibe/dyn/state.hpp:53-56
// with StepContext, IntegratorStats and the exceptions of dycore.hpp.
// Header-only; link against libvibe_mi.so.  The state lives on the GPU
// between advance() calls; step() uploads and downloads a host LiteralState
// as the reference's Integrator::step(State&) does.
#pragma once

#include <array>
#include <chrono>
#include <cstddef>
#include <functional>
#include <vector>

#include "vibe_mi/dycore.hpp"
#include "vibe_mi/semi_implicit.h"

namespace vibe_mi {

/// The reference's 12 species (u, v, w, rho, theta, pi, qv, qc, qr, qi, qs,
/// qg) as dense interior arrays, i fastest: u [nz][ny][nx+1], v
/// [nz][ny+1][nx], w [nz+1][ny][nx], the others [nz][ny][nx]; full values.
struct LiteralState {
  int nx = 0, ny = 0, nz = 0;
  std::array<std::vector<double>, 12> q;
  double time = 0.0;
  int step = 0;
  LiteralState() = default;
  LiteralState(int nx_, int ny_, int nz_) : nx(nx_), ny(ny_), nz(nz_) {
    for (int s = 0; s < 12; ++s) q[s].assign(size(s), 0.0);
  }
  std::size_t size(int species) const {
    const std::size_t c = static_cast<std::size_t>(nx) * ny * nz;
    if (species == 0) return c + static_cast<std::size_t>(ny) * nz;
    if (species == 1) return c + static_cast<std::size_t>(nx) * nz;
    if (species == 2) return c + static_cast<std::size_t>(nx) * ny;
    return c;
  }
};

/// Grid, reference state and numerics of the reference's Equations, the
/// Damping sponge and the Helmholtz solver settings (host arrays copied at
/// construction).
struct SemiImplicitSpec {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, ztop = 0;
  std::vector<double> zeta, zs, rho0, theta0, pi0, sponge;  // zs, sponge empty: flat, none
  int order = 2, coriolis = 0, div_order = 2, moist = 0;
  double latitude = 45.0, div_coeff = 0.0;
  int solver = VIBE_MI_HELM_MULTIGRID;  // numerics.helmholtz_solver "krylov_multigrid" (the default)
  double tol = 1e-8;                    // numerics.helmholtz_tol
  int max_iter = 200;                   // numerics.helmholtz_max_iter
};

/// The reference's semi-implicit step on one GPU (SemiImplicitIntegrator role).
class SemiImplicitIntegrator {
 public:
  using BoundaryCallback = std::function<void(LiteralState &, const StepContext &)>;

  explicit SemiImplicitIntegrator(const SemiImplicitSpec &p) : nx_(p.nx), ny_(p.ny), nz_(p.nz) {
    vibe_mi_si_desc d{};
    d.eq.nx = p.nx;
    d.eq.ny = p.ny;
    d.eq.nz = p.nz;
    d.eq.dx = p.dx;
    d.eq.dy = p.dy;
    d.eq.z_top = p.ztop;
    d.eq.zeta = p.zeta.data();
    d.eq.zs = p.zs.empty() ? nullptr : p.zs.data();
    d.eq.rho0 = p.rho0.data();
    d.eq.theta0 = p.theta0.data();
    d.eq.pi0 = p.pi0.data();
    d.eq.order = p.order;
    d.eq.coriolis = p.coriolis;
    d.eq.latitude = p.latitude;
    d.eq.div_coeff = p.div_coeff;
    d.eq.div_order = p.div_order;
    d.eq.moist = p.moist;
    d.sponge = p.sponge.empty() ? nullptr : p.sponge.data();
    d.solver = p.solver;
    d.tol = p.tol;
    d.max_iter = p.max_iter;
    d.time_filter = 0.0;
    check(vibe_mi_si_create(&d, &h_), "vibe_mi_si_create");
  }
  ~SemiImplicitIntegrator() {
    if (h_ != nullptr) vibe_mi_si_destroy(h_);
  }
  SemiImplicitIntegrator(const SemiImplicitIntegrator &) = delete;
  SemiImplicitIntegrator &operator=(const SemiImplicitIntegrator &) = delete;

  const char *name() const noexcept { return "semi_implicit_mi355x"; }
  /// Robert-Asselin-Williams filter strength (0: off), from the next step on.
  void set_time_filter(double alpha) { check(vibe_mi_si_set_time_filter(h_, alpha), "vibe_mi_si_set_time_filter"); }
  void set_boundary_callback(BoundaryCallback cb) { boundary_ = std::move(cb); }
  /// Physics rates (theta, qv cells; u, v faces) added every step; empty
  /// vectors are zero, all empty switches the coupling off.
  void set_physics_tendency(const std::vector<double> &theta, const std::vector<double> &qv,
                            const std::vector<double> &u, const std::vector<double> &v) {
    auto p = [](const std::vector<double> &x) { return x.empty() ? nullptr : x.data(); };
    check(vibe_mi_si_set_physics_tendency(h_, p(theta), p(qv), p(u), p(v)), "vibe_mi_si_set_physics_tendency");
  }

  void upload(const LiteralState &s) {
    std::array<const double *, 12> p{};
    for (int q = 0; q < 12; ++q) p[q] = s.q[q].data();
    check(vibe_mi_si_set_state(h_, p.data()), "vibe_mi_si_set_state");
  }
  void download(LiteralState &s) const {
    std::array<double *, 12> p{};
    for (int q = 0; q < 12; ++q) p[q] = s.q[q].data();
    check(vibe_mi_si_get_state(h_, p.data()), "vibe_mi_si_get_state");
  }
  /// nsteps steps of the device state (no host transfer, no boundary callback).
  void advance(double dt, int nsteps = 1) {
    const int rc = vibe_mi_si_step(h_, dt, nsteps);
    refresh();
    check(rc, "vibe_mi_si_step");
  }
  /// Integrator::step semantics: upload s, one step, download into s, then
  /// the boundary callback on s (semi_implicit.cpp:167-226).
  void step(LiteralState &s, const StepContext &ctx) {
    upload(s);
    advance(ctx.dt, 1);
    download(s);
    s.time = ctx.time + ctx.dt;
    s.step += 1;
    if (boundary_) boundary_(s, ctx);
  }
  const IntegratorStats &stats() const noexcept { return stats_; }
  void reset_stats() {
    check(vibe_mi_si_reset_stats(h_), "vibe_mi_si_reset_stats");
    refresh();
  }
  /// 0.6 / (advective CFL at dt = 1) of s, or 600 s at rest
  /// (semi_implicit.cpp:228-233); s becomes the device state.
  double max_stable_dt(const LiteralState &s) {
    upload(s);
    double dt = 0.0;
    check(vibe_mi_si_max_stable_dt(h_, &dt), "vibe_mi_si_max_stable_dt");
    return dt;
  }
  /// pi_new of the last step's Helmholtz solve (cells).
  std::vector<double> pressure() const {
    std::vector<double> x(static_cast<std::size_t>(nx_) * ny_ * nz_);
    check(vibe_mi_si_pressure(h_, x.data()), "vibe_mi_si_pressure");
    return x;
  }
  LiteralState make_state() const { return LiteralState(nx_, ny_, nz_); }

 private:
  void refresh() {
    vibe_mi_si_stats st{};
    if (vibe_mi_si_get_stats(h_, &st) == 0) {
      stats_.steps = st.steps;
      stats_.acoustic_substeps = st.acoustic_substeps;
      stats_.wall_time = st.wall_time;
      stats_.helmholtz_iterations = st.helmholtz_iterations;
      stats_.helmholtz_residual = st.helmholtz_residual;
      stats_.diverged = st.diverged != 0;
    }
  }

  vibe_mi_si *h_ = nullptr;
  int nx_, ny_, nz_;
  IntegratorStats stats_{};
  BoundaryCallback boundary_;
};

}  // namespace vibe_mi
