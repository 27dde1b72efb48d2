// vibe_mi/dycore.hpp -- C++20 host interface over the C ABI (dycore.h).
//
// Mirrors the roles of the reference's time-integration interface:
//   vibe::timeint::StepContext          include/vibe/time/integrator.hpp:34-50
//   vibe::timeint::IntegratorStats      include/vibe/time/integrator.hpp:53-59
//   vibe::timeint::Integrator           include/vibe/time/integrator.hpp:62-82
//     (step, name, stats, reset_stats, max_stable_dt, set_boundary_callback)
//   RungeKutta3Integrator::max_stable_dt  src/time/runge_kutta.cpp:121-141
//   vibe::dyn::State (12 species, halo'd fields) <-> the dense perturbation
//     state                             include/vibe/dyn/state.hpp:53-56,
//                                       include/vibe/grid/field.hpp:34-96
//   vibe::timeint::RungeKutta3Integrator  include/vibe/time/runge_kutta.hpp:46-81
//   vibe::dyn::energy_budget            src/dyn/diagnostics.cpp:255-277
//   vibe::timeint::TimeStepController::inspect  src/time/timestep_control.cpp:23-56
//   vibe::NumericalError                include/vibe/common/error.hpp:35
//   Integrator::set_physics + RungeKutta3Integrator::slow_tendency's physics
//   coupling                            include/vibe/time/integrator.hpp:74,
//                                       src/time/runge_kutta.cpp:46-63
//   vibe::io::RestartInfo, write_restart, read_restart
//                                       include/vibe/io/field_io.hpp:82-94
//   vibe::IoError                       include/vibe/common/error.hpp:41
// Header-only; link against libvibe_mi.so.  The state lives on the GPU; the
// host-side State is used only to upload initial conditions and download
// results.
#pragma once

#include <array>
#include <chrono>
#include <cstddef>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vibe_mi/c_kernels.h"
#include "vibe_mi/dycore.h"
#include "vibe_mi/vibebin.h"

namespace vibe_mi {

/// Non-finite state (runge_kutta.cpp:103-106) or a device failure.
class NumericalError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};
class DeviceError : public std::runtime_error {
 public:
  DeviceError(const std::string &what, int code)
      : std::runtime_error(what + " failed with code " + std::to_string(code)), code_(code) {}
  int code() const noexcept { return code_; }

 private:
  int code_;
};

/// A .vibebin file could not be opened, parsed or written, or lacks a record.
class IoError : public std::runtime_error {
 public:
  IoError(const std::string &what, int code)
      : std::runtime_error(what + " failed with code " + std::to_string(code)), code_(code) {}
  int code() const noexcept { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char *what) {
  if (rc == VIBE_MI_ERR_NONFINITE) throw NumericalError(std::string(what) + ": non-finite model state");
  if (rc == VIBE_MI_ERR_PHYSICS) throw std::runtime_error(std::string(what) + ": the physics callback failed");
  if (rc == VIBE_MI_ERR_IO || rc == VIBE_MI_ERR_FORMAT || rc == VIBE_MI_ERR_NOT_FOUND) throw IoError(what, rc);
  if (rc != 0) throw DeviceError(what, rc);
}

/// Same fields and factory as the reference's StepContext.
struct StepContext {
  double dt = 1.0;
  double dt_acoustic = 1.0;
  int acoustic_substeps = 1;
  double time = 0.0;
  bool physics_on = true;  // false: this step runs without the physics coupling
  static StepContext make(double dt, int substeps, double time) {
    StepContext c;
    c.dt = dt;
    c.acoustic_substeps = substeps;
    c.dt_acoustic = dt / static_cast<double>(substeps);
    c.time = time;
    return c;
  }
};

/// Host copy of the prognostic perturbation state (layout of dycore.h).
struct State {
  std::vector<double> u, v, w, rho, theta, pi;
  double time = 0.0;
  int step = 0;
  State() = default;
  State(int nx, int ny, int nz) {
    const std::size_t c = static_cast<std::size_t>(nx) * ny * nz;
    u.assign(c, 0.0);
    v.assign(c, 0.0);
    w.assign(c + static_cast<std::size_t>(nx) * ny, 0.0);
    rho.assign(c, 0.0);
    theta.assign(c, 0.0);
    pi.assign(c, 0.0);
  }
};

/// Geometry + reference state of one domain (host arrays copied at creation).
struct DomainSpec {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, ztop = 0;
  std::vector<double> zeta, zs, rho0, theta0, pi0, sponge;
  double eps_s = 0.1;
  int precision = VIBE_MI_FP64;  // VIBE_MI_FP64, VIBE_MI_MIXED or VIBE_MI_MIXED_COMP
};

/// Same fields as the reference's StabilityReport (timestep_control.hpp:32-42)
/// that a dry state defines.
struct StabilityReport {
  double cfl_advection = 0, cfl_acoustic = 0, cfl_vertical = 0, max_wind = 0, min_dz = 0;
  bool stable = true;
};

struct EnergyBudget {
  double mass = 0, mass_perturbation = 0, kinetic = 0, internal = 0, potential = 0, total = 0;
};

/// Same fields as the reference's IntegratorStats (the Helmholtz ones stay 0:
/// this integrator is split-explicit).
struct IntegratorStats {
  long long steps = 0, acoustic_substeps = 0;
  double wall_time = 0.0;  // seconds of the last step() call
  int helmholtz_iterations = 0;
  double helmholtz_residual = 0.0;
  bool diverged = false;
};

/// Abstract integrator (Integrator role).
class Integrator {
 public:
  virtual ~Integrator() = default;
  virtual void step(State &s, const StepContext &ctx) = 0;
  virtual const char *name() const noexcept = 0;
  virtual const IntegratorStats &stats() const noexcept = 0;
  virtual void reset_stats() = 0;
  /// Largest stable large step for state s (CFL and acoustic limits).
  virtual double max_stable_dt(const State &s) = 0;
  /// Lateral boundary callback, called on the state after every step()
  /// (the reference calls it after the sponge, runge_kutta.cpp:114).
  using BoundaryCallback = std::function<void(State &, const StepContext &)>;
  virtual void set_boundary_callback(BoundaryCallback cb) { boundary_ = std::move(cb); }

 protected:
  BoundaryCallback boundary_;
};

/// WS-RK3 + HEVI acoustic substeps on one GPU (RungeKutta3Integrator role).
/// The GPU copy of the state is authoritative between calls: step() on the
/// State it was last synchronised with only downloads when asked to.
class RungeKutta3Integrator final : public Integrator {
 public:
  explicit RungeKutta3Integrator(const DomainSpec &spec) : nx_(spec.nx), ny_(spec.ny), nz_(spec.nz) {
    vibe_mi_domain_desc d{};
    d.nx = spec.nx;
    d.ny = spec.ny;
    d.nz = spec.nz;
    d.dx = spec.dx;
    d.dy = spec.dy;
    d.ztop = spec.ztop;
    d.zeta = spec.zeta.data();
    d.zs = spec.zs.data();
    d.rho0 = spec.rho0.data();
    d.theta0 = spec.theta0.data();
    d.pi0 = spec.pi0.data();
    d.sponge = spec.sponge.data();
    d.eps_s = spec.eps_s;
    d.precision = spec.precision;
    check(vibe_mi_domain_create(&d, &h_), "vibe_mi_domain_create");
  }
  ~RungeKutta3Integrator() override {
    if (h_ != nullptr) vibe_mi_domain_destroy(h_);
  }
  RungeKutta3Integrator(const RungeKutta3Integrator &) = delete;
  RungeKutta3Integrator &operator=(const RungeKutta3Integrator &) = delete;

  const char *name() const noexcept override { return "rk3_acoustic_hevi_mi355x"; }

  void upload(const State &s) {
    check(vibe_mi_domain_set_state(h_, s.u.data(), s.v.data(), s.w.data(), s.rho.data(), s.theta.data(),
                                   s.pi.data()),
          "vibe_mi_domain_set_state");
  }
  void download(State &s) const {
    check(vibe_mi_domain_get_state(h_, s.u.data(), s.v.data(), s.w.data(), s.rho.data(), s.theta.data(),
                                   s.pi.data()),
          "vibe_mi_domain_get_state");
  }
  /// Advance the device state by n steps (no host transfer); the boundary
  /// callback, which needs the host state, runs in step() only.
  void advance(const StepContext &ctx, int nsteps = 1) {
    if (!ctx.physics_on) check(vibe_mi_domain_physics_enable(h_, 0), "vibe_mi_domain_physics_enable");
    const int rc = vibe_mi_domain_step(h_, ctx.dt, ctx.acoustic_substeps, nsteps);
    if (!ctx.physics_on) (void)vibe_mi_domain_physics_enable(h_, 1);
    if (rc == VIBE_MI_ERR_NONFINITE) stats_.diverged = true;
    check(rc, "vibe_mi_domain_step");
    stats_.steps += nsteps;
    stats_.acoustic_substeps += 3LL * (ctx.acoustic_substeps < 1 ? 1 : ctx.acoustic_substeps) * nsteps;
  }
  /// Integrator::step semantics: upload s, one step, download into s, then
  /// the boundary callback on s.
  void step(State &s, const StepContext &ctx) override {
    const auto t0 = std::chrono::steady_clock::now();
    upload(s);
    advance(ctx, 1);
    download(s);
    s.time = ctx.time + ctx.dt;
    s.step += 1;
    if (boundary_) boundary_(s, ctx);
    stats_.wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  const IntegratorStats &stats() const noexcept override { return stats_; }
  void reset_stats() override { stats_ = IntegratorStats{}; }
  /// max_stable_dt of the device state (runge_kutta.cpp:121-141).
  double max_stable_dt() const {
    double dt = 0.0;
    check(vibe_mi_domain_max_stable_dt(h_, &dt), "vibe_mi_domain_max_stable_dt");
    return dt;
  }
  /// max_stable_dt of s: s becomes the device state (as in step()).
  double max_stable_dt(const State &s) override {
    upload(s);
    return max_stable_dt();
  }
  EnergyBudget energy_budget() const {
    double e[6];
    check(vibe_mi_domain_energy(h_, e), "vibe_mi_domain_energy");
    return {e[0], e[1], e[2], e[3], e[4], e[5]};
  }
  /// TimeStepController::inspect role (timestep_control.cpp:23-56).
  StabilityReport inspect(double dt) const {
    double o[6];
    check(vibe_mi_domain_stability(h_, dt, o), "vibe_mi_domain_stability");
    return {o[0], o[1], o[2], o[3], o[4], o[5] != 0.0};
  }
  std::pair<double, double> cfl(double dt) const {
    double c = 0, w = 0;
    check(vibe_mi_domain_cfl(h_, dt, &c, &w), "vibe_mi_domain_cfl");
    return {c, w};
  }
  State make_state() const { return State(nx_, ny_, nz_); }
  int nx() const noexcept { return nx_; }
  int ny() const noexcept { return ny_; }
  int nz() const noexcept { return nz_; }

  /// Physics rates held for every following stage: theta' (K/s), u and v
  /// (m/s^2) in the state layout; empty vectors mean zero, all empty
  /// switches the coupling off.
  void set_physics_tendency(const std::vector<double> &theta, const std::vector<double> &u,
                            const std::vector<double> &v) {
    auto p = [](const std::vector<double> &x) { return x.empty() ? nullptr : x.data(); };
    check(vibe_mi_domain_set_physics_tendency(h_, p(theta), p(u), p(v)), "vibe_mi_domain_set_physics_tendency");
  }
  /// Per-stage physics (Integrator::set_physics role): fn is called before
  /// each stage's slow tendency with the stage state and zeroed rate arrays
  /// (host copies when host is true, else device pointers on the domain's
  /// stream); nullptr removes it.
  void set_physics(vibe_mi_physics_fn fn, void *user, bool host = true) {
    check(vibe_mi_domain_set_physics_callback(h_, fn, user, host ? 1 : 0), "vibe_mi_domain_set_physics_callback");
  }

 private:
  vibe_mi_domain *h_ = nullptr;
  int nx_, ny_, nz_;
  IntegratorStats stats_;
};

/// The reference's prognostic State as plain arrays: the 12 species of
/// dyn::Species in order (u, v, w, rho, theta, pi, qv, qc, qr, qi, qs, qg),
/// each a halo'd field of the reference layout (vibe_mi/dycore.h), and the
/// ReferenceState's rho0, theta0, pi0 in the cell layout.
struct ReferenceFields {
  int nx = 0, ny = 0, nz = 0, halo = 4;
  std::array<std::vector<double>, 12> species;
  std::vector<double> rho0, theta0, pi0;
  ReferenceFields() = default;
  ReferenceFields(int nx_, int ny_, int nz_, int halo_) : nx(nx_), ny(ny_), nz(nz_), halo(halo_) {
    for (int q = 0; q < 12; ++q) species[q].assign(vibe_mi_reference_field_size(nx, ny, nz, halo, q), 0.0);
    const std::size_t nc = vibe_mi_reference_field_size(nx, ny, nz, halo, 3);
    rho0.assign(nc, 0.0);
    theta0.assign(nc, 0.0);
    pi0.assign(nc, 0.0);
  }
  std::size_t offset(int i, int j, int k, int sp) const {
    const std::size_t nsx = static_cast<std::size_t>(nx + 2 * halo + (sp == 0 ? 1 : 0));
    const std::size_t nsy = static_cast<std::size_t>(ny + 2 * halo + (sp == 1 ? 1 : 0));
    return (static_cast<std::size_t>(k + halo) * nsy + static_cast<std::size_t>(j + halo)) * nsx +
           static_cast<std::size_t>(i + halo);
  }
};

/// dyn::State totals -> perturbation State (the adapter's step-in).
inline void from_reference(const ReferenceFields &r, State &s) {
  if (s.u.empty()) s = State(r.nx, r.ny, r.nz);
  const double *sp[12];
  for (int q = 0; q < 12; ++q) sp[q] = r.species[q].data();
  check(vibe_mi_state_from_reference(r.nx, r.ny, r.nz, r.halo, sp, r.rho0.data(), r.theta0.data(), r.pi0.data(),
                                     s.u.data(), s.v.data(), s.w.data(), s.rho.data(), s.theta.data(), s.pi.data()),
        "vibe_mi_state_from_reference");
}

/// Perturbation State -> dyn::State totals (the adapter's step-out); the
/// moisture species are left as they are.
inline void to_reference(const State &s, ReferenceFields &r, bool fill_halo = false) {
  double *sp[12];
  for (int q = 0; q < 12; ++q) sp[q] = r.species[q].data();
  check(vibe_mi_state_to_reference(r.nx, r.ny, r.nz, r.halo, s.u.data(), s.v.data(), s.w.data(), s.rho.data(),
                                   s.theta.data(), s.pi.data(), r.rho0.data(), r.theta0.data(), r.pi0.data(),
                                   fill_halo ? 1 : 0, sp),
        "vibe_mi_state_to_reference");
}

/// time and step of a restart (the fields of the reference's RestartInfo that
/// its restart file stores).
struct RestartInfo {
  double time = 0.0;
  int step = 0;
};

/// Writes s as a .vibebin restart file (write_restart role): the reference's
/// 12 species as totals against spec's reference state, plus the exact
/// perturbations.
inline void write_restart(const std::string &path, const DomainSpec &spec, const State &s,
                          const RestartInfo &info) {
  check(vibe_mi_restart_write(path.c_str(), spec.nx, spec.ny, spec.nz, info.time, info.step, s.u.data(),
                              s.v.data(), s.w.data(), s.rho.data(), s.theta.data(), s.pi.data(),
                              spec.rho0.data(), spec.theta0.data(), spec.pi0.data()),
        "vibe_mi_restart_write");
}

/// Reads a restart file into s (read_restart role); s.time and s.step are set
/// from the file as the reference does.
inline RestartInfo read_restart(const std::string &path, const DomainSpec &spec, State &s) {
  if (s.u.empty()) s = State(spec.nx, spec.ny, spec.nz);
  RestartInfo info;
  check(vibe_mi_restart_read(path.c_str(), spec.nx, spec.ny, spec.nz, &info.time, &info.step, s.u.data(),
                             s.v.data(), s.w.data(), s.rho.data(), s.theta.data(), s.pi.data(),
                             spec.rho0.data(), spec.theta0.data(), spec.pi0.data()),
        "vibe_mi_restart_read");
  s.time = info.time;
  s.step = info.step;
  return info;
}

}  // namespace vibe_mi
