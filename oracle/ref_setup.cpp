// ref_setup.cpp -- golden vectors of the reference's grid and reference-state
// setup, computed by the REFERENCE's own code (src/grid/geometry.cpp,
// src/dyn/damping.cpp, src/dyn/reference_state.cpp, compiled from
// /root/reference by oracle/ref.mk).
//
// Test infrastructure: tests/golden/make_setup_golden.py runs it and commits
// the JSON it prints; tests/test_setup_golden.py checks libvibe_mi.so's
// host setup helpers against it.  Cases:
//   make_stretched_zeta(16, 1, 1.08) and (40, 1, 1.05)      geometry.cpp:39-60
//   witch_of_agnesi(8, 6, 1000, 1500, 800, 3000, 3000, 4000)  geometry.cpp:103-117
//   smooth_terrain(that, 8, 6, 2 passes, max_slope 50)       geometry.cpp:70-101
//   Damping sponge profile, 16 stretched levels, ztop 20 km,
//     sponge from 14 km, alpha_max 0.2                        damping.cpp:24-41
//   ReferenceState::isothermal(theta 300 K, p_s 1000 hPa) on the
//     8 x 6 x 16 grid with the Agnesi terrain                 reference_state.cpp:35-59
//   ReferenceState::from_profile of a constant-N sounding (p_s 1000 hPa)
//     and of an uneven sounding with an inversion (985 hPa) on that grid
//                                                          reference_state.cpp:61-110
//   energy_budget of a dry state on that grid and reference state
//     (diagnostics.cpp:255-277): u = 5 + i/4 + j/2 - k/8 (face nx = face 0),
//     v = -2 + i/10 - j/5 + k/20 (face ny = face 0), w = (i - j)/100 + k/500,
//     rho' = 1e-3 (i - j + k) / 16, theta' = (i + 2 j - k) / 8,
//     pi' = 1e-4 (2 i - j + k) / 16; totals = reference + perturbation
//   Advection::cfl_number of that state for dt = 6 s     advection.cpp:325-340
//   TimeStepController::inspect of that state, dt = 6 s  timestep_control.cpp:23-56
// (acoustic_substep.cpp is not linked: it needs helmholtz.cpp, which does not
// compile here.)
//   nest operators (src/grid/nest.cpp:34-128) on a cell field: parent
//     16 x 14 x 3 at 3 km, child 12 x 9 x 3 at 1 km, ratio 3, offsets
//     (12, 15) child cells (inside the parent: no clamping), Davies zone 4;
//     parent p = i/2 - j/4 + k/8 + i j / 100, child c = 10 + i/10 - j/5 + k;
//     interpolate_from_parent(p) and apply_lateral_boundary(c, that),
//     restrict_to_parent(the relaxed child) onto a copy of p, and
//     davies_profile(4, 1) and (5, 0.8)
//   the raw storage of a State's 12 species on a 5 x 4 x 3 grid with halo 2
//     (field.hpp:34-96), filled through its own (i, j, k) indexing
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "vibe/config/config.hpp"
#include "vibe/dyn/advection.hpp"
#include "vibe/dyn/damping.hpp"
#include "vibe/dyn/diagnostics.hpp"
#include "vibe/dyn/state.hpp"
#include "vibe/dyn/reference_state.hpp"
#include "vibe/grid/geometry.hpp"
#include "vibe/grid/nest.hpp"
#include "vibe/time/timestep_control.hpp"

using vibe::Int;
using vibe::Real;
namespace grid = vibe::grid;

namespace {

void put(const char *name, const std::vector<Real> &v, bool last = false) {
  std::printf("  \"%s\": [", name);
  for (size_t q = 0; q < v.size(); ++q) std::printf("%s%.17g", q ? ", " : "", v[q]);
  std::printf("]%s\n", last ? "" : ",");
}

}  // namespace

int main() {
  const Int nx = 8, ny = 6, nz = 16;
  const Real dx = 1000.0, dy = 1500.0, ztop = 20000.0;
  std::printf("{\n");
  const std::vector<Real> z16 = grid::make_stretched_zeta(nz, 1.0, 1.08);
  put("zeta_16_1_1.08", z16);
  put("zeta_40_1_1.05", grid::make_stretched_zeta(40, 1.0, 1.05));
  std::vector<Real> zs = grid::witch_of_agnesi(nx, ny, dx, dy, 800.0, 3000.0, 3000.0, 4000.0);
  put("agnesi_8x6", zs);
  std::vector<Real> sm = zs;
  grid::smooth_terrain(sm, nx, ny, 2, 50.0);
  put("agnesi_8x6_smoothed_2_50", sm);

  grid::Geometry flat;
  flat.nx = nx;
  flat.ny = ny;
  flat.nz = nz;
  flat.dx = dx;
  flat.dy = dy;
  flat.z_top = ztop;
  flat.zeta = z16;
  flat.name = "golden";
  const grid::Grid gflat(flat, grid::Decomposition::make(nx, ny, nz, 1, 1, 2, true, true));
  const vibe::dyn::Damping damp(gflat, 14000.0, ztop, 0.2);
  put("sponge_16_14km_0.2", damp.sponge_profile());

  grid::Geometry hill = flat;
  hill.zs = zs;
  const grid::Grid ghill(hill, grid::Decomposition::make(nx, ny, nz, 1, 1, 2, true, true));
  const auto rs = vibe::dyn::ReferenceState::isothermal(ghill, 300.0, 100000.0, 0.0);
  std::vector<Real> th, pi, rho;
  for (Int k = 0; k < nz; ++k)
    for (Int j = 0; j < ny; ++j)
      for (Int i = 0; i < nx; ++i) {
        th.push_back(rs.theta0()(i, j, k));
        pi.push_back(rs.pi0()(i, j, k));
        rho.push_back(rs.rho0()(i, j, k));
      }
  put("isothermal_theta0", th);
  put("isothermal_pi0", pi);
  put("isothermal_rho0", rho);

  // ReferenceState::from_profile on the same grid: a constant-N sounding
  // (N = 0.01 1/s, 300 K at the ground, 1 km spacing to 20 km) and the
  // reference's own two-layer standard atmosphere (built on from_profile)
  std::vector<Real> zp, thp;
  for (Int q = 0; q <= 20; ++q) {
    zp.push_back(1000.0 * q);
    thp.push_back(300.0 * std::exp(1e-4 / 9.80665 * 1000.0 * q));
  }
  put("profile_n001_z", zp);
  put("profile_n001_theta", thp);
  auto dump3 = [&](const vibe::dyn::ReferenceState &r, const char *tag) {
    std::vector<Real> a, b, c;
    for (Int k = 0; k < nz; ++k)
      for (Int j = 0; j < ny; ++j)
        for (Int i = 0; i < nx; ++i) {
          a.push_back(r.theta0()(i, j, k));
          b.push_back(r.pi0()(i, j, k));
          c.push_back(r.rho0()(i, j, k));
        }
    std::string n(tag);
    put((n + "_theta0").c_str(), a);
    put((n + "_pi0").c_str(), b);
    put((n + "_rho0").c_str(), c);
  };
  dump3(vibe::dyn::ReferenceState::from_profile(ghill, zp, thp, 100000.0, 0.0), "from_profile_n001");
  // an uneven sounding with an inversion, starting above part of the terrain
  // and ending below the lid (both held-value branches of the interpolation)
  const std::vector<Real> zq{200.0, 500.0, 1500.0, 1600.0, 4000.0, 9000.0, 12000.0};
  const std::vector<Real> thq{295.0, 296.0, 299.0, 304.0, 310.0, 325.0, 360.0};
  put("profile_uneven_z", zq);
  put("profile_uneven_theta", thq);
  dump3(vibe::dyn::ReferenceState::from_profile(ghill, zq, thq, 98500.0, 0.0), "from_profile_uneven");

  vibe::dyn::State st(ghill);
  for (Int k = 0; k <= nz; ++k)
    for (Int j = 0; j <= ny; ++j)
      for (Int i = 0; i <= nx; ++i) {
        const Int iw = i == nx ? 0 : i, js = j == ny ? 0 : j;
        if (k < nz && j < ny) st.u()(i, j, k) = 5.0 + 0.25 * iw + 0.5 * j - 0.125 * k;
        if (k < nz && i < nx) st.v()(i, j, k) = -2.0 + 0.1 * i - 0.2 * js + 0.05 * k;
        if (i < nx && j < ny) st.w()(i, j, k) = 0.01 * (i - j) + 0.002 * k;
        if (i < nx && j < ny && k < nz) {
          st.rho()(i, j, k) = rs.rho0()(i, j, k) + 1e-3 * (i - j + k) / 16.0;
          st.theta()(i, j, k) = rs.theta0()(i, j, k) + (i + 2 * j - k) / 8.0;
          st.pi()(i, j, k) = rs.pi0()(i, j, k) + 1e-4 * (2 * i - j + k) / 16.0;
        }
      }
  const vibe::dyn::EnergyBudget e = vibe::dyn::energy_budget(st, rs);
  put("energy_mass_kinetic_internal_potential_total", {e.mass, e.kinetic, e.internal, e.potential, e.total});
  const vibe::dyn::Advection adv(ghill, vibe::dyn::AdvectionScheme::Central2);  // the CFL scan is the same for every scheme
  put("cfl_dt6", {adv.cfl_number(st, 6.0)});
  vibe::config::ModelConfig cfg;
  cfg.time.dt = 6.0;
  const vibe::timeint::TimeStepController tsc(cfg, ghill);
  const vibe::timeint::StabilityReport rep = tsc.inspect(st, 6.0);
  put("inspect_dt6_adv_acoustic_vertical_maxwind_mindz_stable",
      {rep.cfl_advection, rep.cfl_acoustic, rep.cfl_vertical, rep.max_wind, rep.min_dz, rep.stable ? 1.0 : 0.0});

  auto flat_grid = [](Int gx, Int gy, Int gz, Real d) {
    grid::Geometry g;
    g.nx = gx;
    g.ny = gy;
    g.nz = gz;
    g.dx = d;
    g.dy = d;
    g.z_top = 10000.0;
    g.zeta.resize(gz + 1);
    for (Int k = 0; k <= gz; ++k) g.zeta[k] = static_cast<Real>(k) / gz;
    g.name = "nest";
    return grid::Grid(g, grid::Decomposition::make(gx, gy, gz, 1, 1, 2, true, true));
  };
  const grid::Grid gpar = flat_grid(16, 14, 3, 3000.0), gchi = flat_grid(12, 9, 3, 1000.0);
  grid::NestPlacement pl;
  pl.i0 = 12;
  pl.j0 = 15;
  pl.ratio = 3;
  pl.boundary_zone = 4;
  pl.iext = 4;
  pl.jext = 3;
  const grid::Nest nest(gchi, gpar, pl);
  grid::Field<Real> P(gpar, grid::Stagger::Cell, "p"), I(gchi, grid::Stagger::Cell, "i"),
      Cf(gchi, grid::Stagger::Cell, "c");
  for (Int k = 0; k < 3; ++k)
    for (Int j = 0; j < 14; ++j)
      for (Int i = 0; i < 16; ++i) P(i, j, k) = 0.5 * i - 0.25 * j + 0.125 * k + 0.01 * i * j;
  for (Int k = 0; k < 3; ++k)
    for (Int j = 0; j < 9; ++j)
      for (Int i = 0; i < 12; ++i) Cf(i, j, k) = 10.0 + 0.1 * i - 0.2 * j + k;
  std::vector<grid::Field<Real> *> ps{&P}, is{&I}, cs{&Cf};
  nest.interpolate_from_parent(ps, is);
  const std::vector<grid::Field<Real> *> ics{&I};
  nest.apply_lateral_boundary(cs, ics);
  std::vector<Real> vi, vc;
  for (Int k = 0; k < 3; ++k)
    for (Int j = 0; j < 9; ++j)
      for (Int i = 0; i < 12; ++i) {
        vi.push_back(I(i, j, k));
        vc.push_back(Cf(i, j, k));
      }
  put("nest_interpolated_cell", vi);
  put("nest_relaxed_cell_zone4", vc);
  // restrict_to_parent of the relaxed child onto a copy of p; this call reads
  // NestPlacement::i0, j0 as parent cells (nest.cpp:103), so the placement
  // is (4, 5): the parent cells the child's 12 x 9 cells at offset (12, 15)
  // child cells cover
  grid::NestPlacement pr = pl;
  pr.i0 = 4;
  pr.j0 = 5;
  const grid::Nest nest_r(gchi, gpar, pr);
  grid::Field<Real> Pr(gpar, grid::Stagger::Cell, "pr");
  for (Int k = 0; k < 3; ++k)
    for (Int j = 0; j < 14; ++j)
      for (Int i = 0; i < 16; ++i) Pr(i, j, k) = P(i, j, k);
  std::vector<grid::Field<Real> *> prs{&Pr};
  nest_r.restrict_to_parent(cs, prs);
  std::vector<Real> vr;
  for (Int k = 0; k < 3; ++k)
    for (Int j = 0; j < 14; ++j)
      for (Int i = 0; i < 16; ++i) vr.push_back(Pr(i, j, k));
  put("nest_restricted_parent_cell", vr);
  // the State layout (field.hpp:34-96, state.hpp:53-56): a 5 x 4 x 3 grid
  // with halo 2, species q filled through the reference's own indexing,
  // value(q, i, j, k) = 1000 q + 100 k + 10 j + i on every interior point and
  // staggered face (halo points stay 0); raw arrays printed
  {
    grid::Geometry lg;
    lg.nx = 5;
    lg.ny = 4;
    lg.nz = 3;
    lg.dx = lg.dy = 1000.0;
    lg.z_top = 3000.0;
    lg.zeta = {0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0};
    lg.name = "layout";
    const grid::Grid glay(lg, grid::Decomposition::make(5, 4, 3, 1, 1, 2, true, true));
    vibe::dyn::State ls(glay);
    for (int q = 0; q < vibe::dyn::kNumSpecies; ++q) {
      const auto sp = static_cast<vibe::dyn::Species>(q);
      grid::Field<Real> &f = ls.field(sp);
      const Int ni = 5 + (q == 0 ? 1 : 0), nj = 4 + (q == 1 ? 1 : 0), nk = 3 + (q == 2 ? 1 : 0);
      for (Int k = 0; k < nk; ++k)
        for (Int j = 0; j < nj; ++j)
          for (Int i = 0; i < ni; ++i) f(i, j, k) = 1000.0 * q + 100.0 * k + 10.0 * j + i;
      put((std::string("layout_h2_") + vibe::dyn::species_name(sp)).c_str(), f.storage());
    }
  }
  put("davies_4_1", grid::davies_profile(4, 1.0));
  put("davies_5_0.8", grid::davies_profile(5, 0.8), true);
  std::printf("}\n");
  return 0;
}
