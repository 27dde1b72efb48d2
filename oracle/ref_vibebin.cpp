// ref_vibebin.cpp -- writes small .vibebin files with the REFERENCE's own
// BinaryWriter (src/io/binary_writer.cpp, compiled from /root/reference by
// oracle/ref.mk together with src/grid/geometry.cpp for grid::Grid).
//
// Test infrastructure: the files it writes are committed as golden fixtures
// under tests/golden/ (tests/golden/make_vibebin_golden.sh) and compared byte
// for byte with what libvibe_mi.so writes for the same content.
//
//   ref_vibebin generic  OUT   4x3x2 grid, time 3.5: __restart_info__, u, v,
//                              w, zeta, theta and a 40-character name
//   ref_vibebin restart  OUT   the records write_restart writes
//                              (src/io/field_io.cpp:207-227: __restart_info__
//                              [120, 24], then the 12 species of
//                              field_io.cpp:44-57 on their staggers) for a dry
//                              4x3x2 state, then rho_p, theta_p, pi_p
//   ref_vibebin restart13 OUT  the same without the three perturbations
//
// Content (exact binary fractions, so sums are exact in FP64):
//   value(q, i, j, k) = 1000 q + 100 k + 10 j + i + 0.5 (q mod 3)
// with the face at i = nx (u) and j = ny (v) repeating i = 0 / j = 0, and the
// reference state rho0 = 1 + k/4, theta0 = 300 + k/2, pi0 = 3/4 - k/16.
// field_io.cpp itself does not compile here (its logging macro names a
// member that does not exist), so the write_restart sequence is issued
// directly through BinaryWriter.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vibe/grid/field.hpp"
#include "vibe/io/binary_writer.hpp"

using vibe::Int;
using vibe::Real;
namespace grid = vibe::grid;

namespace {

constexpr Int NX = 4, NY = 3, NZ = 2;

grid::Grid make_grid() {
  grid::Geometry g;
  g.nx = NX;
  g.ny = NY;
  g.nz = NZ;
  g.dx = 1000.0;
  g.dy = 1000.0;
  g.z_top = 10000.0;
  g.zeta.resize(NZ + 1);
  for (Int k = 0; k <= NZ; ++k) g.zeta[k] = static_cast<Real>(k) / NZ;
  g.flat_terrain = true;
  g.name = "golden";
  return grid::Grid(g, grid::Decomposition::make(NX, NY, NZ, 1, 1, 2));
}

double value(int q, Int i, Int j, Int k) { return 1000.0 * q + 100.0 * k + 10.0 * j + i + 0.5 * (q % 3); }

double base(int q, Int k) {
  if (q == 3) return 1.0 + 0.25 * k;
  if (q == 4) return 300.0 + 0.5 * k;
  return 0.75 - 0.0625 * k;
}

vibe::io::FieldMeta meta(const std::string &name, grid::Stagger s) {
  vibe::io::FieldMeta m;
  m.name = name;
  m.stagger = s;
  return m;
}

// fills every point the writer emits (ni x nj x nk of the stagger)
void fill(grid::Field<Real> &f, grid::Stagger s, int q, bool total) {
  const Int ni = NX + (s == grid::Stagger::FaceX), nj = NY + (s == grid::Stagger::FaceY),
            nk = NZ + (s == grid::Stagger::FaceZ);
  for (Int k = 0; k < nk; ++k)
    for (Int j = 0; j < nj; ++j)
      for (Int i = 0; i < ni; ++i) {
        const Int ii = i == NX ? 0 : i, jj = j == NY ? 0 : j;
        double x = value(q, ii, jj, k);
        if (total) x = base(q, k) + x;
        f(i, j, k) = x;
      }
}

int generic(const char *out) {
  const grid::Grid g = make_grid();
  vibe::io::BinaryWriter w;
  w.open(out, g);
  w.write_time(3.5);
  grid::Field<Real> info(g, grid::Stagger::Cell, "__restart_info__");
  info.data()[0] = 120.0;
  info.data()[1] = 24.0;
  w.write(meta("__restart_info__", grid::Stagger::Cell), info);
  struct R {
    const char *name;
    grid::Stagger s;
    int q;
  };
  const std::string longname(40, 'a');
  const R recs[] = {{"u", grid::Stagger::FaceX, 0},    {"v", grid::Stagger::FaceY, 1},
                    {"w", grid::Stagger::FaceZ, 2},    {"zeta", grid::Stagger::FaceZ, 7},
                    {"theta", grid::Stagger::Cell, 4}, {longname.c_str(), grid::Stagger::Cell, 9}};
  for (const R &r : recs) {
    grid::Field<Real> f(g, r.s, r.name);
    fill(f, r.s, r.q, false);
    w.write(meta(r.name, r.s), f);
  }
  w.close();
  return 0;
}

int restart(const char *out, bool with_perturbations) {
  const grid::Grid g = make_grid();
  vibe::io::BinaryWriter w;
  w.open(out, g);
  w.write_time(120.0);
  grid::Field<Real> info(g, grid::Stagger::Cell, "__restart_info__");
  info.data()[0] = 120.0;
  info.data()[1] = 24.0;
  w.write(meta("__restart_info__", grid::Stagger::Cell), info);
  static const char *const names[12] = {"u", "v", "w", "rho", "theta", "pi", "qv", "qc", "qr", "qi", "qs", "qg"};
  for (int q = 0; q < 12; ++q) {
    const grid::Stagger s = q == 0 ? grid::Stagger::FaceX
                            : q == 1 ? grid::Stagger::FaceY
                            : q == 2 ? grid::Stagger::FaceZ
                                     : grid::Stagger::Cell;
    grid::Field<Real> f(g, s, names[q]);
    if (q < 6) fill(f, s, q, q >= 3);  // moisture stays zero
    w.write(meta(names[q], s), f);
  }
  if (with_perturbations) {
    static const char *const pn[3] = {"rho_p", "theta_p", "pi_p"};
    for (int q = 3; q < 6; ++q) {
      grid::Field<Real> f(g, grid::Stagger::Cell, pn[q - 3]);
      fill(f, grid::Stagger::Cell, q, false);
      w.write(meta(pn[q - 3], grid::Stagger::Cell), f);
    }
  }
  w.close();
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s generic|restart|restart13 OUT\n", argv[0]);
    return 2;
  }
  if (std::strcmp(argv[1], "generic") == 0) return generic(argv[2]);
  if (std::strcmp(argv[1], "restart") == 0) return restart(argv[2], true);
  if (std::strcmp(argv[1], "restart13") == 0) return restart(argv[2], false);
  std::fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 2;
}
