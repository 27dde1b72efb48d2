// literal_dev.hpp -- device-level access to a vibe_mi_literal handle for
// callers inside the library (the semi-implicit step): its grid, reference
// state and the device copies of the 12-species state and tendency.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "vibe_mi/literal.h"

namespace vibe_mi {

struct LiteralView {
  int nx, ny, nz;
  double dx, dy, ztop;
  const double *zeta, *zs;             // device; zs NULL: flat
  const double *rho0, *theta0, *pi0;   // device cell arrays
  double *s[12], *t[12];               // device state and tendency, species order of literal.h
  size_t n[12];                        // elements of each species
  hipStream_t stream;                  // the handle's stream
};

LiteralView literal_view(vibe_mi_literal *h);
// Equations::tendencies of the device state s into t, enqueued on the
// handle's stream (no transfers).
int literal_tendencies_device(vibe_mi_literal *h);

}  // namespace vibe_mi
