// helmholtz_dev.hpp -- device-array forms of the Helmholtz solver for callers
// inside the library (the semi-implicit step).  Both calls are synchronous
// with respect to the host: they order their work on the handle's own stream
// and return when it is done, so the caller finishes writing a, b or f on its
// stream before the call.
#pragma once

#include "vibe_mi/helmholtz.h"

namespace vibe_mi {

// set_coefficients from device cell arrays (no validation: the caller forms
// a >= 0 and b > 0 itself).
int helmholtz_set_coefficients_device(vibe_mi_helmholtz *h, const double *a, const double *b);
// vibe_mi_helmholtz_solve with device arrays f and x.
int helmholtz_solve_device(vibe_mi_helmholtz *h, const double *f, double *x, int kind, double tol, int max_iter,
                           int *iterations, double *rel_res);

}  // namespace vibe_mi
