// common.hpp -- shared host/device helpers of the vibe_mi HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "vibe_mi/c_kernels.h"

namespace vibe_mi {

// Physical constants (include/vibe/common/constants.hpp:16-26 of the reference).
constexpr double kG = 9.80665;
constexpr double kRd = 287.05;
constexpr double kCp = 1004.64;
constexpr double kCv = 717.63;

// Per-thread error slot read by vibe_mi_last_error().
void set_last_error(int code);
int last_error();

#define VMI_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) {                             \
      ::vibe_mi::set_last_error(VIBE_MI_ERR_HIP);       \
      return VIBE_MI_ERR_HIP;                           \
    }                                                   \
  } while (0)

inline unsigned ceil_div(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Error-free transformation s + e = a + b (Knuth TwoSum).
template <typename T>
__host__ __device__ inline void two_sum(T a, T b, T &s, T &e) {
  s = a + b;
  const T bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

// Compensated accumulator: value = s + c.  add() is Neumaier's step
// (c_kernels.c:190-206); merge() combines two accumulators exactly up to the
// rounding of the compensation terms.
struct NSum {
  double s, c;
  __host__ __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __host__ __device__ void merge(const NSum &o) {
    double t, e;
    two_sum(s, o.s, t, e);
    s = t;
    c = c + o.c + e;
  }
};

}  // namespace vibe_mi
