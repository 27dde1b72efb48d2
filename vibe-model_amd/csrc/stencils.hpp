// stencils.hpp -- per-point stencil formulas shared by the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace vibe_mi {

// ---------------------------------------------------------------------------
// WENO5 (c_kernels.c:64-104).
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ inline void weno5(T qm2, T qm1, T q0, T q1, T q2, T &ql, T &qr) {
  const T eps = T(1e-6);
  {
    const T p0 = (T(2) * qm2 - T(7) * qm1 + T(11) * q0) / T(6);
    const T p1 = (-qm1 + T(5) * q0 + T(2) * q1) / T(6);
    const T p2 = (T(2) * q0 + T(5) * q1 - q2) / T(6);
    const T s0a = qm2 - T(2) * qm1 + q0, s0b = qm2 - T(4) * qm1 + T(3) * q0;
    const T s1a = qm1 - T(2) * q0 + q1, s1b = qm1 - q1;
    const T s2a = q0 - T(2) * q1 + q2, s2b = T(3) * q0 - T(4) * q1 + q2;
    const T be0 = (T(13.0 / 12.0)) * s0a * s0a + T(0.25) * s0b * s0b;
    const T be1 = (T(13.0 / 12.0)) * s1a * s1a + T(0.25) * s1b * s1b;
    const T be2 = (T(13.0 / 12.0)) * s2a * s2a + T(0.25) * s2b * s2b;
    const T w0 = T(0.1) / ((eps + be0) * (eps + be0));
    const T w1 = T(0.6) / ((eps + be1) * (eps + be1));
    const T w2 = T(0.3) / ((eps + be2) * (eps + be2));
    ql = (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
  }
  {
    const T p0 = (T(2) * q2 - T(7) * q1 + T(11) * q0) / T(6);
    const T p1 = (-q1 + T(5) * q0 + T(2) * qm1) / T(6);
    const T p2 = (T(2) * q0 + T(5) * qm1 - qm2) / T(6);
    const T s0a = q2 - T(2) * q1 + q0, s0b = q2 - T(4) * q1 + T(3) * q0;
    const T s1a = q1 - T(2) * q0 + qm1, s1b = q1 - qm1;
    const T s2a = q0 - T(2) * qm1 + qm2, s2b = T(3) * q0 - T(4) * qm1 + qm2;
    const T be0 = (T(13.0 / 12.0)) * s0a * s0a + T(0.25) * s0b * s0b;
    const T be1 = (T(13.0 / 12.0)) * s1a * s1a + T(0.25) * s1b * s1b;
    const T be2 = (T(13.0 / 12.0)) * s2a * s2a + T(0.25) * s2b * s2b;
    const T w0 = T(0.1) / ((eps + be0) * (eps + be0));
    const T w1 = T(0.6) / ((eps + be1) * (eps + be1));
    const T w2 = T(0.3) / ((eps + be2) * (eps + be2));
    qr = (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
  }
}

// 5th-order upwind face interpolation (Wicker & Skamarock 2002) for the face
// between q0 and q1: flux = U * c6 - |U| * d5.
template <typename T>
__host__ __device__ inline T c6(T qm2, T qm1, T q0, T q1, T q2, T q3) {
  return (T(37) * (q0 + q1) - T(8) * (qm1 + q2) + (qm2 + q3)) / T(60);
}
template <typename T>
__host__ __device__ inline T d5(T qm2, T qm1, T q0, T q1, T q2, T q3) {
  return (T(10) * (q1 - q0) - T(5) * (q2 - qm1) + (q3 - qm2)) / T(60);
}
template <typename T>
__host__ __device__ inline T up5(T U, T qm2, T qm1, T q0, T q1, T q2, T q3) {
  return U * c6<T>(qm2, qm1, q0, q1, q2, q3) - fabs(U) * d5<T>(qm2, qm1, q0, q1, q2, q3);
}

// WENO5 halves: left-biased value at the face right of q0, and the mirrored
// (right-biased) value at the face left of q0.
template <typename T>
__host__ __device__ inline T weno_left(T qm2, T qm1, T q0, T q1, T q2) {
  T l, r;
  weno5<T>(qm2, qm1, q0, q1, q2, l, r);
  return l;
}
template <typename T>
__host__ __device__ inline T weno_right(T qm2, T qm1, T q0, T q1, T q2) {
  T l, r;
  weno5<T>(qm2, qm1, q0, q1, q2, l, r);
  return r;
}
// Local Lax-Friedrichs face flux, alpha = max(|U|, 1e-3)
// (src/dyn/advection.cpp:151-156 of the reference).
template <typename T>
__host__ __device__ inline T llf(T U, T qm, T qp) {
  const T a = fmax(fabs(U), T(1e-3));
  return T(0.5) * (U * (qm + qp) - a * (qp - qm));
}

}  // namespace vibe_mi
