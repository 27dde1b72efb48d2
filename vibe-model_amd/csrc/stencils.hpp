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
  return (T(37) * (q0 + q1) - T(8) * (qm1 + q2) + (qm2 + q3)) * T(1.0 / 60.0);
}
template <typename T>
__host__ __device__ inline T d5(T qm2, T qm1, T q0, T q1, T q2, T q3) {
  return (T(10) * (q1 - q0) - T(5) * (q2 - qm1) + (q3 - qm2)) * T(1.0 / 60.0);
}
template <typename T>
__host__ __device__ inline T up5(T U, T qm2, T qm1, T q0, T q1, T q2, T q3) {
  return U * c6<T>(qm2, qm1, q0, q1, q2, q3) - fabs(U) * d5<T>(qm2, qm1, q0, q1, q2, q3);
}

// Quotient num / den: correctly rounded in FP64 (the oracle's arithmetic);
// in FP32 the hardware reciprocal (1 ulp) times num, which replaces the
// ten-instruction correctly rounded division sequence.
template <typename T>
__device__ __forceinline__ T quot(T num, T den) {
  if constexpr (sizeof(T) == 4)
    return num * __builtin_amdgcn_rcpf(den);
  else
    return num / den;
}

// WENO5 value at the face right of c from the stencil (a, b, c, d, e), same
// weights as weno5() (d = 0.1, 0.6, 0.3; eps = 1e-6) but with one division:
// alpha_k = d_k/(eps+beta_k)^2 scaled by prod_j (eps+beta_j)^2.
template <typename T>
__device__ inline T weno_face(T a, T b, T c, T d, T e) {
  const T eps = T(1e-6);
  const T p0 = T(2) * a - T(7) * b + T(11) * c;
  const T p1 = -b + T(5) * c + T(2) * d;
  const T p2 = T(2) * c + T(5) * d - e;
  const T s0a = a - T(2) * b + c, s0b = a - T(4) * b + T(3) * c;
  const T s1a = b - T(2) * c + d, s1b = b - d;
  const T s2a = c - T(2) * d + e, s2b = T(3) * c - T(4) * d + e;
  const T e0 = eps + ((T(13.0 / 12.0)) * s0a * s0a + T(0.25) * s0b * s0b);
  const T e1 = eps + ((T(13.0 / 12.0)) * s1a * s1a + T(0.25) * s1b * s1b);
  const T e2 = eps + ((T(13.0 / 12.0)) * s2a * s2a + T(0.25) * s2b * s2b);
  const T q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const T w0 = T(0.1) * (q1 * q2), w1 = T(0.6) * (q0 * q2), w2 = T(0.3) * (q0 * q1);
  return quot<T>(w0 * p0 + w1 * p1 + w2 * p2, T(6) * (w0 + w1 + w2));
}

// Upwind WENO5 flux U * q through the face between cells f-1 and f from the
// six cells q[f-3..f+2]: U >= 0 uses the left-biased value, U < 0 the mirrored
// right-biased one.  The stencil is selected per lane, so a wavefront always
// evaluates exactly one reconstruction per face.
template <typename T>
__device__ inline T upwind_weno(T U, T qm3, T qm2, T qm1, T q0, T q1, T q2) {
  const bool pos = U >= T(0);
  const T a = pos ? qm3 : q2, b = pos ? qm2 : q1, c = pos ? qm1 : q0, d = pos ? q0 : qm1, e = pos ? q1 : qm2;
  return U * weno_face<T>(a, b, c, d, e);
}

}  // namespace vibe_mi
