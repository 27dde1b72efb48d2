// geo.hpp -- device view of one domain's geometry and reference state.
//
// Column fields (terrain depth, slopes, ...) and the 3-D reference state are
// stored compacted: a direction in which neither the terrain nor the
// reference state varies is stored once.  Flat terrain gives 1-D (per-level)
// reference arrays, a y-uniform ridge (x, z) arrays, general terrain full
// 3-D arrays.  Compacted arrays stay resident in L2 across a launch, so the
// reference state costs no HBM traffic on flat and ridge grids.
#pragma once

#include <hip/hip_runtime.h>

namespace vibe_mi {

template <typename T>
struct Geo {
  int nx, ny, nz;
  int hi, hj;  // column index = j*hj + i*hi (0 where the domain is uniform)
  int rk;      // 3-D compact index = k*rk + j*hj + i*hi
  T rdx, rdy;
  const T *zeta;    // nz+1 interfaces
  const T *zc;      // nz   cell-centre zeta
  const T *rdzeta;  // nz   1/(zeta[k+1]-zeta[k])
  const T *rdzw;    // nz+1 1/(zc[k]-zc[k-1]) for 1..nz-1, 0 elsewhere
  const T *Du, *Dv, *sx, *sy, *rDc, *rDu, *rDv, *metx, *mety;  // columns
  const T *rho0, *theta0, *kap, *lam;                           // 3-D reference
  // face and interface products of the reference state (3-D compact, like
  // rho0): cthu = cp (theta0_{i-1} + theta0_i)/2 and rhu = Du (rho0_{i-1} +
  // rho0_i)/2 at u faces, cthv / rhv at v faces, r0w = (rho0_{k-1} + rho0_k)/2
  // at interfaces (index k*rk + column, 0 at k = 0 and nz)
  const T *cthu, *cthv, *rhu, *rhv, *r0w;
  // differences of the face products across a cell, formed in FP64:
  // drhu = rhu(east face) - rhu(west face), drhv = rhv(north) - rhv(south)
  // (the FP32 substep forms its mass-flux divergence from them)
  const T *drhu, *drhv;
  // FP32-mixed only: per-level background x-wind (the level mean of the
  // state set on the domain); the FP32 acoustic buffers hold u - ubg, whose
  // rounding error scales with the deviation instead of the full wind
  const T *ubg;
  const double *ubg64;
};

// Layout of one HEVI table slot (a stage's factorisation for one dtau):
// per = (nz-1)*rk entries each of beta, 1/pivot, c' and gk (interfaces
// 1..nz-1), then nz*rk entries of ck (levels), all [level][compact column].
//   gk = dtau/D cp theta0_w / dzw   (implicit vertical PGF coefficient)
//   ck = kappa dtau/D / dzeta       (vertical flux-divergence coefficient)
__host__ __device__ inline size_t hevi_per(int nz, int rk) { return static_cast<size_t>(nz - 1) * rk; }
__host__ __device__ inline size_t hevi_slot_elems(int nz, int rk) {
  return 4 * hevi_per(nz, rk) + static_cast<size_t>(nz) * rk;
}
template <typename T>
struct HeviTab {
  const T *A, *INV, *CP, *GK, *CK;
  __host__ __device__ static HeviTab at(const T *base, int nz, int rk) {
    const size_t per = hevi_per(nz, rk);
    return HeviTab{base, base + per, base + 2 * per, base + 3 * per, base + 4 * per};
  }
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampk(int k, int hi) { return k < 0 ? 0 : (k > hi ? hi : k); }

// Index helpers for the dense state arrays ([k][j][i], periodic i and j).
struct Acc {
  int nx, ny;
  size_t nxy;
  __device__ __forceinline__ size_t h(int i, int j) const {
    return static_cast<size_t>(wrapi(j, ny)) * nx + wrapi(i, nx);
  }
  __device__ __forceinline__ size_t at(int i, int j, int k) const { return static_cast<size_t>(k) * nxy + h(i, j); }
};

// Global-address-space views of a pointer.  Pointers read from device
// memory (the batched kernels' member descriptors) carry no address space,
// so accesses through them would be generic (flat) instructions; the views
// make every access a global load/store with scalar base and vector offset.
#define VMI_GLOBAL __attribute__((address_space(1)))
template <typename P>
__device__ __forceinline__ const VMI_GLOBAL P *G(const P *p) {
  return (const VMI_GLOBAL P *)p;
}
template <typename P>
__device__ __forceinline__ VMI_GLOBAL P *G(P *p) {
  return (VMI_GLOBAL P *)p;
}

// Constant-address-space view for arrays no kernel writes while it reads
// them (geometry, reference state, HEVI factorisation): wave-uniform
// accesses through it become scalar loads.
#define VMI_CONST __attribute__((address_space(4)))
template <typename P>
__device__ __forceinline__ const VMI_CONST P *C(const P *p) {
  return (const VMI_CONST P *)p;
}

// Loads/stores addressed as (wave-uniform base of a level, 32-bit per-lane
// element index): the base lives in scalar registers and the byte offset fits
// 32 bits, so the compiler emits the scalar-base + 32-bit vector-offset form
// instead of 64-bit per-lane address arithmetic.  A level of one field must
// stay below 4 GiB (nx*ny*8 bytes), far above any single-GPU domain here.
template <typename P>
__device__ __forceinline__ P ld32(const P *__restrict__ f, size_t base, uint32_t idx) {
  const VMI_GLOBAL char *b = (const VMI_GLOBAL char *)(f + base);
  return *(const VMI_GLOBAL P *)(b + idx * static_cast<uint32_t>(sizeof(P)));
}
template <typename P>
__device__ __forceinline__ P ldc32(const P *__restrict__ f, size_t base, uint32_t idx) {
  const VMI_CONST char *b = (const VMI_CONST char *)(f + base);
  return *(const VMI_CONST P *)(b + idx * static_cast<uint32_t>(sizeof(P)));
}
template <typename P>
__device__ __forceinline__ void st32(P *__restrict__ f, size_t base, uint32_t idx, P v) {
  VMI_GLOBAL char *b = (VMI_GLOBAL char *)(f + base);
  *(VMI_GLOBAL P *)(b + idx * static_cast<uint32_t>(sizeof(P))) = v;
}

// Loads/stores at (array base, 32-bit byte offset).  The callers form the
// offset as (per-lane column offset + wave-uniform level offset) in 32-bit
// arithmetic next to the access, so the compiler emits the scalar-base +
// 32-bit vector-offset form (one 32-bit add per distinct offset, no 64-bit
// per-lane address).  An array addressed this way stays below 4 GiB
// (domain creation checks it).
template <typename P>
__device__ __forceinline__ P gld(const P *__restrict__ f, uint32_t off) {
  return *(const VMI_GLOBAL P *)((const VMI_GLOBAL char *)f + off);
}
// the same through the constant-address-space view (read-only arrays: a
// wave-uniform offset becomes a scalar load)
template <typename P>
__device__ __forceinline__ P gldc(const P *__restrict__ f, uint32_t off) {
  return *(const VMI_CONST P *)((const VMI_CONST char *)f + off);
}
template <typename P>
__device__ __forceinline__ void gst(P *__restrict__ f, uint32_t off, P v) {
  *(VMI_GLOBAL P *)((VMI_GLOBAL char *)f + off) = v;
}

template <typename T>
__device__ __forceinline__ int gcol(const Geo<T> &g, int i, int j) {
  return wrapi(j, g.ny) * g.hj + wrapi(i, g.nx) * g.hi;
}
// Geometry compaction fixed at compile time (kernels specialised on it):
//   kGeoFlat  terrain and reference state uniform in x and y (rk = 1): every
//             slope and metric term is zero, every geometry load is uniform;
//   kGeoX     uniform in y (hj = 0; rk = nx, or 1 for flat members): the
//             y slope and metric terms are zero;
//   kGeoAny   general (runtime hi, hj).
constexpr int kGeoFlat = 0, kGeoX = 1, kGeoAny = 2;
template <int GM, typename T>
__device__ __forceinline__ int gcolm(const Geo<T> &g, int i, int j) {
  if constexpr (GM == kGeoFlat) {
    return 0;
  } else if constexpr (GM == kGeoX) {
    return wrapi(i, g.nx) * g.hi;
  } else {
    return gcol(g, i, j);
  }
}

template <typename T>
__device__ __forceinline__ int gref(const Geo<T> &g, int i, int j, int k) {
  return k * g.rk + gcol(g, i, j);
}

}  // namespace vibe_mi
