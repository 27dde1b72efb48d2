// ckernels.hip -- the reference's C11 hot-kernel ABI (c_kernels.h) on gfx950.
//
// Every vibe_*_f64 entry point keeps the reference's signature and semantics
// (src/common/c_kernels.c, cited per kernel) and computes on the GPU.  The
// contract of c_kernels.h:13-19 holds: no entry point allocates (no hipMalloc,
// hipFree, stream creation or host heap), and none keeps state that changes
// its results.  Arrays the GPU can dereference (device memory of the current
// GPU, hipHostMalloc'd host memory) are used in place; pageable host arrays
// are copied in chunks through a statically sized device array of the
// library, one per GPU, taken under a per-GPU lock for the length of a call
// (concurrent calls on one GPU serialise and stay correct).  Work runs on
// hipStreamPerThread.  The vibe_mi_*_batched entry points take device
// pointers and a stream and use the caller's workspace.
//
// Built with -ffp-contract=off: per-point arithmetic is the same IEEE
// sequence as the C code, so WENO5, the CFL scan, flux advection and the
// bilinear stencil are bit-identical to the CPU; Thomas differs from it only
// in the (documented) corrected recurrence; the Neumaier sum is a fixed-order
// tree and agrees to a few ulp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "stencils.hpp"

namespace vibe_mi {

namespace {
// errno-like slot of the calling thread: how the reference's void entry
// points (weno5, flux advection) report a device failure
thread_local int t_last_error = 0;
}  // namespace

void set_last_error(int code) { t_last_error = code; }
int last_error() { return t_last_error; }

constexpr int kMaxDevices = 64;

// ---------------------------------------------------------------------------
// Thomas: one system per lane, interleaved k-major layout.  Contract of
// c_kernels.h:38-55; recurrence of helmholtz.cpp:57-86; exact-zero pivot
// check as c_kernels.c:34/:40.  The forward sweep stores c'_k for k < n-1
// only, so the caller's workspace row n-1 is free for the singular flag.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_thomas(const T *__restrict__ a, const T *__restrict__ b,
                                                const T *__restrict__ c, const T *__restrict__ d,
                                                int n, int nsys, T *__restrict__ x,
                                                T *__restrict__ ws, int *__restrict__ flag) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const size_t ns = static_cast<size_t>(nsys);
  T denom = b[s];
  if (denom == T(0)) {
    atomicOr(flag, 1);
    return;
  }
  if (n == 1) {
    x[s] = d[s] / denom;
    return;
  }
  T cp = (c != nullptr ? c[s] : T(0)) / denom;
  T dp = d[s] / denom;
  ws[s] = cp;
  x[s] = dp;
  for (int k = 1; k < n; ++k) {
    const size_t o = static_cast<size_t>(k) * ns + s;
    const T am = (a != nullptr) ? a[o - ns] : T(0);
    denom = b[o] - am * cp;
    if (denom == T(0)) {
      atomicOr(flag, 1);
      return;
    }
    if (k < n - 1) {
      cp = (c != nullptr ? c[o] : T(0)) / denom;
      ws[o] = cp;
    }
    dp = (d[o] - am * dp) / denom;
    x[o] = dp;
  }
  T xn = dp;
  for (int k = n - 2; k >= 0; --k) {
    const size_t o = static_cast<size_t>(k) * ns + s;
    xn = x[o] - ws[o] * xn;
    x[o] = xn;
  }
}

// The singular flag of a Thomas launch: row n-1 of the workspace (never
// written by k_thomas).
template <typename T>
int *thomas_flag(T *ws, int n, int nsys) {
  return reinterpret_cast<int *>(ws + static_cast<size_t>(n - 1) * static_cast<size_t>(nsys));
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation batched Thomas entry point: one statically
// sized device array per GPU.  Bytes [0, 256) hold the singular flag, the rest
// the forward-sweep workspace c' (n * nsys values).  In host code the array's
// name is only a stand-in for the symbol: its address on each device is taken
// with hipGetSymbolAddress, once per device with that device current, and that
// pointer is what host code passes to kernels and copies.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr size_t kScratchHead = 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
int scratch_address(void **out) {
  int dev = 0;
  VMI_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return VIBE_C_ERR_BAD_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    VMI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return VIBE_MI_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Clears the flag with a kernel (not a memset node) so the enqueue sequence
// stays capturable.
__global__ void k_clear_flag(int *flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

// ---------------------------------------------------------------------------
// Staging of the host-pointer entry points: a second statically sized device
// array per GPU (256 MiB), laid out as a 64 KiB head (flags, reduction
// results, partial sums) and the body that pageable inputs are copied
// through.  Same symbol-address rule as `scratch`; a per-GPU lock is held for
// the length of a call.
// ---------------------------------------------------------------------------
constexpr size_t kStageBytes = size_t(256) << 20;
constexpr size_t kStageHead = size_t(64) << 10;
constexpr size_t kStageBody = kStageBytes - kStageHead;
static __device__ __align__(256) unsigned char staging[kStageBytes];

struct StageSlot {
  std::mutex mu;
  void *base = nullptr;
};
StageSlot g_stage[kMaxDevices];

class Stage {
 public:
  // Locks the current GPU's staging array; 0 or an error code.
  int open() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return VIBE_MI_ERR_NO_DEVICE;
    VMI_HIP(hipGetDevice(&dev_));
    if (dev_ < 0 || dev_ >= kMaxDevices) return VIBE_C_ERR_BAD_ARG;
    lock_ = std::unique_lock<std::mutex>(g_stage[dev_].mu);
    if (g_stage[dev_].base == nullptr) {
      void *p = nullptr;
      VMI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(staging)));
      if (p == nullptr) return VIBE_MI_ERR_HIP;
      g_stage[dev_].base = p;
    }
    base_ = static_cast<unsigned char *>(g_stage[dev_].base);
    return 0;
  }
  unsigned char *head() const { return base_; }
  unsigned char *body() const { return base_ + kStageHead; }
  int device() const { return dev_; }

 private:
  std::unique_lock<std::mutex> lock_;
  unsigned char *base_ = nullptr;
  int dev_ = 0;
};

// Where the current GPU dereferences p, or nullptr when it cannot (pageable
// host memory, memory of another GPU).  Pinned host memory is used in place
// only where its device address is its host address (hipHostMalloc on this
// platform); anything else is staged.
template <typename T>
T *gpu_view(T *p, int dev) {
  if (p == nullptr) return nullptr;
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory: the query fails; clear its error
    return nullptr;
  }
  if (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged) return at.device == dev ? p : nullptr;
  if (at.type == hipMemoryTypeHost && at.devicePointer != nullptr && at.hostPointer != nullptr) {
    const char *dp = static_cast<const char *>(at.devicePointer) +
                     (reinterpret_cast<const char *>(p) - static_cast<const char *>(at.hostPointer));
    return dp == reinterpret_cast<const char *>(p) ? p : nullptr;
  }
  return nullptr;
}

// Copies between host and staging; hipMemcpyDefault lets the runtime tell
// pageable, pinned and device memory (of any GPU) apart.
#define VMI_COPY(dst, src, bytes) VMI_HIP(hipMemcpyAsync((dst), (src), (bytes), hipMemcpyDefault, hipStreamPerThread))

// Small staged calls: the host inputs are gathered into one block on the
// caller's stack (at 256-byte offsets) and sent with one copy.
struct Pack {
  static constexpr size_t kMax = size_t(48) << 10;
  alignas(256) unsigned char buf[kMax];
  size_t used = 0;
  static constexpr size_t align(size_t o) { return (o + 255) / 256 * 256; }
  // offset of bytes from src (zeros when src is NULL)
  size_t put(const void *src, size_t bytes) {
    const size_t o = align(used);
    if (src != nullptr)
      std::memcpy(buf + o, src, bytes);
    else
      std::memset(buf + o, 0, bytes);
    used = o + bytes;
    return o;
  }
};

// ---------------------------------------------------------------------------
// WENO5 (c_kernels.c:64-104): batched over stencils, and the scalar form on
// its by-value arguments.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_weno(const T *__restrict__ q, size_t n, T *__restrict__ ql,
                                              T *__restrict__ qr) {
  const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const T *p = q + 5 * s;
  T l, r;
  weno5<T>(p[0], p[1], p[2], p[3], p[4], l, r);
  ql[s] = l;
  qr[s] = r;
}

template <typename T>
__global__ void k_weno_scalar(T qm2, T qm1, T q0, T q1, T q2, T *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  T l, r;
  weno5<T>(qm2, qm1, q0, q1, q2, l, r);
  out[0] = l;
  out[1] = r;
}

// ---------------------------------------------------------------------------
// Bilinear (c_kernels.c:108-144).  bilinear_weights is the index and weight
// step, bilinear_mix the interpolation of the four corner values.
// ---------------------------------------------------------------------------
struct BilinearCorners {
  int ia, ib, ja, jb;
  double wx, wy;
};

__host__ __device__ inline BilinearCorners bilinear_weights(int halo, int nsx, int nsy, double xi, double yj) {
  const int nxi = nsx - 2 * halo, nyi = nsy - 2 * halo;
  const int i0 = static_cast<int>(floor(xi)), j0 = static_cast<int>(floor(yj));
  BilinearCorners c;
  c.wx = xi - static_cast<double>(i0);
  c.wy = yj - static_cast<double>(j0);
  auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  c.ia = cl(i0, nxi - 1);
  c.ib = cl(i0 + 1, nxi - 1);
  c.ja = cl(j0, nyi - 1);
  c.jb = cl(j0 + 1, nyi - 1);
  return c;
}

__device__ inline double bilinear_mix(double q00, double q10, double q01, double q11, double wx, double wy) {
  const double lo = q00 + wx * (q10 - q00);
  const double hi = q01 + wx * (q11 - q01);
  return lo + wy * (hi - lo);
}

// f: the halo'd field, readable by the GPU; level k (clamped by the host).
__global__ void k_bilinear(const double *__restrict__ f, int halo, int nsx, int nsy, int k, double xi,
                           double yj, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const BilinearCorners c = bilinear_weights(halo, nsx, nsy, xi, yj);
  const double *plane = f + static_cast<size_t>(k + halo) * nsx * nsy;
  auto at = [&](int i, int j) { return plane[static_cast<size_t>(j + halo) * nsx + (i + halo)]; };
  *out = bilinear_mix(at(c.ia, c.ja), at(c.ib, c.ja), at(c.ia, c.jb), at(c.ib, c.jb), c.wx, c.wy);
}

// Pageable field: the host hands over the four corner values it selected with
// bilinear_weights; the weights are formed again here from (xi, yj).
__global__ void k_bilinear_corners(int halo, int nsx, int nsy, double xi, double yj, double q00, double q10,
                                   double q01, double q11, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const BilinearCorners c = bilinear_weights(halo, nsx, nsy, xi, yj);
  *out = bilinear_mix(q00, q10, q01, q11, c.wx, c.wy);
}

// ---------------------------------------------------------------------------
// CFL scan (c_kernels.c:148-187): per-point values identical to the C code;
// maxima/minima of non-negative doubles combined through their IEEE bit
// patterns (monotonic as unsigned integers), so the result is exact and does
// not depend on how the points are split into launches.  Point t of the
// launch is global point beg + t; zs_mode 0: no terrain, 1: zs[column],
// 2: zs[t] (staged in point order).
// ---------------------------------------------------------------------------
__device__ inline void atomic_max_pos(unsigned long long *p, double v) {
  atomicMax(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}
__device__ inline void atomic_min_pos(unsigned long long *p, double v) {
  atomicMin(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}

__global__ void k_cfl_init(unsigned long long *res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  res[0] = 0ull;
  res[1] = 0ull;
  res[2] = static_cast<unsigned long long>(__double_as_longlong(1e30));
}

__global__ void __launch_bounds__(256) k_cfl(const double *__restrict__ u, const double *__restrict__ v,
                                             const double *__restrict__ w, const double *__restrict__ dzeta,
                                             const double *__restrict__ zs, int zs_mode, size_t nxy,
                                             size_t beg, size_t len, double dx, double dy, double z_top,
                                             double dt, unsigned long long *__restrict__ res) {
  double worst = 0.0, wind = 0.0, mdz = 1e30;
  for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < len;
       t += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t idx = beg + t, k = idx / nxy, col = idx - k * nxy;
    const double scale = zs_mode == 0 ? z_top : z_top - (zs_mode == 1 ? zs[col] : zs[t]);
    double dz = dzeta[k] * scale;
    if (dz < 1.0) dz = 1.0;
    if (dz < mdz) mdz = dz;
    const double uu = u[t], vv = v[t], ww = w[t];
    const double speed = sqrt(uu * uu + vv * vv + ww * ww);
    if (speed > wind) wind = speed;
    const double cfl = dt * (fabs(uu) / dx + fabs(vv) / dy + fabs(ww) / dz);
    if (cfl > worst) worst = cfl;
  }
  __shared__ double sw[256], sv[256], sz[256];
  sw[threadIdx.x] = worst;
  sv[threadIdx.x] = wind;
  sz[threadIdx.x] = mdz;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(h)) {
      sw[threadIdx.x] = fmax(sw[threadIdx.x], sw[threadIdx.x + h]);
      sv[threadIdx.x] = fmax(sv[threadIdx.x], sv[threadIdx.x + h]);
      sz[threadIdx.x] = fmin(sz[threadIdx.x], sz[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomic_max_pos(&res[0], sw[0]);
    atomic_max_pos(&res[1], sv[0]);
    atomic_min_pos(&res[2], sz[0]);
  }
}

// ---------------------------------------------------------------------------
// Neumaier sum (c_kernels.c:190-206) as a fixed-order reduction: the input is
// cut into chunks of kSumChunk values; in each chunk block b sums the
// contiguous sub-chunk b with per-lane Neumaier accumulators over a fixed
// stride and merges lanes by a fixed TwoSum tree, one block merges the block
// partials into the chunk's accumulator, and one block merges the chunk
// accumulators in order.  The order depends on n only, so every entry point
// (in place or staged) gives the same bits.
// ---------------------------------------------------------------------------
constexpr int kSumThreads = 256;
constexpr int kSumBlocks = 1024;
constexpr size_t kSumChunk = size_t(1) << 23;
// head layout: [0, 256) results, [256, 256 + kSumBlocks NSum) block
// partials, then the chunk accumulators
constexpr size_t kSumPartOff = 256;
constexpr size_t kSumSlotOff = kSumPartOff + kSumBlocks * sizeof(NSum);
constexpr size_t kSumMaxChunks = (kStageHead - kSumSlotOff) / sizeof(NSum);

template <typename T>
__global__ void __launch_bounds__(kSumThreads) k_nsum_partial(const T *__restrict__ x, size_t n,
                                                              size_t chunk, NSum *__restrict__ part) {
  const size_t beg = static_cast<size_t>(blockIdx.x) * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  NSum acc{0.0, 0.0};
  for (size_t i = beg + threadIdx.x; i < end; i += kSumThreads) acc.add(static_cast<double>(x[i]));
  __shared__ NSum sh[kSumThreads];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kSumThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(h)) sh[threadIdx.x].merge(sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Merges part[0..nparts) by the fixed tree into *slot; with total != nullptr
// also writes slot's value s + c there.
__global__ void __launch_bounds__(kSumThreads) k_nsum_merge(const NSum *__restrict__ part, int nparts,
                                                            NSum *__restrict__ slot, double *__restrict__ total) {
  __shared__ NSum sh[kSumThreads];
  NSum acc{0.0, 0.0};
  for (int p = threadIdx.x; p < nparts; p += kSumThreads) acc.merge(part[p]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kSumThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(h)) sh[threadIdx.x].merge(sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *slot = sh[0];
    if (total != nullptr) *total = sh[0].s + sh[0].c;
  }
}

// ---------------------------------------------------------------------------
// Flux-form scalar advection (c_kernels.c:209-286).  The launch covers levels
// [k0, k0 + gridDim.z); q, rho, u, v, w hold levels from kin (the staged slab:
// index c - kin*nxy), dqdt levels from kout.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_flux_adv(const double *__restrict__ q, const double *__restrict__ rho,
                                                 const double *__restrict__ u, const double *__restrict__ v,
                                                 const double *__restrict__ w,
                                                 const double *__restrict__ dzeta,
                                                 const double *__restrict__ zs, int nx, int ny, int nz,
                                                 double dx, double dy, double z_top, int k0, int kin, int kout,
                                                 double *__restrict__ dqdt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = k0 + static_cast<int>(blockIdx.z);
  if (i >= nx) return;
  const size_t nxy = static_cast<size_t>(nx) * ny;
  const double rdx = 1.0 / dx, rdy = 1.0 / dy;
  const size_t col = static_cast<size_t>(j) * nx + i;
  const double scale = zs != nullptr ? z_top - zs[col] : z_top;
  double dz = dzeta[k] * scale;
  if (dz < 1.0) dz = 1.0;
  const size_t base = static_cast<size_t>(k - kin) * nxy;
  const size_t c = base + col;
  const size_t e = base + static_cast<size_t>(j) * nx + static_cast<size_t>((i + 1) % nx);
  const size_t wv = base + static_cast<size_t>(j) * nx + static_cast<size_t>((i + nx - 1) % nx);
  const size_t nn = base + static_cast<size_t>((j + 1) % ny) * nx + i;
  const size_t ss = base + static_cast<size_t>((j + ny - 1) % ny) * nx + i;
  const double fxe = u[e] * (0.5 * (rho[c] + rho[e])) * (0.5 * (q[c] + q[e]));
  const double fxw = u[c] * (0.5 * (rho[wv] + rho[c])) * (0.5 * (q[wv] + q[c]));
  const double fyn = v[nn] * (0.5 * (rho[c] + rho[nn])) * (0.5 * (q[c] + q[nn]));
  const double fys = v[c] * (0.5 * (rho[ss] + rho[c])) * (0.5 * (q[ss] + q[c]));
  double fzt = 0.0, fzb = 0.0;
  if (k > 0 && k + 1 < nz) {
    const size_t t = c + nxy, b = c - nxy;
    fzt = w[t] * (0.5 * (rho[c] + rho[t])) * (0.5 * (q[c] + q[t]));
    fzb = w[c] * (0.5 * (rho[b] + rho[c])) * (0.5 * (q[b] + q[c]));
  }
  const double div = (fxe - fxw) * rdx + (fyn - fys) * rdy + (fzt - fzb) / dz;
  const double rr = rho[c] > 1e-12 ? rho[c] : 1e-12;
  dqdt[static_cast<size_t>(k - kout) * nxy + col] = -div / rr;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
int thomas_launch(const T *a, const T *b, const T *c, const T *d, int n, int nsys, T *x, T *ws, int *flag,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, flag);
  hipLaunchKernelGGL(k_thomas<T>, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, a, b, c, d, n, nsys, x, ws, flag);
  VMI_HIP(hipGetLastError());
  return 0;
}

// c_kernels.h:38-55 for one system of host or device arrays.
template <typename T>
int thomas_one(const T *a, const T *b, const T *c, const T *d, int n, T *x, T *ws) {
  if (n <= 0 || b == nullptr || d == nullptr || x == nullptr || ws == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (n == 1) a = c = nullptr;  // a and c have n - 1 = 0 entries
  Stage sg;
  if (int rc = sg.open()) return rc;
  const int dev = sg.device();
  hipStream_t st = hipStreamPerThread;
  const size_t nn = static_cast<size_t>(n), nm = n > 1 ? nn - 1 : 0;
  const T *ga = gpu_view(a, dev), *gb = gpu_view(b, dev), *gc = gpu_view(c, dev), *gd = gpu_view(d, dev);
  T *gx = gpu_view(x, dev), *gw = gpu_view(ws, dev);
  const bool in_place = (a == nullptr || ga) && gb && (c == nullptr || gc) && gd && gx && gw;
  int flag = 0;
  if (in_place) {
    int *dflag = thomas_flag(gw, n, 1);
    if (int rc = thomas_launch<T>(ga, gb, gc, gd, n, 1, gx, gw, dflag, st)) return rc;
    VMI_COPY(&flag, dflag, sizeof(int));
    VMI_HIP(hipStreamSynchronize(st));
    return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
  }
  const bool use_a = a != nullptr, use_c = c != nullptr;
  if ((2 * nm + 3 * nn + 1) * sizeof(T) + 5 * 256 <= Pack::kMax) {
    // one copy in, one out: body = [x | flag] [a] [c] [b] [d] [ws], the
    // flag zeroed by the copy
    Pack pk;
    const size_t ox = pk.put(nullptr, nn * sizeof(T));
    const size_t of = pk.put(nullptr, sizeof(T));
    const size_t oa = use_a ? pk.put(a, nm * sizeof(T)) : 0;
    const size_t oc = use_c ? pk.put(c, nm * sizeof(T)) : 0;
    const size_t ob = pk.put(b, nn * sizeof(T)), od = pk.put(d, nn * sizeof(T));
    const size_t ow = Pack::align(pk.used);
    unsigned char *bd = sg.body();
    VMI_COPY(bd + of, pk.buf + of, pk.used - of);
    T *sx = reinterpret_cast<T *>(bd + ox);
    int *dflag = reinterpret_cast<int *>(bd + of);
    hipLaunchKernelGGL(k_thomas<T>, dim3(1), dim3(64), 0, st, use_a ? reinterpret_cast<T *>(bd + oa) : nullptr,
                       reinterpret_cast<T *>(bd + ob), use_c ? reinterpret_cast<T *>(bd + oc) : nullptr,
                       reinterpret_cast<T *>(bd + od), n, 1, sx, reinterpret_cast<T *>(bd + ow), dflag);
    VMI_HIP(hipGetLastError());
    VMI_COPY(pk.buf, bd, of + sizeof(int));
    VMI_HIP(hipStreamSynchronize(st));
    std::memcpy(x, pk.buf + ox, nn * sizeof(T));
    std::memcpy(&flag, pk.buf + of, sizeof(int));
    return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
  }
  // staged: body = [a | c | b | d] [x | ws]
  if ((2 * nm + 4 * nn) * sizeof(T) > kStageBody) {
    set_last_error(VIBE_MI_ERR_TOO_LARGE);
    return VIBE_MI_ERR_TOO_LARGE;
  }
  T *sa = reinterpret_cast<T *>(sg.body()), *sc = sa + nm, *sb = sc + nm, *sd = sb + nn, *sx = sd + nn,
    *sw = sx + nn;
  if (use_a) VMI_COPY(sa, a, nm * sizeof(T));
  if (use_c) VMI_COPY(sc, c, nm * sizeof(T));
  VMI_COPY(sb, b, nn * sizeof(T));
  VMI_COPY(sd, d, nn * sizeof(T));
  int *dflag = thomas_flag(sw, n, 1);
  if (int rc = thomas_launch<T>(use_a ? sa : nullptr, sb, use_c ? sc : nullptr, sd, n, 1, sx, sw, dflag, st))
    return rc;
  VMI_COPY(x, sx, nn * sizeof(T));
  VMI_COPY(&flag, dflag, sizeof(int));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
}

template <typename T>
int weno_one(T qm2, T qm1, T q0, T q1, T q2, T *ql, T *qr) {
  Stage sg;
  if (int rc = sg.open()) return rc;
  T *dout = reinterpret_cast<T *>(sg.head());
  hipLaunchKernelGGL(k_weno_scalar<T>, dim3(1), dim3(64), 0, hipStreamPerThread, qm2, qm1, q0, q1, q2, dout);
  VMI_HIP(hipGetLastError());
  T out[2];
  VMI_COPY(out, dout, sizeof(out));
  VMI_HIP(hipStreamSynchronize(hipStreamPerThread));
  *ql = out[0];
  *qr = out[1];
  return 0;
}

// Reduces x[0..n) (GPU-readable, or pageable and staged chunk by chunk) into
// *out_dev (a double in the head) on stream st.
template <typename T>
int nsum_enqueue(const T *x, bool staged, size_t n, Stage &sg, double *out_dev, hipStream_t st) {
  const size_t nchunks = n == 0 ? 1 : (n + kSumChunk - 1) / kSumChunk;
  if (nchunks > kSumMaxChunks) return VIBE_MI_ERR_TOO_LARGE;
  NSum *part = reinterpret_cast<NSum *>(sg.head() + kSumPartOff);
  NSum *slots = reinterpret_cast<NSum *>(sg.head() + kSumSlotOff);
  T *body = reinterpret_cast<T *>(sg.body());
  for (size_t ch = 0; ch < nchunks; ++ch) {
    const size_t beg = ch * kSumChunk, len = std::min(kSumChunk, n - std::min(n, beg));
    const T *src = x + beg;
    if (staged && len > 0) {
      VMI_HIP(hipMemcpyAsync(body, src, len * sizeof(T), hipMemcpyDefault, st));
      src = body;
    }
    const size_t sub = len == 0 ? 1 : (len + kSumBlocks - 1) / kSumBlocks;
    const int blocks = len == 0 ? 1 : static_cast<int>((len + sub - 1) / sub);
    hipLaunchKernelGGL(k_nsum_partial<T>, dim3(blocks), dim3(kSumThreads), 0, st, src, len, sub, part);
    // one chunk: its accumulator is the result (merging one slot again
    // would give the same bits)
    hipLaunchKernelGGL(k_nsum_merge, dim3(1), dim3(kSumThreads), 0, st, part, blocks, slots + ch,
                       nchunks == 1 ? out_dev : static_cast<double *>(nullptr));
  }
  if (nchunks > 1)
    hipLaunchKernelGGL(k_nsum_merge, dim3(1), dim3(kSumThreads), 0, st, slots, static_cast<int>(nchunks), part,
                       out_dev);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
double nsum_one(const T *x, size_t n) {
  set_last_error(0);
  if (x == nullptr) return 0.0;
  Stage sg;
  if (int rc = sg.open()) {
    set_last_error(rc);
    return NAN;
  }
  const T *gx = gpu_view(x, sg.device());
  double *dres = reinterpret_cast<double *>(sg.head());
  if (int rc = nsum_enqueue<T>(gx ? gx : x, gx == nullptr, n, sg, dres, hipStreamPerThread)) {
    set_last_error(rc);
    return NAN;
  }
  double r = NAN;
  if (hipMemcpyAsync(&r, dres, sizeof(double), hipMemcpyDeviceToHost, hipStreamPerThread) != hipSuccess ||
      hipStreamSynchronize(hipStreamPerThread) != hipSuccess) {
    set_last_error(VIBE_MI_ERR_HIP);
    return NAN;
  }
  return r;
}

double bilinear_one(const double *f, int halo, int nsx, int nsy, int nz, double xi, double yj, int k) {
  Stage sg;
  if (int rc = sg.open()) {
    set_last_error(rc);
    return NAN;
  }
  hipStream_t st = hipStreamPerThread;
  double *dout = reinterpret_cast<double *>(sg.head());
  if (const double *gf = gpu_view(f, sg.device())) {
    hipLaunchKernelGGL(k_bilinear, dim3(1), dim3(64), 0, st, gf, halo, nsx, nsy, k, xi, yj, dout);
  } else {
    // pageable field: the four corner values the stencil reads
    const BilinearCorners c = bilinear_weights(halo, nsx, nsy, xi, yj);
    const double *plane = f + static_cast<size_t>(k + halo) * nsx * nsy;
    auto at = [&](int i, int j) { return plane[static_cast<size_t>(j + halo) * nsx + (i + halo)]; };
    hipLaunchKernelGGL(k_bilinear_corners, dim3(1), dim3(64), 0, st, halo, nsx, nsy, xi, yj, at(c.ia, c.ja),
                       at(c.ib, c.ja), at(c.ia, c.jb), at(c.ib, c.jb), dout);
  }
  double r = NAN;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&r, dout, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    set_last_error(VIBE_MI_ERR_HIP);
    return NAN;
  }
  return r;
}

int cfl_one(const double *u, const double *v, const double *w, const double *dzeta, const double *zs, int nx,
            int ny, int nz, double dx, double dy, double z_top, double dt, double out[3]) {
  Stage sg;
  if (int rc = sg.open()) return rc;
  hipStream_t st = hipStreamPerThread;
  const int dev = sg.device();
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * static_cast<size_t>(nz);
  unsigned long long *res = reinterpret_cast<unsigned long long *>(sg.head());
  hipLaunchKernelGGL(k_cfl_init, dim3(1), dim3(64), 0, st, res);
  const double *gu = gpu_view(u, dev), *gv = gpu_view(v, dev), *gw = gpu_view(w, dev),
               *gz = gpu_view(dzeta, dev), *gs = gpu_view(zs, dev);
  if (n > 0 && gu && gv && gw && gz && (zs == nullptr || gs)) {
    const unsigned blocks = std::min<size_t>(ceil_div(n, 256), 2048);
    hipLaunchKernelGGL(k_cfl, dim3(blocks), dim3(256), 0, st, gu, gv, gw, gz, gs, zs ? 1 : 0, nxy, size_t(0), n,
                       dx, dy, z_top, dt, res);
  } else if (n > 0 && (3 * n + nz + (zs ? nxy : 0)) * sizeof(double) + 6 * 256 <= Pack::kMax) {
    Pack pk;
    const size_t ou = pk.put(u, n * sizeof(double)), ov = pk.put(v, n * sizeof(double)),
                 ow = pk.put(w, n * sizeof(double)), oz = pk.put(dzeta, nz * sizeof(double));
    const size_t os = zs ? pk.put(zs, nxy * sizeof(double)) : 0;
    unsigned char *bd = sg.body();
    VMI_COPY(bd, pk.buf, pk.used);
    auto at = [&](size_t o) { return reinterpret_cast<const double *>(bd + o); };
    hipLaunchKernelGGL(k_cfl, dim3(ceil_div(n, 256)), dim3(256), 0, st, at(ou), at(ov), at(ow), at(oz),
                       zs ? at(os) : nullptr, zs ? 1 : 0, nxy, size_t(0), n, dx, dy, z_top, dt, res);
  } else if (n > 0) {
    // staged: body = [dzeta nz] [zs by column, when it fits in a quarter of
    // the body] [chunk of u | v | w (| zs in point order)]
    const size_t al = 256;
    const size_t fz = (static_cast<size_t>(nz) * sizeof(double) + al - 1) / al * al;
    const size_t fcol = (nxy * sizeof(double) + al - 1) / al * al;
    const bool zs_cols = zs != nullptr && fcol <= kStageBody / 4;
    const size_t fixed = fz + (zs_cols ? fcol : 0);
    const int per_point = zs != nullptr && !zs_cols ? 4 : 3;
    if (fixed + per_point * 256 * sizeof(double) > kStageBody) return VIBE_MI_ERR_TOO_LARGE;
    const size_t chunk = std::min(n, (kStageBody - fixed) / (per_point * sizeof(double)));
    double *sz = reinterpret_cast<double *>(sg.body());
    double *scol = reinterpret_cast<double *>(sg.body() + fz);
    double *su = reinterpret_cast<double *>(sg.body() + fixed), *sv = su + chunk, *sw = sv + chunk,
           *ss = sw + chunk;
    VMI_COPY(sz, dzeta, nz * sizeof(double));
    if (zs_cols) VMI_COPY(scol, zs, nxy * sizeof(double));
    for (size_t beg = 0; beg < n; beg += chunk) {
      const size_t len = std::min(chunk, n - beg);
      VMI_COPY(su, u + beg, len * sizeof(double));
      VMI_COPY(sv, v + beg, len * sizeof(double));
      VMI_COPY(sw, w + beg, len * sizeof(double));
      if (zs != nullptr && !zs_cols) {
        // zs of points beg..beg+len-1 in point order: whole column runs
        for (size_t t = 0; t < len;) {
          const size_t col = (beg + t) % nxy, run = std::min(len - t, nxy - col);
          VMI_COPY(ss + t, zs + col, run * sizeof(double));
          t += run;
        }
      }
      const unsigned blocks = std::min<size_t>(ceil_div(len, 256), 2048);
      const double *zarg = zs == nullptr ? nullptr : (zs_cols ? scol : ss);
      const int zmode = zs == nullptr ? 0 : (zs_cols ? 1 : 2);
      hipLaunchKernelGGL(k_cfl, dim3(blocks), dim3(256), 0, st, su, sv, sw, sz, zarg, zmode, nxy, beg, len, dx, dy,
                         z_top, dt, res);
      VMI_HIP(hipGetLastError());
      // the next chunk's copies are ordered after this launch on the stream
    }
  }
  VMI_HIP(hipGetLastError());
  unsigned long long r[3];
  VMI_COPY(r, res, sizeof(r));
  VMI_HIP(hipStreamSynchronize(st));
  for (int q = 0; q < 3; ++q) out[q] = __builtin_bit_cast(double, r[q]);
  return 0;
}

int flux_one(const double *q, const double *rho, const double *u, const double *v, const double *w,
             const double *dzeta, const double *zs, int nx, int ny, int nz, double dx, double dy, double z_top,
             double *dqdt) {
  Stage sg;
  if (int rc = sg.open()) return rc;
  hipStream_t st = hipStreamPerThread;
  const int dev = sg.device();
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * nz;
  const double *gq = gpu_view(q, dev), *gr = gpu_view(rho, dev), *gu = gpu_view(u, dev), *gv = gpu_view(v, dev),
               *gw = gpu_view(w, dev), *gz = gpu_view(dzeta, dev), *gs = gpu_view(zs, dev);
  double *go = gpu_view(dqdt, dev);
  if (gq && gr && gu && gv && gw && gz && (zs == nullptr || gs) && go) {
    hipLaunchKernelGGL(k_flux_adv, dim3(ceil_div(nx, 64), ny, nz), dim3(64), 0, st, gq, gr, gu, gv, gw, gz, gs,
                       nx, ny, nz, dx, dy, z_top, 0, 0, 0, go);
    VMI_HIP(hipGetLastError());
    VMI_HIP(hipStreamSynchronize(st));
    return 0;
  }
  if ((5 * n + nz + (zs ? nxy : 0)) * sizeof(double) + 8 * 256 <= Pack::kMax) {
    Pack pk;
    const size_t oq = pk.put(q, n * sizeof(double)), orr = pk.put(rho, n * sizeof(double)),
                 ou = pk.put(u, n * sizeof(double)), ov = pk.put(v, n * sizeof(double)),
                 ow = pk.put(w, n * sizeof(double)), oz = pk.put(dzeta, nz * sizeof(double));
    const size_t os = zs ? pk.put(zs, nxy * sizeof(double)) : 0;
    const size_t oo = Pack::align(pk.used);
    unsigned char *bd = sg.body();
    VMI_COPY(bd, pk.buf, pk.used);
    auto at = [&](size_t o) { return reinterpret_cast<const double *>(bd + o); };
    hipLaunchKernelGGL(k_flux_adv, dim3(ceil_div(nx, 64), ny, nz), dim3(64), 0, st, at(oq), at(orr), at(ou), at(ov),
                       at(ow), at(oz), zs ? at(os) : nullptr, nx, ny, nz, dx, dy, z_top, 0, 0, 0,
                       reinterpret_cast<double *>(bd + oo));
    VMI_HIP(hipGetLastError());
    VMI_COPY(dqdt, bd + oo, n * sizeof(double));
    VMI_HIP(hipStreamSynchronize(st));
    return 0;
  }
  // staged by level slabs [k0, k1): q, rho, u, v, w over [k0-1, k1+1) clipped
  // to the column, dqdt over [k0, k1); body = [dzeta] [zs] [5 slabs] [out]
  const size_t al = 256;
  const size_t fz = (nz * sizeof(double) + al - 1) / al * al, fs = (nxy * sizeof(double) + al - 1) / al * al;
  const size_t plane = nxy * sizeof(double);
  if (fz + fs + 16 * plane > kStageBody) return VIBE_MI_ERR_TOO_LARGE;  // one level with its neighbours
  const size_t levels = std::min<size_t>(nz, ((kStageBody - fz - fs) / plane - 10) / 6);
  unsigned char *p = sg.body();
  double *sz = reinterpret_cast<double *>(p), *ss = reinterpret_cast<double *>(p + fz);
  double *slab = reinterpret_cast<double *>(p + fz + fs);
  const size_t sl = (levels + 2) * nxy;  // values per staged slab
  double *sq = slab, *sr = sq + sl, *su = sr + sl, *sv = su + sl, *sw = sv + sl, *so = sw + sl;
  VMI_COPY(sz, dzeta, nz * sizeof(double));
  if (zs != nullptr) VMI_COPY(ss, zs, plane);
  for (int k0 = 0; k0 < nz; k0 += static_cast<int>(levels)) {
    const int k1 = std::min(nz, k0 + static_cast<int>(levels));
    const int ka = std::max(0, k0 - 1), kb = std::min(nz, k1 + 1);
    const size_t off = static_cast<size_t>(ka) * nxy, cnt = static_cast<size_t>(kb - ka) * nxy;
    VMI_COPY(sq, q + off, cnt * sizeof(double));
    VMI_COPY(sr, rho + off, cnt * sizeof(double));
    VMI_COPY(su, u + off, cnt * sizeof(double));
    VMI_COPY(sv, v + off, cnt * sizeof(double));
    VMI_COPY(sw, w + off, cnt * sizeof(double));
    hipLaunchKernelGGL(k_flux_adv, dim3(ceil_div(nx, 64), ny, k1 - k0), dim3(64), 0, st, sq, sr, su, sv, sw, sz,
                       zs ? ss : nullptr, nx, ny, nz, dx, dy, z_top, k0, ka, k0, so);
    VMI_HIP(hipGetLastError());
    VMI_COPY(dqdt + static_cast<size_t>(k0) * nxy, so, static_cast<size_t>(k1 - k0) * plane);
  }
  VMI_HIP(hipStreamSynchronize(st));
  return 0;
}

template <typename T>
int thomas_batched(const T *a, const T *b, const T *c, const T *d, int n, int nsys, T *x, T *ws,
                          void *stream) {
  if (n <= 0 || nsys <= 0 || b == nullptr || d == nullptr || x == nullptr || ws == nullptr)
    return VIBE_C_ERR_BAD_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int *dflag = thomas_flag(ws, n, nsys);
  if (int rc = thomas_launch<T>(a, b, c, d, n, nsys, x, ws, dflag, st)) return rc;
  int flag = 0;
  VMI_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
}

}  // namespace vibe_mi

using namespace vibe_mi;

extern "C" {

int vibe_mi_last_error(void) { return last_error(); }

int vibe_tridiagonal_solve_f64(const double *a, const double *b, const double *c, const double *d,
                               int n, double *x, double *ws) {
  return thomas_one<double>(a, b, c, d, n, x, ws);
}
int vibe_tridiagonal_solve_f32(const float *a, const float *b, const float *c, const float *d, int n,
                               float *x, float *ws) {
  return thomas_one<float>(a, b, c, d, n, x, ws);
}

int vibe_mi_tridiagonal_solve_batched_f64(const double *a, const double *b, const double *c,
                                          const double *d, int n, int nsys, double *x, double *ws,
                                          void *stream) {
  return thomas_batched<double>(a, b, c, d, n, nsys, x, ws, stream);
}

int vibe_mi_tridiagonal_solve_batched_f32(const float *a, const float *b, const float *c,
                                          const float *d, int n, int nsys, float *x, float *ws,
                                          void *stream) {
  return thomas_batched<float>(a, b, c, d, n, nsys, x, ws, stream);
}

int vibe_mi_tridiagonal_solve_batched_noalloc_f64(const double *a, const double *b, const double *c,
                                                  const double *d, int n, int nsys, double *x, void *stream) {
  if (n <= 0 || nsys <= 0 || b == nullptr || d == nullptr || x == nullptr) return VIBE_C_ERR_BAD_ARG;
  const size_t need = static_cast<size_t>(n) * static_cast<size_t>(nsys) * sizeof(double);
  if (need > kScratchBytes - kScratchHead) {
    set_last_error(VIBE_MI_ERR_TOO_LARGE);
    return VIBE_MI_ERR_TOO_LARGE;
  }
  void *base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int *dflag = static_cast<int *>(base);
  double *ws = reinterpret_cast<double *>(static_cast<unsigned char *>(base) + kScratchHead);
  if (int rc = thomas_launch<double>(a, b, c, d, n, nsys, x, ws, dflag, st)) return rc;
  int flag = 0;
  VMI_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
}

size_t vibe_mi_noalloc_scratch_capacity(void) { return (kScratchBytes - kScratchHead) / sizeof(double); }

void vibe_weno5_reconstruct_f64(double qm2, double qm1, double q0, double q1, double q2, double *q_left,
                                double *q_right) {
  if (q_left == nullptr || q_right == nullptr) return;
  set_last_error(0);
  if (int rc = weno_one<double>(qm2, qm1, q0, q1, q2, q_left, q_right)) {
    set_last_error(rc);
    *q_left = *q_right = NAN;
  }
}
void vibe_weno5_reconstruct_f32(float qm2, float qm1, float q0, float q1, float q2, float *q_left,
                                float *q_right) {
  if (q_left == nullptr || q_right == nullptr) return;
  set_last_error(0);
  if (int rc = weno_one<float>(qm2, qm1, q0, q1, q2, q_left, q_right)) {
    set_last_error(rc);
    *q_left = *q_right = NAN;
  }
}

int vibe_mi_weno5_reconstruct_batched_f64(const double *q, size_t n, double *q_left, double *q_right,
                                          void *stream) {
  if (q == nullptr || q_left == nullptr || q_right == nullptr) return VIBE_C_ERR_BAD_ARG;
  if (n == 0) return VIBE_C_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_weno<double>, dim3(ceil_div(n, 256)), dim3(256), 0, st, q, n, q_left, q_right);
  VMI_HIP(hipGetLastError());
  VMI_HIP(hipStreamSynchronize(st));
  return VIBE_C_OK;
}

double vibe_bilinear_f64(const double *f, int halo, int nsx, int nsy, int nz, double xi, double yj, int k) {
  set_last_error(0);
  if (f == nullptr || nz <= 0) return 0.0;
  if (k < 0) k = 0;
  if (k >= nz) k = nz - 1;
  return bilinear_one(f, halo, nsx, nsy, nz, xi, yj, k);
}

double vibe_cfl_scan_f64(const double *u, const double *v, const double *w, const double *dzeta,
                         const double *zs, int nx, int ny, int nz, double dx, double dy, double z_top,
                         double dt, double *max_wind_out, double *min_dz_out) {
  set_last_error(0);
  if (u == nullptr || v == nullptr || w == nullptr || dzeta == nullptr) return 0.0;
  double r[3];
  if (int rc = cfl_one(u, v, w, dzeta, zs, nx > 0 ? nx : 0, ny > 0 ? ny : 0, nz > 0 ? nz : 0, dx, dy, z_top, dt,
                       r)) {
    set_last_error(rc);
    return NAN;
  }
  if (max_wind_out != nullptr) *max_wind_out = r[1];
  if (min_dz_out != nullptr) *min_dz_out = (r[2] > 1e29) ? 0.0 : r[2];
  return r[0];
}

double vibe_neumaier_sum_f64(const double *x, size_t n) { return nsum_one<double>(x, n); }
double vibe_neumaier_sum_f32(const float *x, size_t n) { return nsum_one<float>(x, n); }

int vibe_mi_neumaier_sum_device_f64(const double *x, size_t n, double *out, void *stream) {
  if (out == nullptr || (x == nullptr && n > 0)) return VIBE_C_ERR_BAD_ARG;
  Stage sg;
  if (int rc = sg.open()) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double *dres = reinterpret_cast<double *>(sg.head());
  if (int rc = nsum_enqueue<double>(x, false, n, sg, dres, st)) return rc;
  VMI_HIP(hipMemcpyAsync(out, dres, sizeof(double), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return 0;
}

void vibe_advect_scalar_flux_f64(const double *q, const double *rho, const double *u, const double *v,
                                 const double *w, const double *dzeta, const double *zs, int nx, int ny,
                                 int nz, double dx, double dy, double z_top, double *dqdt) {
  set_last_error(0);
  if (q == nullptr || rho == nullptr || u == nullptr || v == nullptr || w == nullptr || dzeta == nullptr ||
      dqdt == nullptr)
    return;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  if (int rc = flux_one(q, rho, u, v, w, dzeta, zs, nx, ny, nz, dx, dy, z_top, dqdt)) set_last_error(rc);
}

}  // extern "C"
