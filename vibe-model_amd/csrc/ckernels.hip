// ckernels.hip -- the reference's C11 hot-kernel ABI (c_kernels.h) on gfx950.
//
// Every vibe_*_f64 entry point keeps the reference's signature and semantics
// (src/common/c_kernels.c, cited per kernel) and computes on the GPU: host
// arrays are staged through the calling thread's device arena.  The
// vibe_mi_*_batched entry points take device pointers and a stream.
//
// Built with -ffp-contract=off: per-point arithmetic is the same IEEE
// sequence as the C code, so WENO5, the CFL scan, flux advection and the
// bilinear stencil are bit-identical to the CPU; Thomas differs from it only
// in the (documented) corrected recurrence; the Neumaier sum is a fixed-order
// tree and agrees to a few ulp.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "stencils.hpp"

namespace vibe_mi {

namespace {
thread_local int t_last_error = 0;
thread_local Arena t_arena;
}  // namespace

void set_last_error(int code) { t_last_error = code; }
int last_error() { return t_last_error; }

int arena_get(size_t bytes, void **out, hipStream_t *stream) {
  int dev = 0, count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    set_last_error(VIBE_MI_ERR_NO_DEVICE);
    return VIBE_MI_ERR_NO_DEVICE;
  }
  VMI_HIP(hipGetDevice(&dev));
  Arena &a = t_arena;
  if (a.device != dev) {  // first use on this device by this thread
    a.ptr = nullptr;
    a.bytes = 0;
    a.stream = nullptr;
    a.device = dev;
  }
  if (a.stream == nullptr) VMI_HIP(hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking));
  if (bytes > a.bytes) {
    if (a.ptr != nullptr) VMI_HIP(hipFree(a.ptr));
    a.ptr = nullptr;
    const size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes + bytes / 4;
    VMI_HIP(hipMalloc(&a.ptr, want));
    a.bytes = want;
  }
  *out = a.ptr;
  *stream = a.stream;
  return 0;
}

// ---------------------------------------------------------------------------
// Thomas: one system per lane, interleaved k-major layout.  Contract of
// c_kernels.h:38-55; recurrence of helmholtz.cpp:57-86; exact-zero pivot
// check as c_kernels.c:34/:40.
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
    if (k < n - 1) cp = (c != nullptr ? c[o] : T(0)) / denom;
    dp = (d[o] - am * dp) / denom;
    ws[o] = cp;
    x[o] = dp;
  }
  T xn = dp;
  for (int k = n - 2; k >= 0; --k) {
    const size_t o = static_cast<size_t>(k) * ns + s;
    xn = x[o] - ws[o] * xn;
    x[o] = xn;
  }
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

constexpr int kMaxDevices = 64;
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

// ---------------------------------------------------------------------------
// Bilinear (c_kernels.c:108-144) on one staged level of the halo'd field.
// ---------------------------------------------------------------------------
__global__ void k_bilinear(const double *__restrict__ plane, int halo, int nsx, int nsy, double xi,
                           double yj, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nxi = nsx - 2 * halo, nyi = nsy - 2 * halo;
  const int i0 = static_cast<int>(floor(xi)), j0 = static_cast<int>(floor(yj));
  const double wx = xi - static_cast<double>(i0), wy = yj - static_cast<double>(j0);
  int ia = i0, ib = i0 + 1, ja = j0, jb = j0 + 1;
  ia = ia < 0 ? 0 : (ia > nxi - 1 ? nxi - 1 : ia);
  ib = ib < 0 ? 0 : (ib > nxi - 1 ? nxi - 1 : ib);
  ja = ja < 0 ? 0 : (ja > nyi - 1 ? nyi - 1 : ja);
  jb = jb < 0 ? 0 : (jb > nyi - 1 ? nyi - 1 : jb);
  auto at = [&](int i, int j) { return plane[static_cast<size_t>(j + halo) * nsx + (i + halo)]; };
  const double q00 = at(ia, ja), q10 = at(ib, ja), q01 = at(ia, jb), q11 = at(ib, jb);
  const double lo = q00 + wx * (q10 - q00);
  const double hi = q01 + wx * (q11 - q01);
  *out = lo + wy * (hi - lo);
}

// ---------------------------------------------------------------------------
// CFL scan (c_kernels.c:148-187): per-point values identical to the C code;
// maxima/minima of non-negative doubles combined through their IEEE bit
// patterns (monotonic as unsigned integers), so the result is exact.
// ---------------------------------------------------------------------------
__device__ inline void atomic_max_pos(unsigned long long *p, double v) {
  atomicMax(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}
__device__ inline void atomic_min_pos(unsigned long long *p, double v) {
  atomicMin(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}

__global__ void __launch_bounds__(256) k_cfl(const double *__restrict__ u, const double *__restrict__ v,
                                             const double *__restrict__ w, const double *__restrict__ dzeta,
                                             const double *__restrict__ zs, int nx, int ny, int nz, double dx,
                                             double dy, double z_top, double dt,
                                             unsigned long long *__restrict__ res) {
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * nz;
  double worst = 0.0, wind = 0.0, mdz = 1e30;
  for (size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t k = idx / nxy, col = idx - k * nxy;
    const double scale = zs != nullptr ? z_top - zs[col] : z_top;
    double dz = dzeta[k] * scale;
    if (dz < 1.0) dz = 1.0;
    if (dz < mdz) mdz = dz;
    const double uu = u[idx], vv = v[idx], ww = w[idx];
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
// Neumaier sum (c_kernels.c:190-206) as a fixed-order two-level reduction:
// block b sums the contiguous chunk b with per-lane Neumaier accumulators
// over a fixed stride, merges lanes by a fixed TwoSum tree, and a single
// block merges the block partials by the same tree.
// ---------------------------------------------------------------------------
constexpr int kSumThreads = 256;
constexpr int kSumBlocks = 1024;

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

__global__ void __launch_bounds__(kSumThreads) k_nsum_final(const NSum *__restrict__ part, int nparts,
                                                            int ncomp, double *__restrict__ out) {
  // ncomp independent sums stored as part[comp * nparts + p]
  __shared__ NSum sh[kSumThreads];
  for (int comp = 0; comp < ncomp; ++comp) {
    NSum acc{0.0, 0.0};
    for (int p = threadIdx.x; p < nparts; p += kSumThreads) acc.merge(part[comp * nparts + p]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kSumThreads / 2; h > 0; h >>= 1) {
      if (threadIdx.x < static_cast<unsigned>(h)) sh[threadIdx.x].merge(sh[threadIdx.x + h]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[comp] = sh[0].s + sh[0].c;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Flux-form scalar advection (c_kernels.c:209-286).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_flux_adv(const double *__restrict__ q, const double *__restrict__ rho,
                                                  const double *__restrict__ u, const double *__restrict__ v,
                                                  const double *__restrict__ w,
                                                  const double *__restrict__ dzeta,
                                                  const double *__restrict__ zs, int nx, int ny, int nz,
                                                  double dx, double dy, double z_top,
                                                  double *__restrict__ dqdt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= nx) return;
  const size_t nxy = static_cast<size_t>(nx) * ny;
  const double rdx = 1.0 / dx, rdy = 1.0 / dy;
  const size_t col = static_cast<size_t>(j) * nx + i;
  const size_t c = static_cast<size_t>(k) * nxy + col;
  const double scale = zs != nullptr ? z_top - zs[col] : z_top;
  double dz = dzeta[k] * scale;
  if (dz < 1.0) dz = 1.0;
  const size_t base = static_cast<size_t>(k) * nxy;
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
  dqdt[c] = -div / rr;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
template <typename T>
int thomas_device(const T *a, const T *b, const T *c, const T *d, int n, int nsys, T *x, T *ws,
                  hipStream_t st, int *dflag) {
  hipLaunchKernelGGL(k_thomas<T>, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, a, b, c, d, n, nsys, x,
                     ws, dflag);
  VMI_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int thomas_host(const T *a, const T *b, const T *c, const T *d, int n, T *x, T *ws) {
  if (n <= 0 || b == nullptr || d == nullptr || x == nullptr || ws == nullptr) return VIBE_C_ERR_BAD_ARG;
  const size_t nb = static_cast<size_t>(n) * sizeof(T), nm = static_cast<size_t>(n > 1 ? n - 1 : 1) * sizeof(T);
  void *mem = nullptr;
  hipStream_t st;
  const size_t need = 2 * nm + 4 * nb + 256;
  if (int rc = arena_get(need, &mem, &st)) return rc;
  char *p = static_cast<char *>(mem);
  int *dflag = reinterpret_cast<int *>(p);
  T *da = reinterpret_cast<T *>(p + 256), *dc = da + (nm / sizeof(T));
  T *db = dc + (nm / sizeof(T)), *dd = db + n, *dx = dd + n, *dws = dx + n;
  VMI_HIP(hipMemsetAsync(dflag, 0, sizeof(int), st));
  if (a != nullptr && n > 1) VMI_HIP(hipMemcpyAsync(da, a, (n - 1) * sizeof(T), hipMemcpyHostToDevice, st));
  if (c != nullptr && n > 1) VMI_HIP(hipMemcpyAsync(dc, c, (n - 1) * sizeof(T), hipMemcpyHostToDevice, st));
  VMI_HIP(hipMemcpyAsync(db, b, nb, hipMemcpyHostToDevice, st));
  VMI_HIP(hipMemcpyAsync(dd, d, nb, hipMemcpyHostToDevice, st));
  if (int rc = thomas_device<T>(a ? da : nullptr, db, c ? dc : nullptr, dd, n, 1, dx, dws, st, dflag)) return rc;
  int flag = 0;
  VMI_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipMemcpyAsync(x, dx, nb, hipMemcpyDeviceToHost, st));
  VMI_HIP(hipMemcpyAsync(ws, dws, nb, hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
}

template <typename T>
void weno_host(T qm2, T qm1, T q0, T q1, T q2, T *ql, T *qr) {
  if (ql == nullptr || qr == nullptr) return;
  void *mem = nullptr;
  hipStream_t st;
  set_last_error(0);
  auto fail = [&](int rc) {
    set_last_error(rc);
    *ql = *qr = static_cast<T>(NAN);
  };
  if (int rc = arena_get(8 * sizeof(T), &mem, &st)) return fail(rc);
  T *dq = static_cast<T *>(mem), *dl = dq + 5, *dr = dq + 6;
  const T hq[5] = {qm2, qm1, q0, q1, q2};
  if (hipMemcpyAsync(dq, hq, sizeof(hq), hipMemcpyHostToDevice, st) != hipSuccess) return fail(VIBE_MI_ERR_HIP);
  hipLaunchKernelGGL(k_weno<T>, dim3(1), dim3(64), 0, st, dq, size_t(1), dl, dr);
  T out[2];
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, dl, 2 * sizeof(T), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(VIBE_MI_ERR_HIP);
  *ql = out[0];
  *qr = out[1];
}

template <typename T>
int nsum_device(const T *x, size_t n, double *out_host, hipStream_t st, NSum *part, double *dres) {
  const size_t chunk = (n + kSumBlocks - 1) / kSumBlocks;
  const int blocks = n == 0 ? 1 : static_cast<int>((n + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_nsum_partial<T>, dim3(blocks), dim3(kSumThreads), 0, st, x, n, chunk == 0 ? 1 : chunk,
                     part);
  hipLaunchKernelGGL(k_nsum_final, dim3(1), dim3(kSumThreads), 0, st, part, blocks, 1, dres);
  VMI_HIP(hipGetLastError());
  VMI_HIP(hipMemcpyAsync(out_host, dres, sizeof(double), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return 0;
}

template <typename T>
double nsum_host(const T *x, size_t n) {
  set_last_error(0);
  if (x == nullptr) return 0.0;
  void *mem = nullptr;
  hipStream_t st;
  const size_t head = 256 + kSumBlocks * sizeof(NSum);
  if (int rc = arena_get(head + n * sizeof(T), &mem, &st)) {
    set_last_error(rc);
    return NAN;
  }
  char *p = static_cast<char *>(mem);
  double *dres = reinterpret_cast<double *>(p);
  NSum *part = reinterpret_cast<NSum *>(p + 256);
  T *dx = reinterpret_cast<T *>(p + head);
  if (n > 0 && hipMemcpyAsync(dx, x, n * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) {
    set_last_error(VIBE_MI_ERR_HIP);
    return NAN;
  }
  double r = 0.0;
  if (int rc = nsum_device<T>(dx, n, &r, st, part, dres)) {
    set_last_error(rc);
    return NAN;
  }
  return r;
}

}  // namespace vibe_mi

using namespace vibe_mi;

extern "C" {

int vibe_mi_last_error(void) { return last_error(); }

int vibe_tridiagonal_solve_f64(const double *a, const double *b, const double *c, const double *d,
                               int n, double *x, double *ws) {
  return thomas_host<double>(a, b, c, d, n, x, ws);
}
int vibe_tridiagonal_solve_f32(const float *a, const float *b, const float *c, const float *d, int n,
                               float *x, float *ws) {
  return thomas_host<float>(a, b, c, d, n, x, ws);
}

int vibe_mi_tridiagonal_solve_batched_f64(const double *a, const double *b, const double *c,
                                          const double *d, int n, int nsys, double *x, double *ws,
                                          void *stream) {
  if (n <= 0 || nsys <= 0 || b == nullptr || d == nullptr || x == nullptr || ws == nullptr)
    return VIBE_C_ERR_BAD_ARG;
  void *mem;
  hipStream_t ast;
  if (int rc = arena_get(256, &mem, &ast)) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ast;
  int *dflag = static_cast<int *>(mem);
  VMI_HIP(hipMemsetAsync(dflag, 0, sizeof(int), st));
  if (int rc = thomas_device<double>(a, b, c, d, n, nsys, x, ws, st, dflag)) return rc;
  int flag = 0;
  VMI_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
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
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, dflag);
  VMI_HIP(hipGetLastError());
  if (int rc = thomas_device<double>(a, b, c, d, n, nsys, x, ws, st, dflag)) return rc;
  int flag = 0;
  VMI_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
}

size_t vibe_mi_noalloc_scratch_capacity(void) { return (kScratchBytes - kScratchHead) / sizeof(double); }

int vibe_mi_tridiagonal_solve_batched_f32(const float *a, const float *b, const float *c,
                                          const float *d, int n, int nsys, float *x, float *ws,
                                          void *stream) {
  if (n <= 0 || nsys <= 0 || b == nullptr || d == nullptr || x == nullptr || ws == nullptr)
    return VIBE_C_ERR_BAD_ARG;
  void *mem;
  hipStream_t ast;
  if (int rc = arena_get(256, &mem, &ast)) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ast;
  int *dflag = static_cast<int *>(mem);
  VMI_HIP(hipMemsetAsync(dflag, 0, sizeof(int), st));
  if (int rc = thomas_device<float>(a, b, c, d, n, nsys, x, ws, st, dflag)) return rc;
  int flag = 0;
  VMI_HIP(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
  VMI_HIP(hipStreamSynchronize(st));
  return flag ? VIBE_C_ERR_SINGULAR : VIBE_C_OK;
}

void vibe_weno5_reconstruct_f64(double qm2, double qm1, double q0, double q1, double q2, double *q_left,
                                double *q_right) {
  weno_host<double>(qm2, qm1, q0, q1, q2, q_left, q_right);
}
void vibe_weno5_reconstruct_f32(float qm2, float qm1, float q0, float q1, float q2, float *q_left,
                                float *q_right) {
  weno_host<float>(qm2, qm1, q0, q1, q2, q_left, q_right);
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
  const size_t plane = static_cast<size_t>(nsx) * nsy;
  void *mem;
  hipStream_t st;
  if (int rc = arena_get(256 + plane * sizeof(double), &mem, &st)) {
    set_last_error(rc);
    return NAN;
  }
  double *dout = static_cast<double *>(mem);
  double *dplane = reinterpret_cast<double *>(static_cast<char *>(mem) + 256);
  double r = NAN;
  if (hipMemcpyAsync(dplane, f + static_cast<size_t>(k + halo) * plane, plane * sizeof(double),
                     hipMemcpyHostToDevice, st) != hipSuccess) {
    set_last_error(VIBE_MI_ERR_HIP);
    return NAN;
  }
  hipLaunchKernelGGL(k_bilinear, dim3(1), dim3(64), 0, st, dplane, halo, nsx, nsy, xi, yj, dout);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&r, dout, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    set_last_error(VIBE_MI_ERR_HIP);
    return NAN;
  }
  return r;
}

double vibe_cfl_scan_f64(const double *u, const double *v, const double *w, const double *dzeta,
                         const double *zs, int nx, int ny, int nz, double dx, double dy, double z_top,
                         double dt, double *max_wind_out, double *min_dz_out) {
  set_last_error(0);
  if (u == nullptr || v == nullptr || w == nullptr || dzeta == nullptr) return 0.0;
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * static_cast<size_t>(nz > 0 ? nz : 0);
  const size_t fb = n * sizeof(double);
  void *mem;
  hipStream_t st;
  if (int rc = arena_get(256 + 3 * fb + nz * sizeof(double) + nxy * sizeof(double) + 64, &mem, &st)) {
    set_last_error(rc);
    return NAN;
  }
  char *p = static_cast<char *>(mem);
  unsigned long long *res = reinterpret_cast<unsigned long long *>(p);
  double *du = reinterpret_cast<double *>(p + 256), *dv = du + n, *dw = dv + n, *dzt = dw + n, *dzs = dzt + nz;
  const unsigned long long init[3] = {0ull, 0ull,
                                      static_cast<unsigned long long>(__builtin_bit_cast(long long, 1e30))};
  unsigned long long out[3] = {0, 0, 0};
  bool ok = hipMemcpyAsync(res, init, sizeof(init), hipMemcpyHostToDevice, st) == hipSuccess;
  if (n > 0) {
    ok = ok && hipMemcpyAsync(du, u, fb, hipMemcpyHostToDevice, st) == hipSuccess;
    ok = ok && hipMemcpyAsync(dv, v, fb, hipMemcpyHostToDevice, st) == hipSuccess;
    ok = ok && hipMemcpyAsync(dw, w, fb, hipMemcpyHostToDevice, st) == hipSuccess;
    ok = ok && hipMemcpyAsync(dzt, dzeta, nz * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
    if (zs != nullptr) ok = ok && hipMemcpyAsync(dzs, zs, nxy * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
    const unsigned blocks = std::min<size_t>(ceil_div(n, 256), 2048);
    hipLaunchKernelGGL(k_cfl, dim3(blocks), dim3(256), 0, st, du, dv, dw, dzt, zs ? dzs : nullptr, nx, ny, nz, dx,
                       dy, z_top, dt, res);
    ok = ok && hipGetLastError() == hipSuccess;
  }
  ok = ok && hipMemcpyAsync(out, res, sizeof(out), hipMemcpyDeviceToHost, st) == hipSuccess;
  ok = ok && hipStreamSynchronize(st) == hipSuccess;
  if (!ok) {
    set_last_error(VIBE_MI_ERR_HIP);
    return NAN;
  }
  const double worst = __builtin_bit_cast(double, out[0]);
  const double wind = __builtin_bit_cast(double, out[1]);
  const double mdz = __builtin_bit_cast(double, out[2]);
  if (max_wind_out != nullptr) *max_wind_out = wind;
  if (min_dz_out != nullptr) *min_dz_out = (mdz > 1e29) ? 0.0 : mdz;
  return worst;
}

double vibe_neumaier_sum_f64(const double *x, size_t n) { return nsum_host<double>(x, n); }
double vibe_neumaier_sum_f32(const float *x, size_t n) { return nsum_host<float>(x, n); }

int vibe_mi_neumaier_sum_device_f64(const double *x, size_t n, double *out, void *stream) {
  if (out == nullptr || (x == nullptr && n > 0)) return VIBE_C_ERR_BAD_ARG;
  void *mem;
  hipStream_t ast;
  if (int rc = arena_get(256 + kSumBlocks * sizeof(NSum), &mem, &ast)) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ast;
  char *p = static_cast<char *>(mem);
  return nsum_device<double>(x, n, out, st, reinterpret_cast<NSum *>(p + 256), reinterpret_cast<double *>(p));
}

void vibe_advect_scalar_flux_f64(const double *q, const double *rho, const double *u, const double *v,
                                 const double *w, const double *dzeta, const double *zs, int nx, int ny,
                                 int nz, double dx, double dy, double z_top, double *dqdt) {
  set_last_error(0);
  if (q == nullptr || rho == nullptr || u == nullptr || v == nullptr || w == nullptr || dzeta == nullptr ||
      dqdt == nullptr)
    return;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  const size_t nxy = static_cast<size_t>(nx) * ny, n = nxy * nz, fb = n * sizeof(double);
  void *mem;
  hipStream_t st;
  if (int rc = arena_get(6 * fb + nz * sizeof(double) + nxy * sizeof(double) + 64, &mem, &st)) {
    set_last_error(rc);
    return;
  }
  double *dq = static_cast<double *>(mem), *dr = dq + n, *du = dr + n, *dv = du + n, *dw = dv + n, *dout = dw + n;
  double *dzt = dout + n, *dzs = dzt + nz;
  bool ok = true;
  ok = ok && hipMemcpyAsync(dq, q, fb, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(dr, rho, fb, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(du, u, fb, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(dv, v, fb, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(dw, w, fb, hipMemcpyHostToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(dzt, dzeta, nz * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
  if (zs != nullptr) ok = ok && hipMemcpyAsync(dzs, zs, nxy * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
  hipLaunchKernelGGL(k_flux_adv, dim3(ceil_div(nx, 64), ny, nz), dim3(64), 0, st, dq, dr, du, dv, dw, dzt,
                     zs ? dzs : nullptr, nx, ny, nz, dx, dy, z_top, dout);
  ok = ok && hipGetLastError() == hipSuccess;
  ok = ok && hipMemcpyAsync(dqdt, dout, fb, hipMemcpyDeviceToHost, st) == hipSuccess;
  ok = ok && hipStreamSynchronize(st) == hipSuccess;
  if (!ok) set_last_error(VIBE_MI_ERR_HIP);
}

}  // extern "C"
