// vibebin.cpp -- .vibebin files and dry-state restarts (host code, C ABI in
// include/vibe_mi/vibebin.h).
//
// The byte layout, the name -> shape rule and the reader's checks follow the
// reference's BinaryWriter / BinaryReader (src/io/binary_writer.cpp:117-138,
// 224-320, 333-481) so that files move between the two programs and
// python/vibe_post in both directions.  Data is written in bulk from
// per-level staging buffers (one fwrite per level) instead of value by
// value; the bytes are the same on the little-endian hosts this library
// targets (checked at compile time).
#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <sys/types.h>

#include "vibe_mi/vibebin.h"

static_assert(std::endian::native == std::endian::little, ".vibebin payloads are written in host order");

namespace {

constexpr char kMagic[8] = {'V', 'I', 'B', 'E', 'B', 'I', 'N', '1'};
constexpr int kNameBytes = 32;
constexpr long long kHeaderBytes = 40;
constexpr const char *kRestartInfo = "__restart_info__";

struct Shape {
  long long ni = 0, nj = 0, nk = 0;
  long long count() const { return ni * nj * nk; }
};

// record shape from its name (binary_writer.cpp:117-138)
Shape shape_of(const std::string &name, int nx, int ny, int nz) {
  if (name == kRestartInfo) return {2, 1, 1};
  if (name == "u" || name == "u_facex") return {nx + 1LL, ny, nz};
  if (name == "v" || name == "v_facey") return {nx, ny + 1LL, nz};
  if (name == "w" || name == "zeta" || name == "w_facez") return {nx, ny, nz + 1LL};
  return {nx, ny, nz};
}

struct File {
  std::FILE *fp = nullptr;
  File(const char *path, const char *mode) : fp(std::fopen(path, mode)) {}
  ~File() {
    if (fp != nullptr) std::fclose(fp);
  }
  File(const File &) = delete;
  File &operator=(const File &) = delete;
  // closes and reports whether buffered writes reached the file
  bool close() {
    const bool ok = std::fclose(fp) == 0;
    fp = nullptr;
    return ok;
  }
};

void put_i32(unsigned char *p, int32_t v) {
  const uint32_t u = static_cast<uint32_t>(v);
  for (int i = 0; i < 4; ++i) p[i] = static_cast<unsigned char>((u >> (8 * i)) & 0xFFu);
}
void put_f64(unsigned char *p, double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  for (int i = 0; i < 8; ++i) p[i] = static_cast<unsigned char>((u >> (8 * i)) & 0xFFu);
}
int32_t get_i32(const unsigned char *p) {
  uint32_t u = 0;
  for (int i = 0; i < 4; ++i) u |= static_cast<uint32_t>(p[i]) << (8 * i);
  return static_cast<int32_t>(u);
}
double get_f64(const unsigned char *p) {
  uint64_t u = 0;
  for (int i = 0; i < 8; ++i) u |= static_cast<uint64_t>(p[i]) << (8 * i);
  double v;
  std::memcpy(&v, &u, 8);
  return v;
}

// ------------------------------------------------------------------ writing

struct Writer {
  File f;
  int nx, ny, nz;
  bool ok = true;
  Writer(const char *path, int nx_, int ny_, int nz_) : f(path, "wb"), nx(nx_), ny(ny_), nz(nz_) {}

  void bytes(const void *p, size_t n) {
    if (ok && n > 0 && std::fwrite(p, 1, n, f.fp) != n) ok = false;
  }
  // header (binary_writer.cpp:286-293)
  void header(double time, int nvars) {
    unsigned char h[kHeaderBytes];
    std::memcpy(h, kMagic, 8);
    put_i32(h + 8, nx);
    put_i32(h + 12, ny);
    put_i32(h + 16, nz);
    put_i32(h + 20, 0);  // real_kind: FP64 writer
    put_f64(h + 24, time);
    put_i32(h + 32, nvars);
    put_i32(h + 36, nz);  // nlevels
    bytes(h, sizeof(h));
  }
  // 32-byte NUL-padded name, at most 31 characters (binary_writer.cpp:296-301)
  void name(const char *nm) {
    char b[kNameBytes] = {};
    const size_t n = std::min<size_t>(std::strlen(nm), kNameBytes - 1);
    std::memcpy(b, nm, n);
    bytes(b, sizeof(b));
  }
  void values(const double *p, size_t n) { bytes(p, n * sizeof(double)); }
};

// ------------------------------------------------------------------ reading

struct Record {
  std::string name;
  long long offset = 0;  // first data byte
  Shape shape;
};

struct Index {
  int nx = 0, ny = 0, nz = 0, real_kind = 0, nvars = 0, nlevels = 0;
  double time = 0.0;
  std::vector<Record> recs;
  const Record *find(const std::string &nm) const {
    for (const Record &r : recs)
      if (r.name == nm) return &r;  // first match, like BinaryReader::read
    return nullptr;
  }
};

// Parses the header and the record table (BinaryReader::open,
// binary_writer.cpp:333-402); unlike the reference, a record that runs past
// the end of the file is reported here rather than when it is read.
int scan(std::FILE *fp, Index &ix) {
  unsigned char h[kHeaderBytes];
  if (std::fread(h, 1, sizeof(h), fp) != sizeof(h) || std::memcmp(h, kMagic, 8) != 0) return VIBE_MI_ERR_FORMAT;
  ix.nx = get_i32(h + 8);
  ix.ny = get_i32(h + 12);
  ix.nz = get_i32(h + 16);
  ix.real_kind = get_i32(h + 20);
  ix.time = get_f64(h + 24);
  ix.nvars = get_i32(h + 32);
  ix.nlevels = get_i32(h + 36);
  if (ix.nx <= 0 || ix.ny <= 0 || ix.nz <= 0 || ix.nvars < 0 || ix.nlevels < 0) return VIBE_MI_ERR_FORMAT;
  if (ix.real_kind != 0 && ix.real_kind != 1) return VIBE_MI_ERR_FORMAT;
  if (fseeko(fp, 0, SEEK_END) != 0) return VIBE_MI_ERR_IO;
  const long long size = ftello(fp);
  if (size < 0) return VIBE_MI_ERR_IO;
  long long off = kHeaderBytes;
  ix.recs.clear();
  for (int v = 0; v < ix.nvars; ++v) {
    char nm[kNameBytes];
    if (off + kNameBytes > size || fseeko(fp, static_cast<off_t>(off), SEEK_SET) != 0 ||
        std::fread(nm, 1, kNameBytes, fp) != static_cast<size_t>(kNameBytes))
      return VIBE_MI_ERR_FORMAT;
    nm[kNameBytes - 1] = '\0';
    Record r;
    r.name = nm;
    r.offset = off + kNameBytes;
    r.shape = shape_of(r.name, ix.nx, ix.ny, ix.nz);
    off = r.offset + r.shape.count() * 8;
    if (off > size) return VIBE_MI_ERR_FORMAT;
    ix.recs.push_back(std::move(r));
  }
  return VIBE_C_OK;
}

int open_index(const char *path, File &f, Index &ix) {
  if (f.fp == nullptr) return VIBE_MI_ERR_IO;
  return scan(f.fp, ix);
}

int read_values(std::FILE *fp, const Record &r, long long first, double *out, size_t n) {
  if (fseeko(fp, static_cast<off_t>(r.offset + first * 8), SEEK_SET) != 0) return VIBE_MI_ERR_IO;
  if (std::fread(out, sizeof(double), n, fp) != n) return VIBE_MI_ERR_FORMAT;
  return VIBE_C_OK;
}

}  // namespace

extern "C" {

int vibe_mi_vibebin_header(const char *path, int *nx, int *ny, int *nz, int *real_kind, double *time,
                           int *nvars, int *nlevels) {
  if (path == nullptr) return VIBE_C_ERR_BAD_ARG;
  File f(path, "rb");
  Index ix;
  if (int rc = open_index(path, f, ix)) return rc;
  if (nx) *nx = ix.nx;
  if (ny) *ny = ix.ny;
  if (nz) *nz = ix.nz;
  if (real_kind) *real_kind = ix.real_kind;
  if (time) *time = ix.time;
  if (nvars) *nvars = ix.nvars;
  if (nlevels) *nlevels = ix.nlevels;
  return VIBE_C_OK;
}

int vibe_mi_vibebin_name(const char *path, int index, char name[32]) {
  if (path == nullptr || name == nullptr || index < 0) return VIBE_C_ERR_BAD_ARG;
  File f(path, "rb");
  Index ix;
  if (int rc = open_index(path, f, ix)) return rc;
  if (index >= static_cast<int>(ix.recs.size())) return VIBE_MI_ERR_NOT_FOUND;
  std::memset(name, 0, kNameBytes);
  std::memcpy(name, ix.recs[index].name.c_str(), ix.recs[index].name.size());
  return VIBE_C_OK;
}

int vibe_mi_vibebin_write(const char *path, int nx, int ny, int nz, double time, int nvars,
                          const char *const *names, const double *const *data) {
  if (path == nullptr || nx <= 0 || ny <= 0 || nz <= 0 || nvars < 0) return VIBE_C_ERR_BAD_ARG;
  if (nvars > 0 && (names == nullptr || data == nullptr)) return VIBE_C_ERR_BAD_ARG;
  for (int v = 0; v < nvars; ++v)
    if (names[v] == nullptr || names[v][0] == '\0' || data[v] == nullptr) return VIBE_C_ERR_BAD_ARG;
  Writer w(path, nx, ny, nz);
  if (w.f.fp == nullptr) return VIBE_MI_ERR_IO;
  w.header(time, nvars);
  for (int v = 0; v < nvars; ++v) {
    // the shape follows from the name as the file stores it (truncated)
    const std::string stored(names[v], std::min<size_t>(std::strlen(names[v]), kNameBytes - 1));
    w.name(names[v]);
    w.values(data[v], static_cast<size_t>(shape_of(stored, nx, ny, nz).count()));
  }
  if (!w.ok) return VIBE_MI_ERR_IO;
  return w.f.close() ? VIBE_C_OK : VIBE_MI_ERR_IO;
}

int vibe_mi_vibebin_read(const char *path, const char *name, double *out, size_t count) {
  if (path == nullptr || name == nullptr || out == nullptr) return VIBE_C_ERR_BAD_ARG;
  File f(path, "rb");
  Index ix;
  if (int rc = open_index(path, f, ix)) return rc;
  const Record *r = ix.find(name);
  if (r == nullptr) return VIBE_MI_ERR_NOT_FOUND;
  if (static_cast<long long>(count) != r->shape.count()) return VIBE_C_ERR_BAD_ARG;
  return read_values(f.fp, *r, 0, out, count);
}

int vibe_mi_restart_write(const char *path, int nx, int ny, int nz, double time, int step,
                          const double *u, const double *v, const double *w, const double *rho_p,
                          const double *theta_p, const double *pi_p, const double *rho0,
                          const double *theta0, const double *pi0) {
  if (path == nullptr || nx <= 0 || ny <= 0 || nz <= 0) return VIBE_C_ERR_BAD_ARG;
  for (const double *p : {u, v, w, rho_p, theta_p, pi_p, rho0, theta0, pi0})
    if (p == nullptr) return VIBE_C_ERR_BAD_ARG;
  const size_t sx = static_cast<size_t>(nx), sy = static_cast<size_t>(ny);
  const size_t plane = sx * sy, ncell = plane * static_cast<size_t>(nz);
  static const char *const kMoist[] = {"qv", "qc", "qr", "qi", "qs", "qg"};
  Writer wr(path, nx, ny, nz);
  if (wr.f.fp == nullptr) return VIBE_MI_ERR_IO;
  wr.header(time, 1 + 12 + 3);
  // [time, step] (field_io.cpp:211-223)
  const double info[2] = {time, static_cast<double>(step)};
  wr.name(kRestartInfo);
  wr.values(info, 2);
  std::vector<double> lev((sx + 1) * (sy + 1));
  // u: (nx+1, ny, nz), the face at i = nx repeats i = 0 (periodic)
  wr.name("u");
  for (int k = 0; k < nz && wr.ok; ++k) {
    const double *src = u + k * plane;
    for (size_t j = 0; j < sy; ++j) {
      std::memcpy(&lev[j * (sx + 1)], src + j * sx, sx * sizeof(double));
      lev[j * (sx + 1) + sx] = src[j * sx];
    }
    wr.values(lev.data(), (sx + 1) * sy);
  }
  // v: (nx, ny+1, nz), the face at j = ny repeats j = 0
  wr.name("v");
  for (int k = 0; k < nz && wr.ok; ++k) {
    wr.values(v + k * plane, plane);
    wr.values(v + k * plane, sx);
  }
  wr.name("w");
  wr.values(w, plane * static_cast<size_t>(nz + 1));
  // totals of the cell-centred thermodynamic fields
  const double *pert[3] = {rho_p, theta_p, pi_p}, *base[3] = {rho0, theta0, pi0};
  static const char *const kTot[3] = {"rho", "theta", "pi"};
  for (int q = 0; q < 3; ++q) {
    wr.name(kTot[q]);
    for (int k = 0; k < nz && wr.ok; ++k) {
      const double *a = pert[q] + k * plane, *b = base[q] + k * plane;
      for (size_t n = 0; n < plane; ++n) lev[n] = b[n] + a[n];
      wr.values(lev.data(), plane);
    }
  }
  std::fill(lev.begin(), lev.end(), 0.0);
  for (const char *m : kMoist) {
    wr.name(m);
    for (int k = 0; k < nz && wr.ok; ++k) wr.values(lev.data(), plane);
  }
  // the perturbations themselves (exact restart)
  static const char *const kPert[3] = {"rho_p", "theta_p", "pi_p"};
  for (int q = 0; q < 3; ++q) {
    wr.name(kPert[q]);
    wr.values(pert[q], ncell);
  }
  if (!wr.ok) return VIBE_MI_ERR_IO;
  return wr.f.close() ? VIBE_C_OK : VIBE_MI_ERR_IO;
}

int vibe_mi_restart_read(const char *path, int nx, int ny, int nz, double *time, int *step, double *u,
                         double *v, double *w, double *rho_p, double *theta_p, double *pi_p,
                         const double *rho0, const double *theta0, const double *pi0) {
  if (path == nullptr || nx <= 0 || ny <= 0 || nz <= 0) return VIBE_C_ERR_BAD_ARG;
  for (const double *p : {static_cast<const double *>(u), static_cast<const double *>(v),
                          static_cast<const double *>(w), static_cast<const double *>(rho_p),
                          static_cast<const double *>(theta_p), static_cast<const double *>(pi_p), rho0,
                          theta0, pi0})
    if (p == nullptr) return VIBE_C_ERR_BAD_ARG;
  File f(path, "rb");
  Index ix;
  if (int rc = open_index(path, f, ix)) return rc;
  if (ix.nx != nx || ix.ny != ny || ix.nz != nz) return VIBE_C_ERR_BAD_ARG;
  const size_t sx = static_cast<size_t>(nx), sy = static_cast<size_t>(ny);
  const size_t plane = sx * sy, ncell = plane * static_cast<size_t>(nz);
  const Record *ru = ix.find("u"), *rv = ix.find("v"), *rw = ix.find("w");
  if (ru == nullptr || rv == nullptr || rw == nullptr) return VIBE_MI_ERR_NOT_FOUND;
  // u drops the face at i = nx: read (nx+1)-long rows, keep the first nx
  std::vector<double> lev((sx + 1) * sy);
  for (int k = 0; k < nz; ++k) {
    if (int rc = read_values(f.fp, *ru, static_cast<long long>(k) * (sx + 1) * sy, lev.data(), (sx + 1) * sy))
      return rc;
    for (size_t j = 0; j < sy; ++j) std::memcpy(u + k * plane + j * sx, &lev[j * (sx + 1)], sx * sizeof(double));
  }
  // v drops the face at j = ny: the first ny rows of each level
  for (int k = 0; k < nz; ++k)
    if (int rc = read_values(f.fp, *rv, static_cast<long long>(k) * sx * (sy + 1), v + k * plane, plane)) return rc;
  if (int rc = read_values(f.fp, *rw, 0, w, plane * static_cast<size_t>(nz + 1))) return rc;
  double *pert[3] = {rho_p, theta_p, pi_p};
  const double *base[3] = {rho0, theta0, pi0};
  static const char *const kTot[3] = {"rho", "theta", "pi"};
  static const char *const kPert[3] = {"rho_p", "theta_p", "pi_p"};
  for (int q = 0; q < 3; ++q) {
    if (const Record *rp = ix.find(kPert[q])) {
      if (int rc = read_values(f.fp, *rp, 0, pert[q], ncell)) return rc;
    } else if (const Record *rt = ix.find(kTot[q])) {
      if (int rc = read_values(f.fp, *rt, 0, pert[q], ncell)) return rc;
      for (size_t n = 0; n < ncell; ++n) pert[q][n] -= base[q][n];
    } else {
      return VIBE_MI_ERR_NOT_FOUND;
    }
  }
  double t = 0.0;
  int s = 0;
  if (const Record *ri = ix.find(kRestartInfo)) {
    double info[2];
    if (int rc = read_values(f.fp, *ri, 0, info, 2)) return rc;
    t = info[0];
    s = static_cast<int>(std::llround(info[1]));
  }
  if (time) *time = t;
  if (step) *step = s;
  return VIBE_C_OK;
}

}  // extern "C"
