/* vibe_mi/vibebin.h -- .vibebin state files and restarts (C ABI, host only).
 *
 * Replaces the reference's fallback binary I/O:
 *   BinaryWriter / BinaryReader      include/vibe/io/binary_writer.hpp:24-59,
 *                                    src/io/binary_writer.cpp:201-489
 *   read_binary_header               src/io/binary_writer.cpp:495-526
 *   write_restart / read_restart     src/io/field_io.cpp:207-250
 * and reads and writes files byte for byte like them (the "classic" layout
 * that python/vibe_post/src/vibe_post/io.py also reads):
 *
 *   offset  type       field
 *   0       char[8]    magic "VIBEBIN1"
 *   8       int32 x3   nx, ny, nz
 *   20      int32      real_kind (0: written by an FP64 model)
 *   24      float64    time
 *   32      int32      nvars
 *   36      int32      nlevels (= nz)
 *   40      nvars records: char name[32] (NUL padded, at most 31 characters)
 *           + float64 data[ni*nj*nk], i fastest
 *
 * all little-endian.  A record's shape follows from its name alone
 * (binary_writer.cpp:117-138): "u"/"u_facex" (nx+1, ny, nz), "v"/"v_facey"
 * (nx, ny+1, nz), "w"/"zeta"/"w_facez" (nx, ny, nz+1), "__restart_info__"
 * (2, 1, 1), any other name (nx, ny, nz).
 *
 * The functions are host code (no GPU needed); the domain state moves through
 * vibe_mi_domain_get_state / vibe_mi_domain_set_state.
 * Return codes: 0 ok, VIBE_C_ERR_BAD_ARG (2) for invalid arguments, and the
 * codes below.  They replace the reference's vibe::IoError exceptions.
 */
#ifndef VIBE_MI_VIBEBIN_H
#define VIBE_MI_VIBEBIN_H

#include <stddef.h>

#include "vibe_mi/c_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VIBE_MI_ERR_IO 110        /* the file could not be opened, read or written */
#define VIBE_MI_ERR_FORMAT 111    /* not a .vibebin file, bad header or truncated */
#define VIBE_MI_ERR_NOT_FOUND 112 /* the file holds no record of that name */

/* Header of a file (read_binary_header, binary_writer.cpp:495-526).  Every
 * output pointer may be NULL.  Also checks that every record lies inside the
 * file (VIBE_MI_ERR_FORMAT if not). */
int vibe_mi_vibebin_header(const char *path, int *nx, int *ny, int *nz, int *real_kind, double *time,
                           int *nvars, int *nlevels);

/* Name of record `index` (0-based, file order) into name[32]
 * (BinaryReader::variables, binary_writer.cpp:411-416). */
int vibe_mi_vibebin_name(const char *path, int index, char name[32]);

/* Writes a file of nvars records (BinaryWriter::write + close,
 * binary_writer.cpp:224-320).  data[v] holds the record's ni*nj*nk values in
 * the file's order; names longer than 31 characters are truncated; empty
 * names, NULL arrays and non-positive sizes are VIBE_C_ERR_BAD_ARG.  real_kind
 * is written as 0 (the model state is FP64). */
int vibe_mi_vibebin_write(const char *path, int nx, int ny, int nz, double time, int nvars,
                          const char *const *names, const double *const *data);

/* Reads the first record called `name` (BinaryReader::read,
 * binary_writer.cpp:425-481).  count must equal the record's size
 * (VIBE_C_ERR_BAD_ARG otherwise, the reference's dimension check). */
int vibe_mi_vibebin_read(const char *path, const char *name, double *out, size_t count);

/* Restart file of one domain (write_restart, field_io.cpp:207-227): records
 * "__restart_info__" = [time, step], then the reference's 12 species
 * u, v, w, rho, theta, pi, qv, qc, qr, qi, qs, qg on their staggered shapes
 * (state.hpp:53-56, field_io.cpp:44-57) as totals: rho = rho0 + rho', theta
 * and pi likewise, u and v repeated at the periodic face nx (ny), moisture
 * zero (the state is dry); then "rho_p", "theta_p", "pi_p", the perturbations
 * exactly, so that a restart reproduces the state bit for bit.  Arrays in the
 * layout of dycore.h (u, v, rho', theta', pi', rho0, theta0, pi0
 * [nz][ny][nx], w [nz+1][ny][nx]). */
int vibe_mi_restart_write(const char *path, int nx, int ny, int nz, double time, int step,
                          const double *u, const double *v, const double *w, const double *rho_p,
                          const double *theta_p, const double *pi_p, const double *rho0,
                          const double *theta0, const double *pi0);

/* Reads a restart file (read_restart, field_io.cpp:229-250) into the layout
 * of dycore.h.  The header's nx, ny, nz must match (VIBE_C_ERR_BAD_ARG).
 * Perturbations come from "rho_p"/"theta_p"/"pi_p" where present, otherwise
 * from the totals minus rho0/theta0/pi0 (a file the reference wrote); u and v
 * drop the periodic face nx (ny).  time and step come from
 * "__restart_info__", 0 when it is absent (field_io.cpp:233-243); either
 * pointer may be NULL.  A missing prognostic is VIBE_MI_ERR_NOT_FOUND (the
 * reference keeps the target's old values there; a restart here needs all). */
int vibe_mi_restart_read(const char *path, int nx, int ny, int nz, double *time, int *step, double *u,
                         double *v, double *w, double *rho_p, double *theta_p, double *pi_p,
                         const double *rho0, const double *theta0, const double *pi0);

#ifdef __cplusplus
}
#endif

#endif /* VIBE_MI_VIBEBIN_H */
