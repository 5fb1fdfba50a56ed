/*
 * opentbl_io.h -- restart files of the OpenTBL boundary-layer DNS
 * (libopentbl_io.so, host code over the HDF5 C library).
 *
 * One file per field, <base>.u.h5, .v.h5, .w.h5, .p.h5, as the reference
 * names them (escru.F90:44-48 writes chfile.NNN.u, getstartzy-bg.F90:55-58
 * reads chinit.u, both adding ".h5").  Each file holds:
 *   "value"    the field rounded to 4-byte reals (escru.F90:180-236), a
 *              Fortran (nz1, R, nx) array, i.e. C order [nx][R][nz1]: the
 *              planes layout of opentbl.h, real index kr fastest;
 *              R = ny+1 for u, w and ny for v, p;
 *   the header of writeheader (escru.F90:950-997): "Variable" (string),
 *   "tiempo", "cfl", "Re", "lx", "ly", "lz", "timeinit", "dt" (8-byte reals,
 *   one each), "nx", "ny", "nz2", "xout", "procs" (4-byte ints, one each),
 *   "y" (ny+2 reals, y(0:ny+1)) and "um" (ny+1 reals).
 * Every entry point returns 0 or a non-zero status with
 * otbl_io_last_error() giving the message.  Plain host pointers only.
 */
#ifndef OPENTBL_IO_H
#define OPENTBL_IO_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct otbl_io_header {
  double tiempo, cfl, re, lx, ly, lz; /* lx = ax pi, ly = ay pi, lz = 2 pi az */
  int nx, ny, nz2, xout, procs;
  double timeinit, dt;
} otbl_io_header;

/* writeheader + the serial field write of escribezy (escru.F90:159-251,
 * 950-997): creates (truncates) path and writes the FP64 field of
 * nx*R*(2 nz2 + 2) doubles rounded to FP32, the header, y (ny+2) and um
 * (ny+1).  variable is "u", "v", "w" or "p" (any short name is stored). */
int otbl_io_write_field(const char *path, const char *variable, const otbl_io_header *h, const double *y,
                        const double *um, const double *field, int R);

/* readheader (getstartzy-bg.F90:490-581): the header, the first
 * varcap-1 characters of "Variable" (NUL-terminated; variable may be null),
 * y (ny+2 doubles, may be null) and um (ny+1 doubles, may be null), with
 * ny the file's (read the header with null y and um first to size them). */
int otbl_io_read_header(const char *path, otbl_io_header *h, char *variable, size_t varcap, double *y,
                        double *um);

/* The field read of getstartzy (getstartzy-bg.F90:180-275): the file's
 * "value" array widened to FP64 into field (nx*R*nz1 doubles for the
 * caller's nz1): the first min(nz1, file nz1) reals of every row are the
 * file's, the rest are zero (a restart read at another spanwise
 * resolution).  The file's nx and rows per plane must be nx and R (the
 * reference stops on a different ny and cannot read another nx on one
 * rank). */
int otbl_io_read_field(const char *path, double *field, int nx, int R, int nz1);

/* The enstrophy file of escribezy (escru.F90:159-170): "value" the
 * enstrophy field ens(nz, ny, nx) of the last statistics sample (4-byte
 * reals, C order [nx][ny][nz]) with the writeheader header, Variable
 * "enstrophy". */
int otbl_io_write_ens(const char *path, const otbl_io_header *h, const double *y, const double *um, const float *ens,
                      int nz);

/* The statistics files of escrst (escru.F90:263-341, 461-516), Fortran
 * sequential unformatted records, big-endian:
 *   <st_path>  (the reference's .st): record 1 tiempo, cfl, Re, ax, ay, az
 *              (8-byte reals), nx, ny, nz2, ical (4-byte integers); record
 *              2 y(0:ny+1); record 3 one 8-byte 0; then one record per
 *              plane i: us(1:ny+1), vs(1:ny), ws(1:ny+1), uv(1:ny),
 *              ua(1:ny+1), va(1:ny), wa(1:ny+1), pp, pm, vortx, vorty,
 *              vortz, vortxa, vortya, vortza, uw, vw (1:ny each);
 *   <budget_path> (.budget): records 1-2 as above, then per plane dispu,
 *              dispv, dispw, dispuv, pup, pvp, pdudx, pdvdy, pdudy, pdvdx,
 *              pdwdz, v3, u2v, v2u, w2v, dudx0, dudy0, dvdx0, dvdy0, dwdx0,
 *              dwdy0 (1:ny each), u3(1:ny+1), w2u(1:ny+1).
 * stats holds the 40 accumulators of otbl_stats_get in the library's order
 * (include/opentbl.h), each (ny+1, nx).  Either path may be null. */
typedef struct otbl_io_stats_header {
  double tiempo, cfl, re, ax, ay, az; /* ax, ay, az as bl_1 passes them (lengths / pi, / 2 pi) */
  int nx, ny, nz2, ical;
} otbl_io_stats_header;
int otbl_io_write_stats(const char *st_path, const char *budget_path, const otbl_io_stats_header *h,
                        const double *y, const double *stats);

const char *otbl_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
