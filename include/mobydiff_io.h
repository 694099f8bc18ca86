/*
 * mobydiff_io.h -- C ABI of the field-file library (libmobydiff_io.so): checkpoint files of the
 * flow state in HDF5, written and read on the host.
 *
 * The file layout is the reference's (field_hdf5.c:149-204): datasets "un", "vn", "wn", "pn" of
 * shape [nx+2][ny+2][nz+2] (ghosted, k fastest: element (i, j, k) of the Fortran array q(0:nx+1,
 * 0:ny+1,0:nz+1) at ((i*(ny+2) + j)*(nz+2) + k)), FP64, and twelve scalar attributes on the root
 * group: nx, ny, nz (int), lx, ly, lz, dx, dy, dz, re, dt, t_current (double).
 *
 * Arrays cross the boundary as host pointers in the Fortran layout q[i + (nx+2)*(j + (ny+2)*k)],
 * as in libmobydiff_hip; the driver downloads the device fields first (mdx_download_fields) and
 * uploads them after a read (mdx_upload_fields), the `target update` pair of io.f90:124-127,161-164.
 * Functions return 0 on success and 1 on failure, as the reference's do; mdx_io_last_error()
 * then names the failing step.
 */
#ifndef MOBYDIFF_IO_H
#define MOBYDIFF_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error text of the last failing call on this thread */
const char *mdx_io_last_error(void);

/* fdm_h5_write_field (field_hdf5.c:149-184, called by write_field, io.f90:105-137): creates or
 * truncates `path` and writes the twelve attributes and the four ghosted fields */
int mdx_h5_write_field(const char *path, int nx, int ny, int nz, double lx, double ly, double lz,
                       double dx, double dy, double dz, double re, double dt, double t_current,
                       const double *un, const double *vn, const double *wn, const double *pn);

/* fdm_h5_read_field (field_hdf5.c:186-204, called by read_field, io.f90:139-166): reads the four
 * ghosted fields; fails when a dataset is missing or its shape is not [nx+2][ny+2][nz+2] */
int mdx_h5_read_field(const char *path, int nx, int ny, int nz, double *un, double *vn, double *wn,
                      double *pn);

/* The attributes a restart needs (no counterpart in the reference, whose reader ignores them):
 * grid size, time step and simulated time of the file.  Any output pointer may be NULL. */
int mdx_h5_read_field_info(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, double *dt,
                           double *t_current);

#ifdef __cplusplus
}
#endif
#endif
