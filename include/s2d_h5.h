/*
 * s2d_h5.h -- C ABI of libs2d_h5, the serial HDF5 / Xdmf writers of the
 * reference's output files (one process). Separate from libs2d so that the
 * time-step library does not depend on HDF5. Every function returns 0 or
 * one of the S2D_H5_ERR_* codes.
 */
#ifndef S2D_H5_H
#define S2D_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define S2D_H5_ERR_ARG 1    /* null pointer or bad size */
#define S2D_H5_ERR_HDF5 2   /* an HDF5 call failed */
#define S2D_H5_ERR_ALLOC 3  /* host allocation failed */
#define S2D_H5_ERR_IO 4     /* the text file could not be written */

/* WriteRestartHDF5, src/IO.F90:306-424 (one process): REST file at path
 * from four physical fields in the reference real layout (rISize, nyG),
 * Fortran order, of which the first nxG of each row are written. */
int s2d_h5_write_rest(const char *path, int nxG, int nyG, int rISize, int nadv, double time,
                      double nu, const double *omega, const double *psi, const double *u,
                      const double *v);
/* WriteGridHDF5, src/IO.F90:158-285 (one process). */
int s2d_h5_write_grid(const char *path, int nxG, int nyG);
/* WriteXMF, src/IO.F90:583-650: the Xdmf index of REST_<num>.h5. */
int s2d_h5_write_xmf(const char *path, int nxG, int nyG, int num);

#ifdef __cplusplus
}
#endif
#endif
