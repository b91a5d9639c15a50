/*
 * s2d_h5.c -- the reference's HDF5 / Xdmf output files, written by one
 * process through the serial HDF5 C library (libs2d_h5.so). It serves the
 * Fortran driver's IO_m; spectral2d_amd/restart_h5.py writes the same files
 * from Python.
 *
 * Layouts (one process, so each dataset is the whole field):
 *   REST_########.h5  src/IO.F90:306-424: dataset "Header" (null dataspace,
 *                     native int) with scalar attributes nx, ny, nadv (int),
 *                     time, nu (double); group "FlowData" with Omega, Psi, u,
 *                     v as native double, Fortran dimensions (nxG, nyG),
 *                     i.e. C dimensions (nyG, nxG), chunked by the whole field
 *                     (src/IO.F90:439-496).
 *   GRID.h5           src/IO.F90:158-285: "Header" with nx, ny; group
 *                     "Domain_00001" with x(i, j) = i 2 pi / nxG and
 *                     y(i, j) = j 2 pi / nyG, i, j from 1.
 *   REST_########.xmf src/IO.F90:583-650.
 */
#include "s2d_h5.h"

#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int attr(hid_t obj, const char *name, hid_t type, const void *val)
{
    hid_t sp = H5Screate(H5S_SCALAR);
    if (sp < 0) return -1;
    hid_t a = H5Acreate2(obj, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    int rc = a < 0 ? -1 : (H5Awrite(a, type, val) < 0 ? -1 : 0);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return rc;
}

/* "Header": a null-dataspace int dataset that carries the attributes */
static hid_t header(hid_t f)
{
    hid_t sp = H5Screate(H5S_NULL);
    if (sp < 0) return -1;
    hid_t d = H5Dcreate2(f, "Header", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(sp);
    return d;
}

/* one (ny, nx) native-double dataset, chunked by the whole array */
static int dataset(hid_t g, const char *name, int nx, int ny, const double *a)
{
    hsize_t dims[2] = {(hsize_t)ny, (hsize_t)nx};
    hid_t sp = H5Screate_simple(2, dims, NULL);
    if (sp < 0) return -1;
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    int rc = -1;
    if (pl >= 0 && H5Pset_chunk(pl, 2, dims) >= 0) {
        hid_t d = H5Dcreate2(g, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
        if (d >= 0) {
            rc = H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a) < 0 ? -1 : 0;
            H5Dclose(d);
        }
    }
    if (pl >= 0) H5Pclose(pl);
    H5Sclose(sp);
    return rc;
}

/* the first nxG of every padded row (ld doubles apart) of a real field */
static double *compact(const double *f, int nxG, int nyG, int ld)
{
    double *out = (double *)malloc(sizeof(double) * (size_t)nxG * (size_t)nyG);
    if (!out) return NULL;
    for (int j = 0; j < nyG; ++j)
        memcpy(out + (size_t)j * nxG, f + (size_t)j * ld, sizeof(double) * (size_t)nxG);
    return out;
}

int s2d_h5_write_rest(const char *path, int nxG, int nyG, int rISize, int nadv, double time,
                      double nu, const double *omega, const double *psi, const double *u,
                      const double *v)
{
    if (!path || !omega || !psi || !u || !v || nxG < 1 || nyG < 1 || rISize < nxG)
        return S2D_H5_ERR_ARG;
    if (H5open() < 0) return S2D_H5_ERR_HDF5;
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return S2D_H5_ERR_HDF5;
    int rc = 0;
    hid_t h = header(f);
    if (h < 0) rc = -1;
    if (!rc) rc = attr(h, "nx", H5T_NATIVE_INT, &nxG);
    if (!rc) rc = attr(h, "ny", H5T_NATIVE_INT, &nyG);
    if (!rc) rc = attr(h, "nadv", H5T_NATIVE_INT, &nadv);
    if (!rc) rc = attr(h, "time", H5T_NATIVE_DOUBLE, &time);
    if (!rc) rc = attr(h, "nu", H5T_NATIVE_DOUBLE, &nu);
    if (h >= 0) H5Dclose(h);
    hid_t g = rc ? -1 : H5Gcreate2(f, "FlowData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) rc = -1;
    const char *names[4] = {"Omega", "Psi", "u", "v"};
    const double *fields[4] = {omega, psi, u, v};
    for (int k = 0; k < 4 && !rc; ++k) {
        double *c = compact(fields[k], nxG, nyG, rISize);
        if (!c) {
            H5Gclose(g);
            H5Fclose(f);
            return S2D_H5_ERR_ALLOC;
        }
        rc = dataset(g, names[k], nxG, nyG, c);
        free(c);
    }
    if (g >= 0) H5Gclose(g);
    if (H5Fclose(f) < 0) rc = -1;
    return rc ? S2D_H5_ERR_HDF5 : 0;
}

int s2d_h5_write_grid(const char *path, int nxG, int nyG)
{
    if (!path || nxG < 1 || nyG < 1) return S2D_H5_ERR_ARG;
    const size_t n = (size_t)nxG * (size_t)nyG;
    double *x = (double *)malloc(sizeof(double) * n), *y = (double *)malloc(sizeof(double) * n);
    if (!x || !y) {
        free(x);
        free(y);
        return S2D_H5_ERR_ALLOC;
    }
    const double twopi = 2.0 * 3.14159265358979323846264338327950288;
    for (int j = 0; j < nyG; ++j)
        for (int i = 0; i < nxG; ++i) {
            x[(size_t)j * nxG + i] = (double)(i + 1) * twopi / (double)nxG;
            y[(size_t)j * nxG + i] = (double)(j + 1) * twopi / (double)nyG;
        }
    int rc = H5open() < 0 ? -1 : 0;
    hid_t f = rc ? -1 : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) rc = -1;
    hid_t h = rc ? -1 : header(f);
    if (!rc && h < 0) rc = -1;
    if (!rc) rc = attr(h, "nx", H5T_NATIVE_INT, &nxG);
    if (!rc) rc = attr(h, "ny", H5T_NATIVE_INT, &nyG);
    if (h >= 0) H5Dclose(h);
    hid_t g = rc ? -1 : H5Gcreate2(f, "Domain_00001", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (!rc && g < 0) rc = -1;
    if (!rc) rc = dataset(g, "x", nxG, nyG, x);
    if (!rc) rc = dataset(g, "y", nxG, nyG, y);
    if (g >= 0) H5Gclose(g);
    if (f >= 0 && H5Fclose(f) < 0) rc = -1;
    free(x);
    free(y);
    return rc ? S2D_H5_ERR_HDF5 : 0;
}

int s2d_h5_write_xmf(const char *path, int nxG, int nyG, int num)
{
    if (!path || nxG < 1 || nyG < 1 || num < 0) return S2D_H5_ERR_ARG;
    FILE *fh = fopen(path, "w");
    if (!fh) return S2D_H5_ERR_IO;
    char dims[32], item[160], rest[32];
    snprintf(dims, sizeof dims, "%04d %04d", nxG, nyG);
    snprintf(item, sizeof item,
             "            <DataItem Dimensions=\"%s\" NumberType=\"Float\" Precision=\"8\" "
             "Format=\"HDF\">",
             dims);
    snprintf(rest, sizeof rest, "REST_%08d.h5", num);
    fprintf(fh, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
                "<Xdmf Version=\"2.1\">\n   <Domain Name=\"Domain 1\">\n"
                "      <Grid GridType=\"Uniform\" Name=\"Block 1\">\n");
    fprintf(fh, "         <Topology TopologyType=\"2DSMesh\" NumberOfElements=\"%s\"/>\n", dims);
    fprintf(fh, "         <Geometry GeometryType=\"X_Y\">\n");
    const char *axes[2] = {"x", "y"};
    for (int k = 0; k < 2; ++k)
        fprintf(fh, "%s\n               GRID.h5:/Domain_00001/%s\n            </DataItem>\n", item,
                axes[k]);
    fprintf(fh, "         </Geometry>\n");
    const char *names[4] = {"Omega", "Psi", "u", "v"};
    const char *labels[4] = {"Omega [1/s]", "Psi [m^2/s]", "u [m/s]", "v [m/s]"};
    for (int k = 0; k < 4; ++k)
        fprintf(fh,
                "         <Attribute AttributeType=\"Scalar\" Center=\"Node\" Name=\"%s\">\n"
                "%s\n               %s:/FlowData/%s\n            </DataItem>\n"
                "         </Attribute>\n",
                labels[k], item, rest, names[k]);
    fprintf(fh, "      </Grid>\n   </Domain>\n</Xdmf>\n");
    return fclose(fh) == 0 ? 0 : S2D_H5_ERR_IO;
}
