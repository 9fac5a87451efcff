/* Field files in the reference's HDF5 layout (declarations and format in
 * include/powerllel_io.h).  Host-side only: the caller copies a field off the
 * device first.  The values go plane by plane through one nx*ny buffer, which
 * also strips the halo and applies the u_crf shift, so a field of any size
 * needs only one plane of extra host memory.  (The reference moves a field
 * in one transfer, or in batches of planes above 2 GiB,
 * src/mod_hdf5.f90:494-500; the file contents do not depend on either.)
 */
#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/powerllel_io.h"

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *pll_io_last_error(void) { return g_err; }

static int check_args(const char *path, const char *tag, const void *var, int nx, int ny, int nz,
                      int halo)
{
    if (!path || !tag || !*tag || !var) return fail(PLL_IO_EINVAL, "NULL path, tag or array");
    if (nx < 1 || ny < 1 || nz < 1) return fail(PLL_IO_EINVAL, "bad shape %d x %d x %d", nx, ny, nz);
    if (halo != 0 && halo != 1) return fail(PLL_IO_EINVAL, "halo must be 0 or 1");
    return PLL_IO_OK;
}

/* element (i, j, k), 0-based interior indices, of a host array */
static size_t at(int i, int j, int k, int nx, int ny, int halo)
{
    const size_t px = (size_t)nx + 2 * halo, py = (size_t)ny + 2 * halo;
    return ((size_t)(k + halo) * py + (size_t)(j + halo)) * px + (size_t)(i + halo);
}

int pll_io_write_field(const char *path, int nt, const char *tag, const double *var, int nx,
                       int ny, int nz, int halo, double shift)
{
    int rc = check_args(path, tag, var, nx, ny, nz, halo);
    if (rc) return rc;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    double *plane = malloc(sizeof(double) * (size_t)nx * ny);
    if (!plane) return fail(PLL_IO_EINVAL, "out of host memory for one plane");
    hid_t f = -1, s1 = -1, d1 = -1, fs = -1, ms = -1, d3 = -1;
    rc = PLL_IO_OK;
    if ((f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)) < 0) {
        rc = fail(PLL_IO_EFILE, "cannot create %s", path);
        goto done;
    }
    /* writeAttribute_int: a one-element integer dataset named "nt" */
    const hsize_t one = 1;
    if ((s1 = H5Screate_simple(1, &one, NULL)) < 0 ||
        (d1 = H5Dcreate2(f, "nt", H5T_NATIVE_INT, s1, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)) < 0 ||
        H5Dwrite(d1, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &nt) < 0) {
        rc = fail(PLL_IO_EFILE, "cannot write \"nt\" to %s", path);
        goto done;
    }
    /* write3d_col: Fortran (nx, ny, nz) = HDF5 [nz][ny][nx] */
    const hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    const hsize_t pdims[3] = {1, (hsize_t)ny, (hsize_t)nx};
    if ((fs = H5Screate_simple(3, dims, NULL)) < 0 || (ms = H5Screate_simple(3, pdims, NULL)) < 0 ||
        (d3 = H5Dcreate2(f, tag, H5T_NATIVE_DOUBLE, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)) < 0) {
        rc = fail(PLL_IO_EFILE, "cannot create dataset \"%s\" in %s", tag, path);
        goto done;
    }
    for (int k = 0; k < nz; k++) {
        for (int j = 0; j < ny; j++) {
            const double *src = var + at(0, j, k, nx, ny, halo);
            double *dst = plane + (size_t)j * nx;
            if (shift != 0.0)
                for (int i = 0; i < nx; i++) dst[i] = src[i] + shift;
            else
                memcpy(dst, src, sizeof(double) * (size_t)nx);
        }
        const hsize_t start[3] = {(hsize_t)k, 0, 0};
        if (H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, pdims, NULL) < 0 ||
            H5Dwrite(d3, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, plane) < 0) {
            rc = fail(PLL_IO_EFILE, "cannot write plane %d of \"%s\" to %s", k + 1, tag, path);
            goto done;
        }
    }
done:
    if (d3 >= 0) H5Dclose(d3);
    if (ms >= 0) H5Sclose(ms);
    if (fs >= 0) H5Sclose(fs);
    if (d1 >= 0) H5Dclose(d1);
    if (s1 >= 0) H5Sclose(s1);
    if (f >= 0 && H5Fclose(f) < 0 && rc == PLL_IO_OK) rc = fail(PLL_IO_EFILE, "cannot close %s", path);
    free(plane);
    return rc;
}

/* opens `path` and dataset `tag`, checks FP64 and rank 3, reports the shape */
static int open_field(const char *path, const char *tag, hid_t *f, hid_t *d, hid_t *fs,
                      hsize_t dims[3])
{
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    *f = *d = *fs = -1;
    if ((*f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)) < 0)
        return fail(PLL_IO_EFILE, "cannot open %s", path);
    if ((*d = H5Dopen2(*f, tag, H5P_DEFAULT)) < 0)
        return fail(PLL_IO_EFORMAT, "%s has no dataset \"%s\"", path, tag);
    hid_t t = H5Dget_type(*d);
    const int fp64 = t >= 0 && H5Tget_class(t) == H5T_FLOAT && H5Tget_size(t) == 8;
    if (t >= 0) H5Tclose(t);
    if (!fp64) return fail(PLL_IO_EFORMAT, "dataset \"%s\" of %s is not FP64", tag, path);
    if ((*fs = H5Dget_space(*d)) < 0 || H5Sget_simple_extent_ndims(*fs) != 3 ||
        H5Sget_simple_extent_dims(*fs, dims, NULL) != 3)
        return fail(PLL_IO_EFORMAT, "dataset \"%s\" of %s is not three-dimensional", tag, path);
    return PLL_IO_OK;
}

static int read_nt(hid_t f, const char *path, int *nt)
{
    hid_t d = H5Dopen2(f, "nt", H5P_DEFAULT);
    if (d < 0) return fail(PLL_IO_EFORMAT, "%s has no dataset \"nt\"", path);
    hid_t s = H5Dget_space(d);
    const int ok = s >= 0 && H5Sget_simple_extent_npoints(s) == 1 &&
                   H5Dread(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, nt) >= 0;
    if (s >= 0) H5Sclose(s);
    H5Dclose(d);
    return ok ? PLL_IO_OK : fail(PLL_IO_EFORMAT, "cannot read \"nt\" of %s", path);
}

static void close_field(hid_t f, hid_t d, hid_t fs)
{
    if (fs >= 0) H5Sclose(fs);
    if (d >= 0) H5Dclose(d);
    if (f >= 0) H5Fclose(f);
}

int pll_io_field_info(const char *path, const char *tag, int shape[3], int *nt)
{
    if (!path || !tag) return fail(PLL_IO_EINVAL, "NULL path or tag");
    hid_t f, d, fs;
    hsize_t dims[3];
    int rc = open_field(path, tag, &f, &d, &fs, dims);
    if (rc == PLL_IO_OK && shape) {
        shape[0] = (int)dims[2];
        shape[1] = (int)dims[1];
        shape[2] = (int)dims[0];
    }
    if (rc == PLL_IO_OK && nt) rc = read_nt(f, path, nt);
    close_field(f, d, fs);
    return rc;
}

int pll_io_read_field(const char *path, int *nt, const char *tag, double *var, int nx, int ny,
                      int nz, int halo, double shift)
{
    int rc = check_args(path, tag, var, nx, ny, nz, halo);
    if (rc) return rc;
    hid_t f, d, fs, ms = -1;
    hsize_t dims[3];
    double *plane = NULL;
    if ((rc = open_field(path, tag, &f, &d, &fs, dims))) goto done;
    if (dims[0] != (hsize_t)nz || dims[1] != (hsize_t)ny || dims[2] != (hsize_t)nx) {
        rc = fail(PLL_IO_EFORMAT, "dataset \"%s\" of %s is (%llu, %llu, %llu), expected (%d, %d, %d)",
                  tag, path, (unsigned long long)dims[2], (unsigned long long)dims[1],
                  (unsigned long long)dims[0], nx, ny, nz);
        goto done;
    }
    if (nt && (rc = read_nt(f, path, nt))) goto done;
    const hsize_t pdims[3] = {1, (hsize_t)ny, (hsize_t)nx};
    if (!(plane = malloc(sizeof(double) * (size_t)nx * ny))) {
        rc = fail(PLL_IO_EINVAL, "out of host memory for one plane");
        goto done;
    }
    if ((ms = H5Screate_simple(3, pdims, NULL)) < 0) {
        rc = fail(PLL_IO_EFILE, "cannot create a memory space");
        goto done;
    }
    for (int k = 0; k < nz; k++) {
        const hsize_t start[3] = {(hsize_t)k, 0, 0};
        if (H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, pdims, NULL) < 0 ||
            H5Dread(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, plane) < 0) {
            rc = fail(PLL_IO_EFILE, "cannot read plane %d of \"%s\" from %s", k + 1, tag, path);
            goto done;
        }
        for (int j = 0; j < ny; j++) {
            const double *src = plane + (size_t)j * nx;
            double *dst = var + at(0, j, k, nx, ny, halo);
            if (shift != 0.0)
                for (int i = 0; i < nx; i++) dst[i] = src[i] - shift;
            else
                memcpy(dst, src, sizeof(double) * (size_t)nx);
        }
    }
done:
    if (ms >= 0) H5Sclose(ms);
    close_field(f, d, fs);
    free(plane);
    return rc;
}
