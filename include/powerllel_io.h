/* Field files of PowerLLEL_Channel (checkpoints, instantaneous fields) for
 * the gfx950 library: libpllio.so, built against HDF5 >= 1.10.
 *
 * The format is the reference's (src/mod_dataIO.f90:343-397,
 * src/mod_hdf5.f90:149-215,478-621): one file per field, holding
 *   "nt"  : 32-bit integer dataset of one element (writeAttribute_int
 *           stores the time step as a dataset, not as an HDF5 attribute);
 *   <tag> : FP64 dataset of Fortran shape (nx, ny, nz), i.e. C/HDF5 shape
 *           [nz][ny][nx], x fastest ("u", "v", "w", "p").
 * Checkpoints (outputData, src/mod_dataIO.f90:90-102) store u + u_crf, v, w;
 * a restart (inputData :26-57, then transform2CRF src/mod_calcVel.f90:26-50)
 * reads them and subtracts u_crf.  Files written here are read by the
 * reference's read3d and the other way round.
 *
 * Arrays are host memory, (nz+2) x (ny+2) x (nx+2) with halo = 1 (the
 * interior is read or written; ghosts of a read are left untouched) or
 * nz x ny x nx with halo = 0.  Every entry returns PLL_IO_OK (0) or a
 * negative code; pll_io_last_error() gives the message.
 */
#ifndef POWERLLEL_IO_H
#define POWERLLEL_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum { PLL_IO_OK = 0, PLL_IO_EINVAL = -1, PLL_IO_EFILE = -2, PLL_IO_EFORMAT = -3 };

/* outputField, src/mod_dataIO.f90:372-397 (write3d_col src/mod_hdf5.f90:548-621
 * and writeAttribute_int :190-215 at one rank): create or truncate `path` and
 * write "nt" and the dataset `tag` = interior of var + shift. */
int pll_io_write_field(const char *path, int nt, const char *tag, const double *var, int nx,
                       int ny, int nz, int halo, double shift);

/* inputField, src/mod_dataIO.f90:343-370 (readAttribute_int src/mod_hdf5.f90:149-168
 * and read3d :478-546): read "nt" and the dataset `tag` of shape (nx, ny, nz)
 * into the interior of var, each value minus shift (the transform2CRF of a
 * restart).  PLL_IO_EFORMAT if the dataset is not FP64 of that shape. */
int pll_io_read_field(const char *path, int *nt, const char *tag, double *var, int nx, int ny,
                      int nz, int halo, double shift);

/* Shape (nx, ny, nz) of dataset `tag` and the file's "nt" (either pointer may
 * be NULL), without reading the values. */
int pll_io_field_info(const char *path, const char *tag, int shape[3], int *nt);

const char *pll_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* POWERLLEL_IO_H */
