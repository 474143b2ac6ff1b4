/*
 * dlbc_h5.h -- the reference's HDF5 field and checkpoint files, for the
 * output and restart of device state (libdlbc_h5.so, links the HDF5 C
 * library; no GPU code).
 *
 * Replaces, for one rank:
 *   dumpFieldHDF5          src/dlbc/io/hdf5.d:125-285 (field and checkpoint files)
 *   readFieldHDF5          src/dlbc/io/hdf5.d:395-443
 *   dumpCheckpointGlobals  src/dlbc/io/checkpoint.d:260-310 (globals/dlbc.lb.lb.timestep)
 *   readCheckpointGlobals  src/dlbc/io/checkpoint.d:272-310
 * File layout: dataset /OutArray (float64, or int32 for the mask) with the
 * global lattice lengths (+ the vector length) as dimensions, row-major, no
 * halo; root attribute "input" (parameter file lines); group "metadata";
 * group "globals" for checkpoints.  The device fields reach the host through
 * dlbc_gpu_download_populations / dlbc_gpu_density / ... with halo 0, which
 * is exactly the /OutArray layout.
 */
#ifndef DLBC_H5_H
#define DLBC_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DLBC_H5_F64 = 0, DLBC_H5_I32 = 1 };
enum {
  DLBC_H5_OK = 0,
  DLBC_H5_ERR_ARG = -1,    /* invalid argument */
  DLBC_H5_ERR_IO = -2,     /* the HDF5 library reported an error */
  DLBC_H5_ERR_FORMAT = -3  /* no /OutArray, or another type or shape than asked */
};

/* Write data (row-major, dims[0..rank)) as /OutArray with the root
 * attribute "input" (ninput lines; none if ninput == 0), the metadata group
 * (revision strings set to `revision`) and, if checkpoint != 0, the globals
 * group holding the time step. */
int dlbc_h5_write(const char *path, int dtype, const void *data, int rank, const int64_t *dims,
                  const char *const *input_lines, int ninput, const char *revision,
                  int checkpoint, int timestep);
/* Type, rank and dimensions of /OutArray (dims: room for 8). */
int dlbc_h5_shape(const char *path, int *dtype, int *rank, int64_t *dims);
/* Read /OutArray into data after checking its type and shape; *timestep
 * (may be NULL) receives globals/dlbc.lb.lb.timestep, or -1 without it. */
int dlbc_h5_read(const char *path, int dtype, void *data, int rank, const int64_t *dims,
                 int *timestep);
/* Version string of the HDF5 library in use. */
const char *dlbc_h5_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DLBC_H5_H */
