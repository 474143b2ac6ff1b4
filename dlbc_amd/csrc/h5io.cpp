// h5io.cpp -- the reference's HDF5 field and checkpoint files.
//
// Layout (src/dlbc/io/hdf5.d:33-39,125-285,395-443,453-545 and
// src/dlbc/io/checkpoint.d:243-310):
//   dataset "/OutArray"      float64 (or int32 for the mask), dimensions =
//                            the global lattice lengths, plus the vector
//                            length for vector fields (populations: q),
//                            row-major, no halo;
//   root attribute "input"   the parameter file, one variable-length string
//                            per line;
//   group "metadata"         revisionHash / revisionDesc / revisionBranch /
//                            revisionChanges (strings), revisionChanged (int32);
//   group "globals"          checkpoints only: "dlbc.lb.lb.timestep" (int32).
//
// Serial HDF5 (one rank); the reference's collective MPI-IO writes the same
// file for a single rank.  No device code: this library links the HDF5 C
// library and nothing of HIP.
#include <hdf5.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dlbc_h5.h"

namespace {

struct Silence {  // no HDF5 error stack on stderr: errors are return codes
  Silence() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

hid_t mem_type(int dtype) { return dtype == DLBC_H5_I32 ? H5T_NATIVE_INT32 : H5T_NATIVE_DOUBLE; }
hid_t file_type(int dtype) { return dtype == DLBC_H5_I32 ? H5T_STD_I32LE : H5T_IEEE_F64LE; }

int write_string_attr(hid_t loc, const char *name, const char *const *lines, hsize_t n) {
  hid_t type = H5Tcopy(H5T_C_S1);
  H5Tset_size(type, H5T_VARIABLE);
  hid_t sid = H5Screate_simple(1, &n, nullptr);
  hid_t aid = H5Acreate2(loc, name, type, sid, H5P_DEFAULT, H5P_DEFAULT);
  const herr_t e = aid < 0 ? -1 : H5Awrite(aid, type, lines);
  if (aid >= 0) H5Aclose(aid);
  H5Sclose(sid);
  H5Tclose(type);
  return e < 0 ? DLBC_H5_ERR_IO : DLBC_H5_OK;
}

int write_int_attr(hid_t loc, const char *name, int32_t v) {
  hsize_t one = 1;
  hid_t sid = H5Screate_simple(1, &one, nullptr);
  hid_t aid = H5Acreate2(loc, name, H5T_STD_I32LE, sid, H5P_DEFAULT, H5P_DEFAULT);
  const herr_t e = aid < 0 ? -1 : H5Awrite(aid, H5T_NATIVE_INT32, &v);
  if (aid >= 0) H5Aclose(aid);
  H5Sclose(sid);
  return e < 0 ? DLBC_H5_ERR_IO : DLBC_H5_OK;
}

}  // namespace

extern "C" {

const char *dlbc_h5_version(void) {
  static std::string v;
  if (v.empty()) {
    unsigned maj = 0, min = 0, rel = 0;
    H5get_libversion(&maj, &min, &rel);
    v = "HDF5 " + std::to_string(maj) + "." + std::to_string(min) + "." + std::to_string(rel);
  }
  return v.c_str();
}

int dlbc_h5_write(const char *path, int dtype, const void *data, int rank, const int64_t *dims,
                  const char *const *input_lines, int ninput, const char *revision,
                  int checkpoint, int timestep) {
  static Silence quiet;
  if (!path || !data || !dims || rank < 1 || rank > 8 || ninput < 0 ||
      (dtype != DLBC_H5_F64 && dtype != DLBC_H5_I32))
    return DLBC_H5_ERR_ARG;
  hsize_t d[8];
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 1) return DLBC_H5_ERR_ARG;
    d[i] = (hsize_t)dims[i];
  }
  hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return DLBC_H5_ERR_IO;
  int rc = DLBC_H5_OK;
  hid_t space = H5Screate_simple(rank, d, nullptr);
  hid_t dset = H5Dcreate2(file, "/OutArray", file_type(dtype), space, H5P_DEFAULT, H5P_DEFAULT,
                          H5P_DEFAULT);
  if (dset < 0 || H5Dwrite(dset, mem_type(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    rc = DLBC_H5_ERR_IO;
  if (dset >= 0) H5Dclose(dset);
  H5Sclose(space);
  hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
  if (!rc && ninput > 0) rc = write_string_attr(root, "input", input_lines, (hsize_t)ninput);
  if (!rc) {  // dumpMetadata (hdf5.d:290-298)
    hid_t g = H5Gcreate2(root, "metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const char *rev = revision ? revision : "";
    const char *empty = "";
    rc = write_string_attr(g, "revisionHash", &rev, 1);
    if (!rc) rc = write_string_attr(g, "revisionDesc", &rev, 1);
    if (!rc) rc = write_string_attr(g, "revisionBranch", &empty, 1);
    if (!rc) rc = write_int_attr(g, "revisionChanged", 0);
    if (!rc) rc = write_string_attr(g, "revisionChanges", &empty, 1);
    H5Gclose(g);
  }
  if (!rc && checkpoint) {  // dumpCheckpointGlobals (checkpoint.d:260-310)
    hid_t g = H5Gcreate2(root, "globals", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc = write_int_attr(g, "dlbc.lb.lb.timestep", timestep);
    H5Gclose(g);
  }
  H5Gclose(root);
  if (H5Fclose(file) < 0 && !rc) rc = DLBC_H5_ERR_IO;
  return rc;
}

int dlbc_h5_shape(const char *path, int *dtype, int *rank, int64_t *dims) {
  static Silence quiet;
  if (!path || !dtype || !rank || !dims) return DLBC_H5_ERR_ARG;
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return DLBC_H5_ERR_IO;
  int rc = DLBC_H5_OK;
  hid_t dset = H5Dopen2(file, "/OutArray", H5P_DEFAULT);
  if (dset < 0) {
    rc = DLBC_H5_ERR_FORMAT;
  } else {
    hid_t space = H5Dget_space(dset);
    const int r = H5Sget_simple_extent_ndims(space);
    hsize_t d[8];
    if (r < 1 || r > 8) {
      rc = DLBC_H5_ERR_FORMAT;
    } else {
      H5Sget_simple_extent_dims(space, d, nullptr);
      *rank = r;
      for (int i = 0; i < r; ++i) dims[i] = (int64_t)d[i];
      hid_t t = H5Dget_type(dset);
      if (H5Tequal(t, H5T_IEEE_F64LE) > 0)
        *dtype = DLBC_H5_F64;
      else if (H5Tequal(t, H5T_STD_I32LE) > 0)
        *dtype = DLBC_H5_I32;
      else
        rc = DLBC_H5_ERR_FORMAT;
      H5Tclose(t);
    }
    H5Sclose(space);
    H5Dclose(dset);
  }
  H5Fclose(file);
  return rc;
}

int dlbc_h5_read(const char *path, int dtype, void *data, int rank, const int64_t *dims,
                 int *timestep) {
  static Silence quiet;
  if (!path || !data || !dims || rank < 1 || rank > 8) return DLBC_H5_ERR_ARG;
  int ftype = -1, frank = 0;
  int64_t fdims[8];
  int rc = dlbc_h5_shape(path, &ftype, &frank, fdims);
  if (rc) return rc;
  if (ftype != dtype || frank != rank) return DLBC_H5_ERR_FORMAT;
  for (int i = 0; i < rank; ++i)
    if (fdims[i] != dims[i]) return DLBC_H5_ERR_FORMAT;
  hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return DLBC_H5_ERR_IO;
  hid_t dset = H5Dopen2(file, "/OutArray", H5P_DEFAULT);
  if (dset < 0 || H5Dread(dset, mem_type(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    rc = DLBC_H5_ERR_IO;
  if (dset >= 0) H5Dclose(dset);
  if (!rc && timestep) {  // readCheckpointGlobals: -1 when the file has no globals
    *timestep = -1;
    if (H5Lexists(file, "globals", H5P_DEFAULT) > 0) {
      hid_t g = H5Gopen2(file, "globals", H5P_DEFAULT);
      if (H5Aexists(g, "dlbc.lb.lb.timestep") > 0) {
        hid_t a = H5Aopen(g, "dlbc.lb.lb.timestep", H5P_DEFAULT);
        int32_t v = -1;
        if (H5Aread(a, H5T_NATIVE_INT32, &v) >= 0) *timestep = v;
        H5Aclose(a);
      }
      H5Gclose(g);
    }
  }
  H5Fclose(file);
  return rc;
}

}  // extern "C"
