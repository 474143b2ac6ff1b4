"""The reference's HDF5 output and checkpoint files for device state.

ctypes front end of libdlbc_h5.so (include/dlbc_h5.h).  File names follow
the reference (src/dlbc/io/io.d makeFilenameOutput / makeFilenameCpOutput):

    <field>-<simulation>-t<8-digit step>.h5     output (density-red, population-blue, ...)
    cp-<field>-<simulation>-t<8-digit step>.h5  checkpoint (cp-fluids0, cp-mask)

Fields come off the device with halo 0, which is the /OutArray layout.
"""
import ctypes as C
import os

import numpy as np

from .lattice import DlbcError

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libdlbc_h5.so")
_lib = None
F64, I32 = 0, 1
REVISION = "dlbc_amd"


def lib():
    """Load libdlbc_h5.so (built by dlbc_amd/build.py where HDF5 is installed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise DlbcError("libdlbc_h5.so is not built (HDF5 not found by dlbc_amd/build.py)")
    L = C.CDLL(_LIB)
    i64p = C.POINTER(C.c_int64)
    L.dlbc_h5_write.restype = C.c_int
    L.dlbc_h5_write.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int, i64p,
                                C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_int, C.c_int]
    L.dlbc_h5_shape.restype = C.c_int
    L.dlbc_h5_shape.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p]
    L.dlbc_h5_read.restype = C.c_int
    L.dlbc_h5_read.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int, i64p,
                               C.POINTER(C.c_int)]
    L.dlbc_h5_version.restype = C.c_char_p
    _lib = L
    return L


def available():
    return os.path.exists(_LIB)


def _check(rc, what, path):
    if rc:
        raise DlbcError("%s %s: %s" % (what, path, {-1: "invalid argument", -2: "HDF5 error",
                                                     -3: "unexpected format"}.get(rc, rc)))


def write_field(path, array, timestep=None, input_lines=(), checkpoint=False):
    """dumpFieldHDF5 (src/dlbc/io/hdf5.d:125-285): /OutArray plus the input,
    metadata and (for checkpoints) globals attributes."""
    a = np.ascontiguousarray(array)
    dtype = I32 if a.dtype == np.int32 else F64
    if dtype == F64:
        a = np.ascontiguousarray(a, np.float64)
    dims = (C.c_int64 * a.ndim)(*a.shape)
    lines = [s.encode() for s in input_lines]
    arr = (C.c_char_p * max(1, len(lines)))(*lines) if lines else None
    rc = lib().dlbc_h5_write(os.fsencode(path), dtype, a.ctypes.data, a.ndim, dims, arr,
                             len(lines), REVISION.encode(), int(bool(checkpoint)),
                             int(timestep or 0))
    _check(rc, "write", path)


def read_field(path):
    """readFieldHDF5 (src/dlbc/io/hdf5.d:395-443): (array, timestep), the
    timestep from globals/dlbc.lb.lb.timestep, or None without it."""
    dtype, rank = C.c_int(), C.c_int()
    dims = (C.c_int64 * 8)()
    _check(lib().dlbc_h5_shape(os.fsencode(path), C.byref(dtype), C.byref(rank), dims),
           "shape of", path)
    shape = tuple(dims[k] for k in range(rank.value))
    out = np.empty(shape, np.int32 if dtype.value == I32 else np.float64)
    t = C.c_int()
    _check(lib().dlbc_h5_read(os.fsencode(path), dtype.value, out.ctypes.data, rank.value, dims,
                              C.byref(t)), "read", path)
    return out, (None if t.value < 0 else t.value)


def output_name(path, field, simulation, t):
    return os.path.join(path, "%s-%s-t%08d.h5" % (field, simulation, t))


def checkpoint_name(path, field, simulation, t):
    return os.path.join(path, "cp-%s-%s-t%08d.h5" % (field, simulation, t))


def dump(L, path, simulation, t, names=("red", "blue"), fields=("density", "population"),
         input_lines=()):
    """The output files of dumpLBData (src/dlbc/lb/io.d:56-101) for the
    fields asked: density-<name>, population-<name>, velocity-<name> per
    component, colour-<red>-<blue>.  Returns the file names."""
    out = []
    for f in range(L.nc):
        for fld in fields:
            if fld == "density":
                a = L.density(f)
            elif fld == "population":
                a = L.population(f)
            elif fld == "velocity":
                a = L.velocity(f)
            else:
                continue
            fn = output_name(path, "%s-%s" % (fld, names[f]), simulation, t)
            write_field(fn, a, input_lines=input_lines)
            out.append(fn)
    if "colour" in fields and L.nc == 2:
        fn = output_name(path, "colour-%s-%s" % (names[0], names[1]), simulation, t)
        write_field(fn, L.colour(), input_lines=input_lines)
        out.append(fn)
    return out


def checkpoint(L, path, simulation, t, mask, input_lines=()):
    """dumpCheckpoint (src/dlbc/io/checkpoint.d:80-157): cp-fluids<f> per
    component and cp-mask, each with the time step in globals.  mask: the
    physical int32 mask the lattice was given (the device keeps a byte copy)."""
    out = []
    for f in range(L.nc):
        fn = checkpoint_name(path, "fluids%d" % f, simulation, t)
        write_field(fn, L.population(f), t, input_lines, checkpoint=True)
        out.append(fn)
    fn = checkpoint_name(path, "mask", simulation, t)
    write_field(fn, np.ascontiguousarray(mask, np.int32), t, input_lines, checkpoint=True)
    out.append(fn)
    return out


def restore(L, path, simulation, t):
    """readCheckpoint (src/dlbc/io/checkpoint.d:248-313): load cp-mask and
    cp-fluids<f> into the lattice; returns the stored time step."""
    mask, tm = read_field(checkpoint_name(path, "mask", simulation, t))
    L.upload_mask(mask)
    step = tm
    for f in range(L.nc):
        pop, tf = read_field(checkpoint_name(path, "fluids%d" % f, simulation, t))
        L.upload_population(f, pop)
        step = tf if step is None else step
    return step
