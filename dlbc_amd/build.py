"""Build libdlbc_hip.so (gfx950) in-tree with hipcc.

FP64 kernels must reproduce the reference's IEEE double arithmetic bit for
bit, so every translation unit is compiled with -ffp-contract=off (no fused
multiply-add unless written explicitly with fma()).  Packed FP32 instructions
are disabled for all device code (-packed-fp32-ops).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdlbc_hip.so")
ARCH = os.environ.get("DLBC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["dlbc_hip.hip", "host_init.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"] + DEVICE
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        err = "\n".join(l for l in r.stderr.splitlines()
                        if "packed-fp32-ops" not in l and l.strip())
        if r.returncode != 0:
            raise RuntimeError("compile of %s failed:\n%s" % (src, r.stderr))
        if err and verbose:
            print(err, file=sys.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr)
    os.replace(tmp, LIB)
    return LIB


H5_LIB = os.path.join(HERE, "libdlbc_h5.so")
H5_PREFIXES = [os.environ.get("DLBC_HDF5_PREFIX"), "/opt/conda", "/usr"]


def hdf5_prefix():
    for p in H5_PREFIXES:
        if p and os.path.exists(os.path.join(p, "include", "hdf5.h")):
            for lib in ("lib", "lib64", "lib/x86_64-linux-gnu"):
                if os.path.exists(os.path.join(p, lib, "libhdf5.so")):
                    return p, os.path.join(p, lib)
    return None, None


def build_h5(verbose=False):
    """libdlbc_h5.so (the reference's HDF5 file layout) with the host C++
    compiler against the HDF5 C library; None when no HDF5 is installed."""
    prefix, libdir = hdf5_prefix()
    if prefix is None:
        if verbose:
            print("HDF5 not found: libdlbc_h5.so not built", file=sys.stderr)
        return None
    tmp = H5_LIB + ".tmp"
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           os.path.join(CSRC, "h5io.cpp"), "-I", os.path.join(prefix, "include"),
           "-L", libdir, "-Wl,-rpath," + libdir, "-lhdf5", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build of libdlbc_h5.so failed:\n%s" % r.stderr)
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_h5(verbose=True))
