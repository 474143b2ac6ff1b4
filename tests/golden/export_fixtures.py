#!/usr/bin/env python3
"""Convert the reference's HDF5 golden files into compact ``.npz`` fixtures.

This script runs in the development container only: it reads the HDF5 files
under ``/root/reference/tests/runnable/tests`` with ``h5dump -b LE`` (the HDF5
command-line tool; h5py is not installed) and writes ``tests/golden/*.npz``
plus ``tests/golden/manifest.json``.  The committed ``.npz`` files are data
(inputs and expected outputs); nothing of the reference's code is kept.

Size control is lossless where possible:

* a field that is bit-for-bit invariant along an axis (lamellae initial
  conditions stay translation invariant for all time) is stored with that
  axis reduced to length 1; the manifest records the full shape and the test
  broadcasts it back;
* ``fluid-init-eqdistperturb-2d`` was written by 4 ranks that all seeded the
  generator identically, so its four quadrants are identical: one is kept;
* the 64^3 three-dimensional runs that are not invariant keep three central
  planes (x, y, z = n/2) plus the exact (``math.fsum``) total of the field.

Run: ``python3 tests/golden/export_fixtures.py`` (needs /opt/conda/bin/h5dump).
"""
import json
import math
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference/tests/runnable/tests"
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_outarray(path):
    """Return /OutArray of an HDF5 file as a numpy array (little endian)."""
    hdr = subprocess.run([H5DUMP, "-H", "-d", "/OutArray", path], check=True,
                         capture_output=True, text=True).stdout
    m = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", hdr)
    shape = tuple(int(s) for s in m.group(1).split(","))
    if "H5T_IEEE_F64LE" in hdr:
        dt = "<f8"
    elif "H5T_STD_I32LE" in hdr:
        dt = "<i4"
    else:
        raise ValueError("unexpected dtype in %s:\n%s" % (path, hdr))
    with tempfile.TemporaryDirectory() as td:
        binf = os.path.join(td, "a.bin")
        subprocess.run([H5DUMP, "-b", "LE", "-d", "/OutArray", "-o", binf, path],
                       check=True, capture_output=True)
        a = np.fromfile(binf, dtype=dt)
    assert a.size == int(np.prod(shape)), (path, a.size, shape)
    return a.reshape(shape)


def read_timestep_attr(path):
    out = subprocess.run([H5DUMP, "-a", "/globals/dlbc.lb.lb.timestep", path],
                         check=True, capture_output=True, text=True).stdout
    return int(re.search(r"\(0\):\s*(-?\d+)", out).group(1))


def reduce_invariant(a, nspatial):
    """Collapse spatial axes along which ``a`` is bitwise constant."""
    reduced = []
    for ax in range(nspatial):
        first = np.take(a, [0], axis=ax)
        bits = a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)
        fbits = first.view(bits.dtype)
        if np.array_equal(bits, np.broadcast_to(fbits, a.shape)):
            a = first
            reduced.append(ax)
    return a, reduced


def find(dirpath, pattern):
    hits = sorted(f for f in os.listdir(dirpath) if re.fullmatch(pattern, f))
    if len(hits) != 1:
        raise RuntimeError("pattern %s in %s matched %s" % (pattern, dirpath, hits))
    return os.path.join(dirpath, hits[0])


MANIFEST = {}


def save(name, arrays, meta):
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
    MANIFEST[name] = meta
    size = os.path.getsize(os.path.join(OUT, name + ".npz"))
    print("%-40s %8d B  %s" % (name, size, {k: v.shape for k, v in arrays.items()}))


def export_full(name, test, fields, steps, nspatial, reduce=True, tol=None, extra=None):
    """Export whole fields ``fields`` at ``steps``; collapse invariant axes."""
    d = os.path.join(REF, test, "reference-data")
    arrays, shapes, reduced = {}, {}, {}
    for fld in fields:
        for t in steps:
            p = find(d, r"%s-.*t%08d\.h5" % (re.escape(fld), t))
            a = read_outarray(p)
            key = "%s_t%d" % (fld.replace("-", "_"), t)
            shapes[key] = list(a.shape)
            if reduce:
                a, red = reduce_invariant(a, nspatial)
                reduced[key] = red
            arrays[key] = a
    meta = {"source": "tests/runnable/tests/%s/reference-data" % test,
            "shapes": shapes, "reduced_axes": reduced, "tolerance": tol}
    if extra:
        meta.update(extra)
    save(name, arrays, meta)


def export_planes(name, test, fields, steps, tol=None, extra=None):
    """Export central planes and the exact sum of 3-D scalar fields."""
    d = os.path.join(REF, test, "reference-data")
    arrays, meta_f = {}, {}
    for fld in fields:
        for t in steps:
            p = find(d, r"%s-.*t%08d\.h5" % (re.escape(fld), t))
            a = read_outarray(p)
            key = "%s_t%d" % (fld.replace("-", "_"), t)
            n = a.shape
            arrays[key + "_x"] = a[n[0] // 2].copy()
            arrays[key + "_y"] = a[:, n[1] // 2].copy()
            arrays[key + "_z"] = a[:, :, n[2] // 2].copy()
            meta_f[key] = {"shape": list(n), "fsum": math.fsum(a.ravel().tolist()),
                           "planes": [n[0] // 2, n[1] // 2, n[2] // 2]}
    meta = {"source": "tests/runnable/tests/%s/reference-data" % test,
            "fields": meta_f, "tolerance": tol}
    if extra:
        meta.update(extra)
    save(name, arrays, meta)


def export_lb3d(name, test, files, steps, nspatial, tol=None):
    """Export the LB3D (Fortran code) outputs a test keeps beside its own
    reference data (``reference-data-lb3d``): ``files`` maps a field name to
    the LB3D file prefix."""
    d = os.path.join(REF, test, "reference-data-lb3d")
    arrays, shapes, reduced = {}, {}, {}
    for fld, prefix in files.items():
        for t in steps:
            a = read_outarray(find(d, r"%s_t%08d-\d+\.h5" % (re.escape(prefix), t)))
            key = "%s_t%d" % (fld.replace("-", "_"), t)
            shapes[key] = list(a.shape)
            a, red = reduce_invariant(a, nspatial)
            reduced[key] = red
            arrays[key] = a
    save(name, arrays, {"source": "tests/runnable/tests/%s/reference-data-lb3d" % test,
                        "shapes": shapes, "reduced_axes": reduced, "tolerance": tol})


# Every lb/init-2d case (tests/runnable/tests/lb/init-2d/*.json): one fluid or
# mask initialisation kind each, written by 4 ranks (nc = [2, 2]) at t = 0.
# The random kinds seed every rank identically (shiftSeedByRank = false), so
# their four quadrants are identical and one is kept.
INIT_2D_RANDOM = ("fluid-init-constrandom-2d", "fluid-init-eqdistrandom-2d",
                  "fluid-init-eqdistperturbfrac-2d", "fluid-init-eqdistperturb-2d")


def export_init_2d():
    d = os.path.join(REF, "lb/init-2d")
    names = sorted(f[:-5] for f in os.listdir(d) if f.endswith(".json"))
    arrays, meta = {}, {}
    for name in names:
        key = name.replace("-", "_")
        fields = {}
        for fld, prefix in (("population_red", "population-red"),
                            ("population_blue", "population-blue"),
                            ("colour", "colour-red-blue"), ("mask", "mask")):
            fields[fld] = read_outarray(os.path.join(
                d, "reference-data", "%s-%s-t00000000.h5" % (prefix, name)))
        shape = list(fields["mask"].shape)
        quadrant = name in INIT_2D_RANDOM
        for fld, a in fields.items():
            if quadrant:
                h0, h1 = a.shape[0] // 2, a.shape[1] // 2
                q = [a[:h0, :h1], a[h0:, :h1], a[:h0, h1:], a[h0:, h1:]]
                bits = [x.view(np.uint64 if x.dtype.itemsize == 8 else np.uint32) for x in q]
                assert all(np.array_equal(bits[0], b) for b in bits[1:]), (name, fld)
                a = q[0].copy()
            arrays["%s__%s" % (key, fld)] = a
        meta[key] = {"params": "lb/init-2d/%s.in" % name, "shape": shape,
                     "quadrant": quadrant}
    save("init_2d", arrays,
         {"source": "tests/runnable/tests/lb/init-2d/reference-data",
          "layout": "np=4, nc=[2,2]; random kinds (shiftSeedByRank=false) keep one "
                    "quadrant, the share of one rank", "cases": meta,
          "tolerance": {"t0": 0.0}})


def export_checkpoint_lowdim():
    """technical/checkpoint-2d (D2Q9) and checkpoint-1d (D1Q3): restart state at
    t = 0 plus the outputs at t = 0 and t = 10 (strict h5diff)."""
    for name, sub, label in (("checkpoint_2d", "technical/checkpoint-2d", "checkpoint-2d"),
                             ("checkpoint_d1q3", "technical/checkpoint-1d", "checkpoint-d1q3")):
        d = os.path.join(REF, sub, "reference-data")
        arrays = {}
        for f, col in ((0, "red"), (1, "blue")):
            cp = find(os.path.join(d, "cp"), r"cp-fluids%d-%s-.*t00000000\.h5" % (f, re.escape(label)))
            assert read_timestep_attr(cp) == 0
            arrays["cp_fluids%d_t0" % f] = read_outarray(cp)
            for t in (0, 10):
                for fld in ("population", "density"):
                    arrays["%s_%s_t%d" % (fld, col, t)] = read_outarray(
                        find(d, r"%s-%s-%s-.*t%08d\.h5" % (fld, col, re.escape(label), t)))
        for t in (0, 10):
            arrays["colour_t%d" % t] = read_outarray(
                find(d, r"colour-red-blue-%s-.*t%08d\.h5" % (re.escape(label), t)))
        arrays["cp_mask_t0"] = read_outarray(
            find(os.path.join(d, "cp"), r"cp-mask-%s-.*t00000000\.h5" % re.escape(label)))
        src = "%s/%s.in.source" % (sub, label)
        save(name, arrays, {"source": "tests/runnable/tests/%s/reference-data" % sub,
                            "tolerance": {"t0": 0.0, "t10": 0.0}, "params": src})


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present")
    # Strict (bit-exact) LB3D comparison cases, BDist2, no Shan-Chen.
    export_full("bdist2_3d", "lb/bdist2-3d",
                ["population-red", "population-blue", "density-red", "density-blue"],
                [0, 10], 3, reduce=False, tol={"t0": 0.0, "t10": 0.0},
                extra={"params": "lb/bdist2-3d/bdist2-3d.in"})
    export_full("bdist2_2d", "lb/bdist2-2d",
                ["population-red", "population-blue", "density-red", "density-blue"],
                [0, 10], 2, reduce=False, tol={"t0": 0.0, "t10": 0.0},
                extra={"params": "lb/bdist2-2d/bdist2-2d.in"})
    # BDist2 + Shan-Chen, lamellae (translation invariant across the lamellae).
    export_full("bdist2_sc_3d", "lb/bdist2-sc-3d",
                ["population-red", "population-blue", "density-red", "density-blue",
                 "colour-red-blue"],
                [0, 10, 100], 3, tol={"t0": 0.0, "t10": 1e-15, "t100": 1e-15},
                extra={"params": "lb/bdist2-sc-3d/bdist2-sc-3d.in"})
    # The same run by LB3D (compare-lb3d.sh: h5diff -d 1e-15 against the above).
    export_lb3d("bdist2_sc_3d_lb3d", "lb/bdist2-sc-3d",
                {"population-red": "od_pop_dlbc-sc", "density-red": "od_dlbc-sc",
                 "population-blue": "wd_pop_dlbc-sc", "density-blue": "wd_dlbc-sc"},
                [0, 100], 3, tol={"t0": 1e-15, "t100": 1e-15})
    export_full("bdist2_sc_2d", "lb/bdist2-sc-2d",
                ["population-red", "population-blue", "density-red", "density-blue",
                 "colour-red-blue"],
                [0, 10, 100], 2, tol={"t0": 0.0, "t10": 1e-15, "t100": 1e-14},
                extra={"params": "lb/bdist2-sc-2d/bdist2-sc-2d.in"})
    export_full("merging_droplets_2d", "lb/merging-droplets-2d",
                ["population-red", "population-blue", "density-red", "density-blue",
                 "colour-red-blue"],
                [0, 100, 1000], 2, reduce=False,
                tol={"t0": 0.0, "t100": 1e-15, "t1000": 1e-14},
                extra={"params": "lb/merging-droplets-2d/merging-droplets-2d.in"})
    export_planes("merging_droplets_3d", "lb/merging-droplets-3d",
                  ["density-red", "density-blue", "colour-red-blue"], [0, 10, 1000],
                  tol={"t0": 0.0, "t10": 1e-15, "t1000": 1e-14},
                  extra={"params": "lb/merging-droplets-3d/merging-droplets-3d.in"})
    export_planes("laplace_droplet_3d", "lb/laplace-droplet-3d",
                  ["density-red", "density-blue", "colour-red-blue"], [0, 10, 100, 10000],
                  tol={"t0": 0.0, "t10": 0.0, "t100": 1e-15, "t10000": 1e-15},
                  extra={"params": "lb/laplace-droplet-3d/laplace-droplet-3d.in"})
    export_full("laplace_droplet_2d", "lb/laplace-droplet-2d",
                ["density-red", "density-blue", "colour-red-blue"], [0, 100, 10000], 2,
                reduce=False, tol={"t0": 0.0, "t100": 1e-15, "t10000": 1e-15},
                extra={"params": "lb/laplace-droplet-2d/laplace-droplet-2d.in"})
    # SecondOrder (the reference default) with walls and a global acceleration.
    export_full("poiseuille_2d", "lb/poiseuille-2d",
                ["density-red", "density-blue", "velocity-red", "velocity-blue", "mask",
                 "force-red", "force-blue"], [0, 1000, 100000], 2, reduce=False,
                tol={"t0": 0.0, "t1000": 0.0, "t100000": 0.0},
                extra={"params": "lb/poiseuille-2d/poiseuille-2d.in"})
    # ThirdOrder + Shan-Chen, D2Q9 and D1Q3.
    # (the quasi-1d data sit one directory deeper than the other tests)
    for lat, nsp in (("d2q9", 2), ("d1q3", 1)):
        d = os.path.join(REF, "lb/laplace-quasi-1d/reference-data", lat)
        arrays, shapes, reduced = {}, {}, {}
        for fld in ("density-red", "density-blue", "colour-red-blue"):
            for t in (0, 1000):
                a = read_outarray(find(d, r"%s-.*t%08d\.h5" % (re.escape(fld), t)))
                key = "%s_t%d" % (fld.replace("-", "_"), t)
                shapes[key] = list(a.shape)
                a, red = reduce_invariant(a, nsp)
                reduced[key] = red
                arrays[key] = a
        save("laplace_quasi_1d_" + lat, arrays,
             {"source": "tests/runnable/tests/lb/laplace-quasi-1d/reference-data/" + lat,
              "shapes": shapes, "reduced_axes": reduced,
              "tolerance": {"t0": 0.0, "t1000": 1e-15},
              "params": "lb/laplace-quasi-1d/laplace-quasi-1d-%s.in" % lat})
    # Restart state and the state five steps later (strict comparison).
    d = os.path.join(REF, "technical/checkpoint-3d/reference-data")
    arrays = {}
    for f, col in ((0, "red"), (1, "blue")):
        cp = os.path.join(d, "cp", "cp-fluids%d-checkpoint-3d-t00000000.h5" % f)
        arrays["cp_fluids%d_t0" % f] = read_outarray(cp)
        assert read_timestep_attr(cp) == 0
        arrays["population_%s_t5" % col] = read_outarray(
            find(d, r"population-%s-checkpoint-3d-t00000005\.h5" % col))
        arrays["density_%s_t5" % col] = read_outarray(
            find(d, r"density-%s-checkpoint-3d-t00000005\.h5" % col))
    arrays["cp_mask_t0"] = read_outarray(os.path.join(d, "cp", "cp-mask-checkpoint-3d-t00000000.h5"))
    save("checkpoint_3d", arrays,
         {"source": "tests/runnable/tests/technical/checkpoint-3d/reference-data",
          "tolerance": {"t5": 0.0}, "params": "technical/checkpoint-3d/checkpoint-3d.in.source"})
    # Seeded perturbed-equilibrium initialisation (Mt19937, 4 ranks, same seed).
    d = os.path.join(REF, "lb/init-2d/reference-data")
    arrays = {}
    for col in ("red", "blue"):
        a = read_outarray(find(d, r"population-%s-fluid-init-eqdistperturb-2d-t00000000\.h5" % col))
        q = [a[:32, :32], a[32:, :32], a[:32, 32:], a[32:, 32:]]
        assert all(np.array_equal(q[0], x) for x in q[1:]), "quadrants differ"
        arrays["population_%s_t0_quadrant" % col] = q[0].copy()
    save("init_eqdistperturb_2d", arrays,
         {"source": "tests/runnable/tests/lb/init-2d/reference-data",
          "params": "lb/init-2d/fluid-init-eqdistperturb-2d.in",
          "layout": "np=4, nc=[2,2], shiftSeedByRank=false: one 32x32 quadrant kept",
          "tolerance": {"t0": 0.0}})
    export_init_2d()
    export_checkpoint_lowdim()
    # Laplace diagnostics written every 100 steps (lb/laplace.d:55-281), printed
    # with 7 significant digits: t, sigma, measuredR, inPres, outPres, D.
    for dim in ("2d", "3d"):
        d = os.path.join(REF, "lb/laplace-droplet-%s/reference-data" % dim)
        path = find(d, r"laplace-laplace-droplet-%s-.*\.asc" % dim)
        rows = []
        with open(path) as fh:
            for line in fh:
                if line.startswith("#") or not line.strip():
                    continue
                v = line.split()
                rows.append([float(v[2]), float(v[3]), float(v[5]), float(v[6]), float(v[7]),
                             float(v[8])])
        a = np.array(rows)
        _, first = np.unique(a[:, 0], return_index=True)  # the final line repeats t = 10000
        save("laplace_trace_%s" % dim, {"trace": a[np.sort(first)]},
             {"source": "tests/runnable/tests/lb/laplace-droplet-%s/reference-data/%s"
                        % (dim, os.path.basename(path)),
              "columns": ["t", "sigma", "measuredR", "inPres", "outPres", "D"],
              "precision": "7 significant digits (%e)",
              "params": "lb/laplace-droplet-%s/laplace-droplet-%s.in" % (dim, dim)})
    with open(os.path.join(OUT, "manifest.json"), "w") as fh:
        json.dump(MANIFEST, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
