"""Reference runnable tests restated as parameter dictionaries.

Each entry restates the ``.in`` file of one reference test
(tests/runnable/tests/<test>/<name>.in) in the keys understood by
``Oracle.from_params`` and ``dlbc_amd.Lattice.from_params``; the golden data
live in tests/golden/<fixture>.npz (written by export_fixtures.py).
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "manifest.json")) as _fh:
    MANIFEST = json.load(_fh)

SC = {"gcc": [[0.0, 3.6], [3.6, 0.0]], "enableShanChen": True, "psiForm": "Exponential"}

CASES = {
    # lb/bdist2-3d/bdist2-3d.in
    "bdist2_3d": dict(d=3, q=19, gn=[4, 4, 4], components=2, tau=[1.0, 1.0],
                      eqDistForm="BDist2", fluidInit=["EqDistLamellae"] * 2,
                      fluidDensities=[[0.7, 0.0], [0.0, 0.7]], lamellaeWidths=[2, 2],
                      initAxis="Z", seed=42),
    # lb/bdist2-2d/bdist2-2d.in
    "bdist2_2d": dict(d=2, q=9, gn=[4, 4], components=2, tau=[1.0, 1.0],
                      eqDistForm="BDist2", fluidInit=["EqDistLamellae"] * 2,
                      fluidDensities=[[0.7, 0.0], [0.0, 0.7]], lamellaeWidths=[2, 2],
                      initAxis="X", seed=42),
    # lb/bdist2-sc-3d/bdist2-sc-3d.in (gn 64^3; x and y invariant)
    "bdist2_sc_3d": dict(d=3, q=19, gn=[64, 64, 64], components=2, tau=[1.0, 1.0],
                         eqDistForm="BDist2", fluidInit=["EqDistLamellae"] * 2,
                         fluidDensities=[[0.7, 0.04], [0.04, 0.7]], lamellaeWidths=[32, 32],
                         initAxis="Z", seed=42, **SC),
    # lb/bdist2-sc-2d/bdist2-sc-2d.in (gn 64^2; y invariant)
    "bdist2_sc_2d": dict(d=2, q=9, gn=[64, 64], components=2, tau=[1.0, 1.0],
                         eqDistForm="BDist2", fluidInit=["EqDistLamellae"] * 2,
                         fluidDensities=[[0.7, 0.04], [0.04, 0.7]], lamellaeWidths=[32, 32],
                         initAxis="X", seed=42, **SC),
    # lb/merging-droplets-2d/merging-droplets-2d.in
    "merging_droplets_2d": dict(d=2, q=9, gn=[64, 64], components=2, tau=[1.0, 1.0],
                                eqDistForm="BDist2", fluidInit=["EqDistTwoSpheresFrac"] * 2,
                                fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.2,
                                initOffset=[0.0, 0.0], initSeparation=[0.4, 0.0],
                                interfaceThickness=5.0, seed=42, **SC),
    # lb/merging-droplets-3d/merging-droplets-3d.in
    "merging_droplets_3d": dict(d=3, q=19, gn=[64, 64, 64], components=2, tau=[1.0, 1.0],
                                eqDistForm="BDist2", fluidInit=["EqDistTwoSpheresFrac"] * 2,
                                fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.2,
                                initOffset=[0.0, 0.0, 0.0], initSeparation=[0.4, 0.0, 0.0],
                                interfaceThickness=5.0, seed=42, **SC),
    # lb/laplace-droplet-3d/laplace-droplet-3d.in
    "laplace_droplet_3d": dict(d=3, q=19, gn=[64, 64, 64], components=2, tau=[1.0, 1.0],
                               eqDistForm="BDist2", fluidInit=["EqDistSphereFrac"] * 2,
                               fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.25,
                               interfaceThickness=5.0, **SC),
    # lb/laplace-droplet-2d/laplace-droplet-2d.in
    "laplace_droplet_2d": dict(d=2, q=9, gn=[64, 64], components=2, tau=[1.0, 1.0],
                               eqDistForm="BDist2", fluidInit=["EqDistSphereFrac"] * 2,
                               fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.25,
                               interfaceThickness=5.0, **SC),
    # lb/poiseuille-2d/poiseuille-2d.in (SecondOrder: the reference default)
    "poiseuille_2d": dict(d=2, q=9, gn=[32, 4], components=2, tau=[1.0, 1.0],
                          globalAcc=[0.0, 0.0001], fluidInit=["EqDist"] * 2,
                          fluidDensities=[[0.7], [1.0]], maskInit="Plates", maskAxis="X"),
    # lb/laplace-quasi-1d/laplace-quasi-1d-d2q9.in (ThirdOrder)
    "laplace_quasi_1d_d2q9": dict(d=2, q=9, gn=[64, 4], components=2, tau=[1.0, 1.0],
                                  eqDistForm="ThirdOrder", fluidInit=["EqDistLamellaeFrac"] * 2,
                                  fluidDensities=[[0.5, 0.7, 0.5], [0.7, 0.5, 0.7]],
                                  lamellaeWidths=[0.25, 0.50, 0.25], initRadius=0.25,
                                  interfaceThickness=5.0, **SC),
    # lb/laplace-quasi-1d/laplace-quasi-1d-d1q3.in (ThirdOrder)
    "laplace_quasi_1d_d1q3": dict(d=1, q=3, gn=[64], components=2, tau=[1.0, 1.0],
                                  eqDistForm="ThirdOrder", fluidInit=["EqDistSphereFrac"] * 2,
                                  fluidDensities=[[0.7, 0.5], [0.5, 0.7]], initRadius=0.25,
                                  interfaceThickness=5.0, **SC),
    # technical/checkpoint-3d/checkpoint-3d.in.source
    "checkpoint_3d": dict(d=3, q=19, gn=[16, 16, 16], components=2, tau=[1.0, 1.0],
                          eqDistForm="BDist2", fluidInit=["EqDistTwoSpheresFrac"] * 2,
                          fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.2,
                          initOffset=[0.0, 0.0, 0.0], initSeparation=[0.4, 0.0, 0.0],
                          interfaceThickness=5.0, seed=42, **SC),
    # lb/init-2d/fluid-init-eqdistperturb-2d.in: one rank's 32x32 share of 64x64
    "init_eqdistperturb_2d": dict(d=2, q=9, gn=[32, 32], components=2, tau=[1.0, 1.0],
                                  eqDistForm="BDist2", fluidInit=["EqDistPerturb"] * 2,
                                  fluidDensities=[[0.5], [1.0]], fluidPerturb=[0.1, 0.2],
                                  randomReseed=[True, True], randomSeed=[42, 43]),
}

CHECKPOINT_LOWDIM = {
    # technical/checkpoint-2d/checkpoint-2d.in.source
    "checkpoint_2d": dict(d=2, q=9, gn=[64, 64], components=2, tau=[1.0, 1.0],
                          eqDistForm="BDist2", fluidInit=["EqDistTwoSpheresFrac"] * 2,
                          fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.2,
                          initOffset=[0.0, 0.0], initSeparation=[0.4, 0.0],
                          interfaceThickness=5.0, seed=42, gcc=[[0.0, 3.6], [3.6, 0.0]],
                          enableShanChen=True, psiForm="Exponential"),
    # technical/checkpoint-1d/checkpoint-d1q3.in.source (BDist2 has no D1Q3 form:
    # eqdist.d:269-273 returns the population itself)
    "checkpoint_d1q3": dict(d=1, q=3, gn=[64], components=2, tau=[1.0, 1.0],
                            eqDistForm="BDist2", fluidInit=["EqDistSphereFrac"] * 2,
                            fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=0.2,
                            initOffset=[0.0], interfaceThickness=5.0, seed=42,
                            gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True,
                            psiForm="Exponential"),
}

# lb/init-2d/<name>.in: D2Q9, two components, t = 0.  Shapes on the reference's
# global coordinates, so one rank over the whole lattice writes what the four
# ranks of the reference wrote; the random kinds keep one rank's 32x32 share.
_SHAPE = dict(eqDistForm="BDist2", fluidDensities=[[1.0, 0.0], [0.0, 1.0]],
              interfaceThickness=4.0)
_RAND = dict(eqDistForm="BDist2", gn=[32, 32], fluidDensities=[[0.5], [1.0]],
             randomReseed=[True, True], randomSeed=[42, 43])
_MASK = dict(fluidInit=["None", "None"], fluidDensities=[[], []])
INIT_2D = {
    "fluid_init_const_2d": dict(fluidInit=["Const"] * 2, fluidDensities=[[0.5], [1.0]]),
    "fluid_init_constrandom_2d": dict(_RAND, fluidInit=["ConstRandom"] * 2, eqDistForm="SecondOrder"),
    "fluid_init_eqdist_2d": dict(eqDistForm="BDist2", fluidInit=["EqDist"] * 2,
                                 fluidDensities=[[0.5], [1.0]]),
    "fluid_init_eqdistrandom_2d": dict(_RAND, fluidInit=["EqDistRandom"] * 2),
    "fluid_init_eqdistperturb_2d": dict(_RAND, fluidInit=["EqDistPerturb"] * 2,
                                        fluidPerturb=[0.1, 0.2]),
    "fluid_init_eqdistperturbfrac_2d": dict(_RAND, fluidInit=["EqDistPerturbFrac"] * 2,
                                            fluidPerturb=[0.2, 0.2]),
    "fluid_init_eqdistsphere_2d": dict(_SHAPE, gn=[128, 64], fluidInit=["EqDistSphere"] * 2,
                                       initRadius=16.0, initOffset=[12.8, 0.0]),
    "fluid_init_eqdistspherefrac_2d": dict(_SHAPE, gn=[128, 64],
                                           fluidInit=["EqDistSphereFrac"] * 2,
                                           initRadius=0.25, initOffset=[0.1, 0.0]),
    "fluid_init_eqdistspherefrac_boundary_2d": dict(_SHAPE, gn=[128, 64],
                                                    fluidInit=["EqDistSphereFrac"] * 2,
                                                    initRadius=0.25, initOffset=[0.5, 0.5]),
    "fluid_init_eqdisttwospheres_2d": dict(_SHAPE, gn=[128, 64],
                                           fluidInit=["EqDistTwoSpheres"] * 2, initRadius=16.0,
                                           initOffset=[12.8, 0.0], initSeparation=[6.4, 0.0]),
    "fluid_init_eqdisttwospheresfrac_2d": dict(_SHAPE, gn=[128, 64],
                                               fluidInit=["EqDistTwoSpheresFrac"] * 2,
                                               initRadius=0.25, initOffset=[0.1, 0.0],
                                               initSeparation=[0.05, 0.0]),
    "fluid_init_none_2d": dict(_MASK),
    "mask_init_none_2d": dict(_MASK),
    "mask_init_box_2d": dict(_MASK, maskInit="Box"),
    "mask_init_walls_2d": dict(_MASK, maskInit="Walls",
                               dimensionHasWall=[[True, False], [False, True]], wallThickness=2),
}
for _ax in "XYZ":
    INIT_2D["fluid_init_eqdistcylinder_%s_2d" % _ax.lower()] = dict(
        _SHAPE, gn=[128, 64], fluidInit=["EqDistCylinder"] * 2, initRadius=16.0,
        initOffset=[12.8, 0.0], initAxis=_ax)
    # lb/init-2d/fluid-init-eqdistcylinderfrac-y-2d.in takes 0.125 (x is the shorter
    # axis other than y), the others 0.25
    INIT_2D["fluid_init_eqdistcylinderfrac_%s_2d" % _ax.lower()] = dict(
        _SHAPE, gn=[128, 64], fluidInit=["EqDistCylinderFrac"] * 2,
        initRadius=0.125 if _ax == "Y" else 0.25, initOffset=[0.1, 0.0], initAxis=_ax)
    INIT_2D["mask_init_tube_%s_2d" % _ax.lower()] = dict(_MASK, maskInit="Tube", maskAxis=_ax)
for _ax in "XY":
    INIT_2D["fluid_init_eqdistlamellae_%s_2d" % _ax.lower()] = dict(
        eqDistForm="BDist2", gn=[128, 64], fluidInit=["EqDistLamellae"] * 2,
        interfaceThickness=4.0, initAxis=_ax,
        **({"lamellaeWidths": [64.0, 32.0, 16.0, 8.0],
            "fluidDensities": [[1.0, 0.0, 1.0, 0.0], [1.0, 1.0, 0.0, 1.0]]} if _ax == "X" else
           {"lamellaeWidths": [32.0, 16.0, 16.0],
            "fluidDensities": [[1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]}))
    INIT_2D["fluid_init_eqdistlamellaefrac_%s_2d" % _ax.lower()] = dict(
        eqDistForm="BDist2", gn=[128, 64], fluidInit=["EqDistLamellaeFrac"] * 2,
        interfaceThickness=4.0, initAxis=_ax,
        **({"lamellaeWidths": [0.5, 0.25, 0.125, 0.0625],
            "fluidDensities": [[1.0, 0.0, 1.0, 0.0], [1.0, 1.0, 0.0, 1.0]]} if _ax == "X" else
           {"lamellaeWidths": [0.5, 0.25, 0.25],
            "fluidDensities": [[1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]}))
    INIT_2D["mask_init_plates_%s_2d" % _ax.lower()] = dict(_MASK, maskInit="Plates",
                                                           maskAxis=_ax, wallOffset=4)
# MaskInit.File and FluidInit.File read the very arrays the reference wrote out
# (mask-init-file-2d.in, fluid-init-file-2d.in), so these two take their input
# from the fixture itself.
INIT_2D["mask_init_file_2d"] = dict(_MASK)
INIT_2D["fluid_init_file_2d"] = dict(eqDistForm="BDist2", gn=[128, 64],
                                     fluidInit=["File"] * 2, fluidDensities=[[], []])


def init_2d_params(name):
    """Parameters of one lb/init-2d case (gn defaults to the 64x64 lattice)."""
    p = dict(d=2, q=9, gn=[64, 64], components=2, tau=[1.0, 1.0])
    p.update(INIT_2D[name])
    return p


def case(name, **over):
    p = dict(CASES[name])
    p.update(over)
    return p


def golden(name):
    """Load a fixture; returns (arrays, manifest entry)."""
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        arrays = {k: z[k] for k in z.files}
    return arrays, MANIFEST.get(name, {})


def expand(arr, full_shape):
    """Broadcast a fixture whose invariant axes were collapsed back to full shape."""
    return np.broadcast_to(arr, tuple(full_shape))


def tolerance(meta, t):
    tol = (meta.get("tolerance") or {}).get("t%d" % t)
    return 0.0 if tol is None else float(tol)


def assert_close(actual, expected, atol, what=""):
    """Absolute comparison like ``h5diff -d atol`` (atol 0: bit for bit)."""
    actual = np.asarray(actual)
    expected = np.asarray(expected)
    assert actual.shape == expected.shape, (what, actual.shape, expected.shape)
    if atol == 0.0:
        diff = actual.view(np.uint64) != expected.view(np.uint64) \
            if actual.dtype == np.float64 else actual != expected
        n = int(np.count_nonzero(diff))
        if n:
            d = np.abs(actual - expected)
            raise AssertionError("%s: %d of %d values differ (max |diff| %.3g)"
                                 % (what, n, actual.size, float(np.max(d))))
    else:
        d = np.abs(actual - expected)
        m = float(np.max(d)) if d.size else 0.0
        assert m <= atol, "%s: max |diff| %.3g > %.3g" % (what, m, atol)
