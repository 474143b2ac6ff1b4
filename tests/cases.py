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
