"""CPU-side checks of the product library (no GPU calls).

* libdlbc_hip.so loads and exports every function include/dlbc_hip.h declares;
* its host-side initial conditions (the reference's init shapes and seeded
  Mt19937 draws) equal the oracle's, which the reference's golden data pin;
* the package has no CPU fallback (it refuses to work without the library).
"""
import os
import re

import numpy as np
import pytest

from cases import CASES, case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    with open(os.path.join(ROOT, "include", "dlbc_hip.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dlbc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    import dlbc_amd
    path = dlbc_amd.lib_path()
    assert os.path.exists(path), "build the library first (dlbc_amd/build.py)"
    lib = ctypes.CDLL(path)
    names = _declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_init_equals_oracle(oracle_mod, name):
    """Product init (C++) == oracle init (C) bit for bit on every reference
    configuration, including the halo-inclusive seeded draw order."""
    import dlbc_amd as D
    p = case(name)
    o = oracle_mod.Oracle.from_params(p)
    d, q, n, nc = p["d"], p["q"], p["gn"], p["components"]
    mask = D.init_mask(d, n, p.get("maskInit", "None"), p.get("maskAxis", "X"),
                       p.get("wallOffset", 0), p.get("dimensionHasWall"), p.get("wallThickness", 1))
    phys = tuple(slice(2, 2 + v) for v in n)
    assert np.array_equal(mask[phys], o.mask())
    grng = D.Rng(p.get("seed", 0))
    for f in range(nc):
        rng = D.Rng(p["randomSeed"][f]) if p.get("randomReseed", [False] * nc)[f] else grng
        pop = D.init_fluid(d, q, n, p["fluidInit"][f], p["fluidDensities"][f],
                           p.get("eqDistForm", "SecondOrder"), p.get("fluidPerturb", [0.0] * nc)[f],
                           rng, p.get("initRadius", 0.0), p.get("initOffset"),
                           p.get("initSeparation"), p.get("initAxis", "X"),
                           p.get("interfaceThickness", 0.0), p.get("lamellaeWidths"))
        D.init_eq_dist_wall(pop, mask, d, q, n)
        assert np.array_equal(pop[phys].view(np.uint64), o.population(f).view(np.uint64))


@pytest.mark.parametrize("kind,extra", [
    ("Box", {}), ("Tube", {"axis": 1}), ("Plates", {"axis": 2, "wall_offset": 1}),
    ("Walls", {"dim_has_wall": [[1, 0], [0, 1], [1, 1]], "wall_thickness": 2})])
def test_mask_shapes_equal_oracle(oracle_mod, kind, extra):
    import dlbc_amd as D
    n = [7, 6, 9]
    m = D.init_mask(3, n, kind, **extra)
    o = oracle_mod.Oracle(3, 19, 1, n)
    o.init_mask(kind, extra.get("axis", 0), extra.get("wall_offset", 0),
                extra.get("dim_has_wall"), extra.get("wall_thickness", 1))
    raw = o.raw("mask")
    assert np.array_equal(m, raw)  # halo included (the reference's own pre-exchange state)


@pytest.mark.parametrize("kind", ["ConstRandom", "EqDistRandom", "EqDistPerturbFrac",
                                  "EqDistCylinder", "EqDistCylinderFrac", "EqDistTwoSpheres",
                                  "EqDistSphere", "Const", "EqDistLamellaeFrac"])
def test_other_fluid_inits_equal_oracle(oracle_mod, kind):
    import dlbc_amd as D
    n = [6, 7, 5]
    kw = dict(densities=[0.6, 0.2], perturb=0.05, radius=2.5, offset=[0.5, -1.0, 0.0],
              separation=[2.0, 0.0, 1.0], axis=1, thickness=2.0, widths=[0.25, 0.75])
    o = oracle_mod.Oracle(3, 19, 1, n)
    o.set_params([1.0], eq_dist_form=2)
    o.init_fluid(0, kind, kw["densities"], kw["perturb"], True, 99, kw["radius"], kw["offset"],
                 kw["separation"], kw["axis"], kw["thickness"], kw["widths"])
    pop = D.init_fluid(3, 19, n, kind, kw["densities"], "BDist2", kw["perturb"], D.Rng(99),
                       kw["radius"], kw["offset"], kw["separation"], kw["axis"], kw["thickness"],
                       kw["widths"])
    assert np.array_equal(pop.view(np.uint64), o.raw("fluid").view(np.uint64))


@pytest.mark.parametrize("kind", ["EqDistPerturb", "ConstRandom", "EqDistSphereFrac",
                                  "EqDistLamellaeFrac"])
def test_d1q5_inits_equal_oracle(oracle_mod, kind):
    """D1Q5 keeps a halo of 4 (src/dlbc/lattice.d:96-103): the random draws
    run over the 4-site halo, and the shapes see the 4-site offset."""
    import dlbc_amd as D
    n = [13]
    assert D.lattice.reference_halo(1, 5) == 4
    kw = dict(densities=[0.6, 0.2], perturb=0.05, radius=0.3, widths=[0.25, 0.75])
    o = oracle_mod.Oracle(1, 5, 1, n)
    o.set_params([1.0], eq_dist_form=1)
    o.init_fluid(0, kind, kw["densities"], kw["perturb"], True, 17, kw["radius"], None, None,
                 0, 2.0, kw["widths"])
    pop = D.init_fluid(1, 5, n, kind, kw["densities"], "ThirdOrder", kw["perturb"], D.Rng(17),
                       kw["radius"], None, None, 0, 2.0, kw["widths"], halo=4)
    assert np.array_equal(pop.view(np.uint64), o.raw("fluid").view(np.uint64))


def test_rng_matches_reference_golden_values():
    """random.d:55-89 golden values through the product's Mt19937."""
    import dlbc_amd as D
    r = D.Rng(42)
    assert [r.uniform(0.0, 2.0) for _ in range(3)] == [
        0.7490802287936861869610538633423857390880584716796875,
        1.5930859687722020989752991226851008832454681396484375,
        1.9014286235676676195538448155275546014308929443359375]


def test_no_cpu_fallback(monkeypatch):
    """Without the shared library the package raises instead of computing."""
    import dlbc_amd.lattice as lat
    monkeypatch.setattr(lat, "_lib", None)
    monkeypatch.setattr(lat, "_LIB", "/nonexistent/libdlbc_hip.so")
    with pytest.raises(lat.DlbcError, match="not built"):
        lat.lib()


def test_exact_constant_division_selftest():
    """The equilibria divide by the c_s^2 products with an FMA-based exact
    quotient (Markstein's correction) instead of IEEE division; on 4 million
    operands, half of them within 4 ulp of an exact multiple of a divisor,
    it equals IEEE division bit for bit."""
    import dlbc_amd as D
    assert D.lib().dlbc_selftest_division(4_000_000, 2026) == 0


def test_site_division_selftest():
    """Site coordinates come from a multiply and a shift by per-lattice
    constants (m = ceil(2^(32+l) / d), 2^l >= d) instead of integer division;
    on 4 million (site, divisor) pairs over the range of lattice indices, half
    of them next to exact multiples, it equals integer division."""
    import dlbc_amd as D
    assert D.lib().dlbc_selftest_site_division(4_000_000, 2027) == 0


@pytest.mark.parametrize("d,q,nc,n", [(2, 9, 2, [64, 64]), (3, 19, 1, [8, 8, 32]),
                                      (3, 19, 2, [8, 3, 32]), (3, 19, 2, [8, 8, 2])])
def test_shifted_storage_shape_checks(d, q, nc, n):
    """The shifted storages exist for the tiled D3Q19 two-component kernel
    (any n >= 4 per axis); other lattices are refused before any device
    call."""
    import dlbc_amd as D
    with pytest.raises(D.DlbcError, match="unsupported"):
        D.Lattice(d, q, nc, n, storage="f32_shift")


def test_invalid_arguments_are_rejected_on_the_host():
    """Entry points check their arguments before any device work: a null
    context (or batch) gives DLBC_ERR_ARG, and the library never exits."""
    import ctypes
    import dlbc_amd
    lib = ctypes.CDLL(dlbc_amd.lib_path())
    ERR_ARG = -1
    for name, args in [("dlbc_gpu_step", (None, 1)), ("dlbc_gpu_step_graph", (None, 1)),
                       ("dlbc_gpu_advect", (None,)), ("dlbc_gpu_synchronize", (None,)),
                       ("dlbc_gpu_set_launch", (None, 0)),
                       ("dlbc_gpu_set_block_order", (None, 0)),
                       ("dlbc_gpu_set_schedule", (None, 0)),
                       ("dlbc_batch_step", (None, 1, 1))]:
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        assert fn(*args) == ERR_ARG, name
    out = ctypes.c_void_p()
    create = lib.dlbc_gpu_create
    create.restype = ctypes.c_int
    n = (ctypes.c_int64 * 3)(8, 8, 8)
    ERR_UNSUPPORTED = -4
    for d, q, comps, storage, rc in [(3, 15, 2, 0, ERR_UNSUPPORTED), (3, 19, 0, 0, ERR_UNSUPPORTED),
                                     (3, 19, 5, 0, ERR_UNSUPPORTED), (3, 19, 2, 7, ERR_ARG),
                                     (2, 9, 2, 1, ERR_UNSUPPORTED)]:
        assert create(ctypes.byref(out), d, q, comps, n, storage, 0) == rc, (d, q, comps, storage)
        assert out.value is None
    small = (ctypes.c_int64 * 3)(1, 8, 8)  # below the reference's two periodic images
    assert create(ctypes.byref(out), 3, 19, 2, small, 0, 0) == ERR_ARG
    d1q5 = (ctypes.c_int64 * 1)(3)  # D1Q5 needs n >= 4 (halo 4)
    assert create(ctypes.byref(out), 1, 5, 2, d1q5, 0, 0) == ERR_ARG
    huge = (ctypes.c_int64 * 3)(2048, 1024, 1024)  # 2^31 sites: beyond the 2^30-site limit
    assert create(ctypes.byref(out), 3, 19, 2, huge, 0, 0) == ERR_UNSUPPORTED
    long_axis = (ctypes.c_int64 * 3)((1 << 20) + 1, 2, 2)  # an axis beyond 2^20
    assert create(ctypes.byref(out), 3, 19, 2, long_axis, 0, 0) == ERR_ARG
    assert out.value is None


def host_init_from_params(D, p, halo=2):
    """The product's host initialisation (initMask, initFluids with
    initEqDistWall, src/main.d:136-138) of a parameter set: the mask and the
    halo-padded population arrays."""
    d, q, n, nc = p["d"], p["q"], p["gn"], p["components"]
    mask = D.init_mask(d, n, p.get("maskInit", "None"), p.get("maskAxis", "X"),
                       p.get("wallOffset", 0), p.get("dimensionHasWall"),
                       p.get("wallThickness", 1), halo=halo)
    if "maskArray" in p:
        mask[...] = 0
        mask[tuple(slice(halo, halo + v) for v in n)] = p["maskArray"]
    grng = D.Rng(p.get("seed", 0))
    pops = []
    for f in range(nc):
        rng = D.Rng(p["randomSeed"][f]) if p.get("randomReseed", [False] * nc)[f] else grng
        pop = D.init_fluid(d, q, n, p["fluidInit"][f], p["fluidDensities"][f],
                           p.get("eqDistForm", "SecondOrder"), p.get("fluidPerturb", [0.0] * nc)[f],
                           rng, p.get("initRadius", 0.0), p.get("initOffset"),
                           p.get("initSeparation"), p.get("initAxis", "X"),
                           p.get("interfaceThickness", 0.0), p.get("lamellaeWidths"), halo=halo)
        D.init_eq_dist_wall(pop, mask, d, q, n, halo=halo)
        pops.append(pop)
    return mask, pops


def _init_2d_names():
    from cases import INIT_2D
    return sorted(INIT_2D)


@pytest.mark.parametrize("name", _init_2d_names())
def test_host_init_equals_init_2d_fixture(name):
    """The product's host initialisation against the reference's own output
    of every lb/init-2d case (every FluidInit and MaskInit kind), bit for bit:
    populations (D's default NaN included), mask.  The File kinds read the
    fixture, as the reference reads its own output there."""
    import dlbc_amd as D
    from cases import golden, init_2d_params
    g, _ = golden("init_2d")
    p = init_2d_params(name)
    if name == "mask_init_file_2d":
        p["maskArray"] = g[name + "__mask"]
    mask, pops = host_init_from_params(D, p)
    phys = tuple(slice(2, 2 + v) for v in p["gn"])
    assert np.array_equal(mask[phys], g[name + "__mask"])
    for f, c in enumerate(("red", "blue")):
        want = g["%s__population_%s" % (name, c)]
        if name == "fluid_init_file_2d":
            continue  # read back from the file itself: nothing computed
        assert np.array_equal(pops[f][phys].view(np.uint64), want.view(np.uint64)), (name, c)
