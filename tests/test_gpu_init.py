"""Initial conditions computed on the device (dlbc_gpu_init_fluid,
dlbc_gpu_init_mask, dlbc_gpu_init_eq_dist_wall) against the host
initialisation followed by an upload, bit for bit.

The host initialisation itself is pinned to the oracle and the reference's
fixtures by the CPU tests (test_library_host.py); every GPU parity test that
starts from Lattice.from_params also starts from the device initialisation.
"""
import numpy as np
import pytest

from cases import case

pytestmark = pytest.mark.gpu

KW = dict(densities=[0.6, 0.2], perturb=0.05, radius=2.5, offset=[0.5, -1.0, 0.25],
          separation=[2.0, 0.0, 1.0], axis=1, thickness=2.0, widths=[0.25, 0.75])
LAMELLAE_KW = dict(KW, densities=[0.6, 0.2, 0.9], widths=[2.0, 3.0, 2.0])
KINDS = ["Const", "EqDist", "EqDistPerturb", "EqDistPerturbFrac", "ConstRandom", "EqDistRandom",
         "EqDistSphere", "EqDistSphereFrac", "EqDistCylinder", "EqDistCylinderFrac",
         "EqDistTwoSpheres", "EqDistTwoSpheresFrac", "EqDistLamellae", "EqDistLamellaeFrac"]


@pytest.fixture(scope="module")
def D():
    import dlbc_amd
    import torch  # noqa: F401  (device visibility check)
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    return dlbc_amd


def _kw(kind, d):
    kw = dict(LAMELLAE_KW if kind == "EqDistLamellae" else KW)
    kw["axis"] = min(kw["axis"], d - 1)
    return kw


def _host_and_device(D, d, q, n, kind, form, storage="f64", halo=2, seed=99):
    """Two lattices, one from the host initialisation uploaded, one
    initialised on the device; shifted storage (two components only) gets
    both components from one generator."""
    kw = _kw(kind, d)
    args = (kind, kw["densities"], form, kw["perturb"])
    rest = (kw["radius"], kw["offset"][:d], kw["separation"][:d], kw["axis"], kw["thickness"],
            kw["widths"])
    nc = 1 if storage == "f64" else 2
    host = D.Lattice(d, q, nc, n, storage=storage)
    dev = D.Lattice(d, q, nc, n, storage=storage)
    for L in (host, dev):
        L.set_params([1.0] * nc, eq_dist_form=form,
                     rho0=[0.7, 0.3] if storage != "f64" else None)
    rh, rd = D.Rng(seed), D.Rng(seed)
    for f in range(nc):
        host.upload_population(f, D.init_fluid(d, q, n, *args, rh, *rest, halo=halo), halo=halo)
        dev.init_fluid(f, *args, rd, *rest, halo=halo)
    return host, dev


def _same_pop(a, b, f=None, what=""):
    for ff in (range(a.nc) if f is None else [f]):
        x, y = a.population(ff), b.population(ff)
        bad = np.nonzero(x.view(np.uint64) != y.view(np.uint64))
        assert bad[0].size == 0, "%s (component %d): %d differences, first at %s: %r vs %r" % (
            what, ff, bad[0].size, tuple(v[0] for v in bad), x[tuple(v[0] for v in bad)],
            y[tuple(v[0] for v in bad)])


@pytest.mark.parametrize("kind", KINDS)
def test_fluid_kinds_d3q19_f64(D, kind):
    host, dev = _host_and_device(D, 3, 19, [6, 7, 5], kind, "BDist2")
    _same_pop(host, dev, what=kind)


@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
@pytest.mark.parametrize("kind", ["EqDistPerturb", "ConstRandom", "EqDistSphereFrac",
                                  "EqDistLamellae"])
def test_fluid_kinds_shifted_storage(D, kind, storage):
    """Shifted storage keeps f - w rho0 rounded once to the storage type on
    both paths."""
    host, dev = _host_and_device(D, 3, 19, [6, 7, 5], kind, "SecondOrder", storage)
    _same_pop(host, dev, what="%s %s" % (kind, storage))


@pytest.mark.parametrize("dq,n", [((2, 9), [9, 7]), ((1, 3), [11])])
@pytest.mark.parametrize("kind", ["EqDistPerturb", "ConstRandom", "EqDistSphere",
                                  "EqDistTwoSpheresFrac", "EqDistLamellaeFrac"])
def test_fluid_kinds_lower_dimensions(D, dq, n, kind):
    """Random draws over a halo-padded 2-D/1-D field and shapes with unused axes."""
    d, q = dq
    host, dev = _host_and_device(D, d, q, n, kind, "ThirdOrder")
    _same_pop(host, dev, what="%s D%dQ%d" % (kind, d, q))


@pytest.mark.parametrize("halo", [0, 1, 3])
def test_random_draw_order_follows_the_halo(D, halo):
    host, dev = _host_and_device(D, 3, 19, [5, 4, 6], "EqDistRandom", "SecondOrder", halo=halo)
    _same_pop(host, dev, what="halo %d" % halo)


def test_random_draws_across_staging_chunks(D):
    """ConstRandom over 120 x-planes of 64 x 64 sites is handed over in two
    chunks (64 MiB of draws each at most); the draw sequence runs on."""
    host, dev = _host_and_device(D, 3, 19, [120, 64, 64], "ConstRandom", "SecondOrder")
    _same_pop(host, dev, what="chunked ConstRandom")


def test_consecutive_components_share_the_generator(D):
    """Component 1 continues the sequence component 0 left off (one
    generator across components, as without randomReseed)."""
    n = [5, 6, 7]
    host = D.Lattice(3, 19, 2, n)
    dev = D.Lattice(3, 19, 2, n)
    r = D.Rng(7)
    for f in range(2):
        host.upload_population(f, D.init_fluid(3, 19, n, "EqDistPerturb", 0.7, "BDist2", 0.01, r),
                               halo=2)
    r = D.Rng(7)
    for f in range(2):
        dev.init_fluid(f, "EqDistPerturb", 0.7, "BDist2", 0.01, r)
    for f in range(2):
        _same_pop(host, dev, f, "component %d" % f)


@pytest.mark.parametrize("kind,extra", [
    ("Box", {}), ("Tube", {"axis": 1}), ("Plates", {"axis": 2, "wall_offset": 1}),
    ("Walls", {"dim_has_wall": [[1, 0], [0, 1], [1, 1]], "wall_thickness": 2})])
def test_mask_and_wall_populations(D, kind, extra):
    """initMask and initEqDistWall on the device: the same solid sites (seen
    through the wall populations [1, 0, ...] and through three Shan-Chen
    steps with bounce-back) as the host mask uploaded."""
    n = [7, 6, 9]
    p = dict(tau=[1.0, 1.0], gcc=[[0.0, 3.6], [3.6, 0.0]], gwc=[0.2, -0.1])
    mask = D.init_mask(3, n, kind, **extra)
    r = D.Rng(5)
    pops = []
    for f in range(2):
        pop = D.init_fluid(3, 19, n, "EqDistPerturb", 0.7, "BDist2", 0.01, r)
        D.init_eq_dist_wall(pop, mask, 3, 19, n)
        pops.append(pop)
    host = D.Lattice(3, 19, 2, n)
    dev = D.Lattice(3, 19, 2, n)
    for L in (host, dev):
        L.set_params(p["tau"], p["gcc"], p["gwc"], None, None, "BDist2", "Exponential", True)
    host.upload_mask(mask, halo=2)
    for f in range(2):
        host.upload_population(f, pops[f], halo=2)
    dev.init_mask(kind, **extra)
    r = D.Rng(5)
    for f in range(2):
        dev.init_fluid(f, "EqDistPerturb", 0.7, "BDist2", 0.01, r)
        dev.init_eq_dist_wall(f)
    solid = mask[2:-2, 2:-2, 2:-2] == 1
    assert solid.any()
    for f in range(2):
        _same_pop(host, dev, f, "%s t0" % kind)
        walls = dev.population(f)[solid]
        assert np.all(walls[:, 0] == 1.0) and np.all(walls[:, 1:] == 0.0)
    host.step(3)
    dev.step(3)
    for f in range(2):
        _same_pop(host, dev, f, "%s t3" % kind)


@pytest.mark.parametrize("name", ["bdist2_sc_3d", "checkpoint_3d", "poiseuille_2d",
                                  "init_eqdistperturb_2d", "laplace_quasi_1d_d1q3"])
def test_from_params_device_and_host_init_agree(D, name):
    """The start-up sequence with device initialisation equals the host one
    after five steps."""
    p = case(name)
    a = D.Lattice.from_params(p, device_init=True)
    b = D.Lattice.from_params(p, device_init=False)
    a.step(5)
    b.step(5)
    for f in range(a.nc):
        _same_pop(a, b, f, "%s component %d" % (name, f))


def test_refusals(D):
    L = D.Lattice(3, 19, 1, [4, 4, 4])
    with pytest.raises(D.DlbcError):
        L.init_fluid(0, 99, [1.0])  # unknown kind
    with pytest.raises(D.DlbcError):
        L.init_fluid(1, "EqDist", [1.0])  # no component 1
    with pytest.raises(D.DlbcError):
        L.init_fluid(0, "EqDistSphere", [1.0])  # a sphere needs two densities
    with pytest.raises(D.DlbcError):
        L.init_fluid(0, "EqDistLamellae", [0.1] * 65, widths=[1.0] * 65)  # above 64 values
    with pytest.raises(D.DlbcError):
        L.init_mask("Plates", axis=3)
    with pytest.raises(D.DlbcError):
        L.init_eq_dist_wall(2)
    with pytest.raises(D.DlbcError):
        L.step(1)  # nothing initialised yet


def _init_2d_names():
    from cases import INIT_2D
    return sorted(INIT_2D)


@pytest.mark.parametrize("name", _init_2d_names())
def test_init_2d_fixture_on_device(D, name):
    """Every lb/init-2d case of the reference (each FluidInit and MaskInit
    kind), initialised on the device and through the host initialisation plus
    upload: populations (D's default NaN included), mask and colour equal the
    reference's own t = 0 output bit for bit.  The File kinds take the
    fixture as the file the reference reads."""
    from cases import assert_close, golden, init_2d_params
    g, _ = golden("init_2d")
    p = init_2d_params(name)
    if name == "mask_init_file_2d":
        p["maskArray"] = g[name + "__mask"]
    for device_init in (True, False):
        L = D.Lattice.from_params(p, device_init=device_init)
        if name == "fluid_init_file_2d":
            for f, c in enumerate(("red", "blue")):
                L.upload_population(f, g["%s__population_%s" % (name, c)])
        what = "%s (%s)" % (name, "device" if device_init else "host + upload")
        for f, c in enumerate(("red", "blue")):
            assert_close(L.population(f), g["%s__population_%s" % (name, c)], 0.0,
                         "%s %s" % (what, c))
        assert np.array_equal(L.mask(), g[name + "__mask"]), what + " mask"
        # colour is computed from the populations: where the reference's is NaN
        # (uninitialised components) the GPU's must be NaN too, with whatever
        # payload its arithmetic gives; elsewhere bit for bit
        col, want = L.colour(), g[name + "__colour"]
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(col), nan), what + " colour NaN sites"
        assert_close(np.where(nan, 0.0, col), np.where(nan, 0.0, want), 0.0, what + " colour")
        L.close()
