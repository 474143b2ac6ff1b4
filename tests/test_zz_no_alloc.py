"""dlbc_gpu_total_mass_noalloc (no device allocation: the block partials live
in a fixed per-GPU array) against dlbc_gpu_total_mass, bit for bit, through
the C ABI: two calls in a row, the second on a smaller lattice (the shared
array keeps the first call's partials), then a lattice beyond its limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import dlbc_amd
    import torch  # noqa: F401  (device visibility check)
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    return dlbc_amd


def _both(D, L, f):
    lib = D.lib()
    ref = np.zeros(2)
    got = np.full(2, np.nan)
    assert lib.dlbc_gpu_total_mass(L.h, f, ref) == 0, lib.dlbc_gpu_last_error(L.h)
    rc = lib.dlbc_gpu_total_mass_noalloc(L.h, f, got)
    assert rc == 0, (rc, lib.dlbc_gpu_last_error(L.h))
    return ref, got


def _droplet(n, storage):
    return dict(d=3, q=19, gn=n, components=2, tau=[1.0, 1.0], eqDistForm="BDist2",
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistSphereFrac"] * 2, fluidDensities=[[0.7, 0.04], [0.04, 0.7]],
                initRadius=0.3, interfaceThickness=5.0, rho0=[0.37, 0.37],
                maskInit="Plates", maskAxis="Z")


@pytest.mark.parametrize("storage", ["f64", "f32_shift"])
def test_total_mass_noalloc_equals_total_mass(D, storage):
    # the first lattice spans 7 partials of 65536 sites, the second 1, the
    # third (D2Q9, one component) 1 again
    big = D.Lattice.from_params(_droplet([96, 80, 60], storage), storage=storage)
    small = D.Lattice.from_params(_droplet([20, 18, 16], storage), storage=storage)
    big.step(7)
    small.step(3)
    for L in (big, small):
        print("noalloc total mass, %s, lattice %s" % (storage, L.n), flush=True)
        for f in range(2):
            ref, got = _both(D, L, f)
            assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), (L.n, f, ref, got)
    if storage == "f64":
        p = dict(d=2, q=9, gn=[31, 17], components=1, tau=[1.0], fluidInit=["EqDistPerturb"],
                 fluidDensities=[[1.0]], fluidPerturb=[1e-3], randomReseed=[True],
                 randomSeed=[42])
        L = D.Lattice.from_params(p)
        L.step(5)
        ref, got = _both(D, L, 0)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), (ref, got)


def test_total_mass_noalloc_rejects_lattices_beyond_its_scratch(D):
    """268,369,920 sites is the limit; a D2Q9 FP64 lattice of 16384 x 16385
    (268,451,840 sites, about 43 GB) is refused with DLBC_ERR_UNSUPPORTED,
    before any device work, while dlbc_gpu_total_mass still serves it."""
    import dlbc_amd.lattice as LM
    lib = D.lib()
    h = ctypes.c_void_p()
    n = np.array([16384, 16385], np.int64)
    rc = lib.dlbc_gpu_create(ctypes.byref(h), 2, 9, 1, n, LM.STORAGE["f64"], 0)
    assert rc == 0
    try:
        out = np.full(2, np.nan)
        rc = lib.dlbc_gpu_total_mass_noalloc(h, 0, out)
        assert rc == -4, rc
        assert b"exceed" in lib.dlbc_gpu_last_error(h)
        assert np.isnan(out).all()
        assert lib.dlbc_gpu_total_mass_noalloc(h, 1, out) == -1  # no component 1
    finally:
        lib.dlbc_gpu_destroy(h)
