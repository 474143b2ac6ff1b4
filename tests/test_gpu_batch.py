"""Batches of lattices (one persistent launch per step over many lattices)
and the tiled shifted-storage kernel on lattices of any size.

A lattice stepped inside a batch must give the same bits as the same
lattice stepped alone, whatever the tile height, chunk length, workgroup
count, graph replay or per-lattice loop; the tiled kernel on unaligned
lattices is compared with the FP64 path (itself bit-identical to the oracle)
within the FP32 bounds of test_gpu_shifted.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import dlbc_amd
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    return dlbc_amd


def droplet(n, tau, g, frac):
    return dict(d=3, q=19, gn=list(n), components=2, tau=[tau, tau], eqDistForm="BDist2",
                gcc=[[0.0, g], [g, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistSphereFrac"] * 2, fluidDensities=[[0.7, 0.04], [0.04, 0.7]],
                initRadius=frac, interfaceThickness=5.0, rho0=[0.37, 0.37])


# uneven shapes: partial tiles in y and z, odd lengths, one aligned lattice
SHAPES = [((33, 40, 37), 1.0, 3.6, 0.25), ((20, 36, 52), 1.3, 4.2, 0.3),
          ((32, 32, 64), 1.1, 3.9, 0.2), ((17, 9, 70), 1.45, 4.5, 0.33)]


def _members(D, storage="f32_shift"):
    return [D.Lattice.from_params(droplet(*s), storage=storage) for s in SHAPES]


def _pops(lats):
    return [[L.population(f) for f in range(2)] for L in lats]


def _same(a, b, what):
    for k, (pa, pb) in enumerate(zip(a, b)):
        for f in range(2):
            assert np.array_equal(pa[f].view(np.uint64), pb[f].view(np.uint64)), (what, k, f)


@pytest.mark.parametrize("kernel", [0, 2], ids=["tiled", "split"])
@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
def test_batch_equals_lattices_stepped_alone(D, storage, kernel):
    """11 steps (odd: the buffers end swapped) in one batch, graph replay on,
    equal bit for bit to stepping every lattice alone; the per-lattice loop
    gives the same bits too."""
    alone = _members(D, storage)
    for L in alone:
        L.step(11)
    ref = _pops(alone)
    lats = _members(D, storage)
    B = D.Batch(lats)
    B.set_config(8, 1 << 20, 512, kernel)
    B.step(11, graph=True)
    _same(_pops(lats), ref, "batch")
    lats2 = _members(D, storage)
    B2 = D.Batch(lats2)
    B2.set_config(8, 1 << 20, 512, kernel)
    B2.step_looped(11)
    _same(_pops(lats2), ref, "looped")


@pytest.mark.parametrize("kernel", [0, 2, 3], ids=["tiled", "split", "split-persistent"])
def test_batch_graph_replays(D, kernel):
    """The captured two-step graph replayed several times, with eager work
    (field downloads, an eager step, set_params) between the replays: every
    replay matches the same sequence run eagerly."""
    a, b = _members(D), _members(D)
    Ba, Bb = D.Batch(a), D.Batch(b)
    Ba.set_config(8, 1 << 20, 512, kernel)
    Bb.set_config(8, 1 << 20, 512, kernel)
    for k in range(4):
        Ba.step(4, graph=True)
        Bb.step(4, graph=False)
        _same(_pops(a), _pops(b), "replay %d" % k)
        for L in a + b:  # eager work between the replays
            L.density(0)
        Ba.step(1, graph=True)
        Bb.step(1, graph=False)
        _same(_pops(a), _pops(b), "odd step %d" % k)
    a[1].set_params([1.2, 1.2], [[0.0, 4.0], [4.0, 0.0]], None, None, None, "BDist2",
                    "Exponential", True, [0.37, 0.37])
    b[1].set_params([1.2, 1.2], [[0.0, 4.0], [4.0, 0.0]], None, None, None, "BDist2",
                    "Exponential", True, [0.37, 0.37])
    Ba.step(6, graph=True)
    Bb.step(6, graph=False)
    _same(_pops(a), _pops(b), "after set_params")


def test_batch_config_and_autotune_do_not_change_results(D):
    ref = None
    for cfg in [(8, 1 << 20, 512, 0), (4, 16, 64, 1), (8, 8, 2048, 1), (4, 32, 8, 0),
                (8, 16, 256, 2), (8, 16, 256, 3)]:
        lats = _members(D)
        B = D.Batch(lats)
        B.set_config(*cfg)
        B.step(5)
        got = _pops(lats)
        if ref is None:
            ref = got
        else:
            _same(got, ref, "config %s" % (cfg,))
    lats = _members(D)
    B = D.Batch(lats)
    cfg, ms = B.autotune()
    assert cfg[0] in (4, 8) and ms > 0
    B.step(5)
    _same(_pops(lats), ref, "autotuned %s" % (cfg,))


def test_unaligned_shifted_lattice_close_to_fp64(D):
    """A lattice whose y and z lengths are not multiples of the tile (partial
    tiles on both axes) against the FP64 path: the same bounds as the
    aligned FP32 droplet tests."""
    p = droplet((44, 42, 50), 1.0, 3.6, 0.25)
    ref = D.Lattice.from_params(p)
    L = D.Lattice.from_params(p, storage="f32_shift")
    m0 = [L.total_mass(f) for f in range(2)]
    ref.step(500)
    L.step(500)
    for f in range(2):
        assert np.abs(L.density(f) - ref.density(f)).max() <= 3e-5
        assert np.abs(L.velocity(f) - ref.velocity(f)).max() <= 1e-5
        assert abs(L.total_mass(f) - m0[f]) <= 1e-4 * m0[f]


def test_batch_refuses_mixed_members(D):
    a = D.Lattice.from_params(droplet((16, 16, 32), 1.0, 3.6, 0.25), storage="f32_shift")
    b = D.Lattice.from_params(droplet((16, 16, 32), 1.0, 3.6, 0.25), storage="f16_shift")
    c = D.Lattice.from_params(droplet((16, 16, 32), 1.0, 3.6, 0.25))
    for members in ([a, b], [a, c], [a, a]):
        with pytest.raises(D.DlbcError):
            D.Batch(members)


@pytest.mark.parametrize("n", [(20, 27, 45), (6, 7, 256), (4, 5, 512)],
                         ids=["unaligned", "z-rows", "z-rows-512"])
@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
@pytest.mark.parametrize("walls", [False, True], ids=["droplet", "walls"])
def test_shifted_split_schedule_equals_tiled(D, storage, walls, n):
    """Shifted storage under the two-pass split schedule gives the same bits
    as the tiled kernel (psi through the P array instead of an LDS ring; the
    per-site arithmetic is shared), with walls, random solids, wall coupling
    and a body force: on an unaligned lattice, and on ones whose z-rows hold
    whole blocks, where the split kernel stages the psi rows in LDS."""
    p = droplet(n, 1.2, 3.8, 0.3)
    if walls:
        rng = np.random.default_rng(5)
        mask = (rng.random(n) < 0.1).astype(np.int32)
        mask[:, :, 0] = 1
        mask[:, :, -1] = 1
        p = dict(p, maskArray=mask, gwc=[0.2, -0.1], bodyForce=[[0.0, 0.0, 0.0], [1e-5, 0, 0]])
    a = D.Lattice.from_params(p, storage=storage, schedule="fused")
    b = D.Lattice.from_params(p, storage=storage, schedule="split")
    a.step(25)
    b.step(25)
    for f in range(2):
        assert np.array_equal(a.population(f).view(np.uint64), b.population(f).view(np.uint64))


@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
def test_shifted_split_whole_solid_lines_equal_tiled(D, storage):
    """Solid boxes covering whole cache lines along z (16 FP32 or 32 FP16
    site pairs a line), which the split collide pass neither reads nor writes,
    and lines of which only part is solid, whose solid sites it copies: the
    same bits as the tiled kernel, which never writes a solid site."""
    n = (6, 7, 256)
    p = droplet(n, 1.2, 3.8, 0.3)
    mask = np.zeros(n, dtype=np.int32)
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    mask[1:3, 2:5, 64:128] = 1
    mask[4, :, 150:157] = 1
    p = dict(p, maskArray=mask, gwc=[0.2, -0.1])
    a = D.Lattice.from_params(p, storage=storage, schedule="fused")
    b = D.Lattice.from_params(p, storage=storage, schedule="split")
    a.step(25)
    b.step(25)
    for f in range(2):
        assert np.array_equal(a.population(f).view(np.uint64), b.population(f).view(np.uint64))


@pytest.mark.parametrize("form,psi,gself", [("SecondOrder", "Linear", 0.0),
                                             ("ThirdOrder", "Exponential", 0.4),
                                             ("BDist2", "Linear", 0.3)],
                         ids=["second-linear", "third-exp-self", "bdist2-linear-self"])
@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
def test_shifted_forms_split_equals_tiled(D, storage, form, psi, gself):
    """The other equilibria (SecondOrder, ThirdOrder: the weights' float-float
    products on rho t + (rho - rho0)), psi = rho, and self-couplings g_ff != 0
    (the second Shan-Chen loop) under the split and tiled schedules: the same
    bits, with walls, random solids and a wall coupling; FP32 also stays within
    2e-4 of the FP64 lattice in density."""
    n = (8, 9, 256)
    rng = np.random.default_rng(9)
    mask = (rng.random(n) < 0.08).astype(np.int32)
    mask[:, :, 0] = 1
    p = dict(droplet(n, 1.1, 2.5, 0.3), eqDistForm=form, psiForm=psi,
             gcc=[[gself, 2.5], [2.5, 0.5 * gself]], maskArray=mask, gwc=[0.1, -0.05])
    a = D.Lattice.from_params(p, storage=storage, schedule="fused")
    b = D.Lattice.from_params(p, storage=storage, schedule="split")
    a.step(20)
    b.step(20)
    for f in range(2):
        assert np.array_equal(a.population(f).view(np.uint64), b.population(f).view(np.uint64))
    if storage == "f32_shift":
        c = D.Lattice.from_params(p)
        c.step(20)
        fluid = mask == 0
        for f in range(2):
            d = np.abs(b.density(f)[fluid] - c.density(f)[fluid]).max()
            assert d < 2e-4, (f, d)


def test_config5_ensemble_16_lattices(D):
    """The batched-ensemble workload cut to its first 16 lattices (L from 32
    to 96, tau, g and radius per lattice from the workload's seed), FP32
    shifted storage, 50 steps: the autotuned batch replayed from its graph
    equals each lattice stepped alone and the per-lattice loop bit for bit,
    and every lattice stays within the FP32 bounds of its own FP64 run
    (largest density and velocity differences; measured on MI355X about a
    third of the bounds)."""
    import bench
    params = bench.ensemble_params(16)
    batch = [D.Lattice.from_params(p, storage="f32_shift") for p in params]
    B = D.Batch(batch)
    cfg, _ = B.autotune()
    B.step(50, graph=True)
    B.synchronize()
    got = _pops(batch)
    alone = [D.Lattice.from_params(p, storage="f32_shift") for p in params]
    for L in alone:
        L.step(50)
    _same(got, _pops(alone), "graph batch vs alone (config %s)" % (cfg,))
    looped = [D.Lattice.from_params(p, storage="f32_shift") for p in params]
    C = D.Batch(looped)
    C.set_config(*cfg)
    C.step_looped(50)
    C.synchronize()
    _same(got, _pops(looped), "graph batch vs looped")
    worst = {"drho": 0.0, "du": 0.0}
    for k, p in enumerate(params):
        ref = D.Lattice.from_params(p)
        ref.step(50)
        for f in range(2):
            worst["drho"] = max(worst["drho"],
                                float(np.abs(batch[k].density(f) - ref.density(f)).max()))
            worst["du"] = max(worst["du"],
                              float(np.abs(batch[k].velocity(f) - ref.velocity(f)).max()))
        ref.close()
    print("config5 16 lattices, 50 steps, FP32 vs FP64: %s (autotuned %s)" % (worst, cfg))
    assert worst["drho"] <= 3e-6 and worst["du"] <= 3e-6, worst


@pytest.mark.parametrize("kernel", [1, 2, 3], ids=["tiled-no-solid", "split", "split-persistent"])
def test_batch_graph_after_member_gains_solids(D, kernel):
    """The captured batch graph is dropped when a member gains solid sites (the
    tiled kernel without bounce-back was captured) or when Shan-Chen is turned
    on for all members: graph replays keep matching each lattice stepped
    alone through the same sequence."""
    batch, alone = _members(D), _members(D)
    B = D.Batch(batch)
    B.set_config(8, 1 << 20, 512, kernel)
    B.step(4, graph=True)
    for L in alone:
        L.step(4)
    _same(_pops(batch), _pops(alone), "before walls")
    walls = [[False, False], [False, False], [True, True]]
    for L in (batch[2], alone[2]):
        L.init_mask("Walls", dim_has_wall=walls)
        for f in range(2):
            L.init_eq_dist_wall(f)
    B.step(6, graph=True)
    for L in alone:
        L.step(6)
    _same(_pops(batch), _pops(alone), "after walls on member 2")
    # Shan-Chen off on every member (no psi pass), then on again
    for on in (False, True):
        for L in batch + alone:
            L.set_params([1.1, 1.1], [[0.0, 3.9], [3.9, 0.0]], None, None, None, "BDist2",
                         "Exponential", on)
        B.step(4, graph=True)
        for L in alone:
            L.step(4)
        _same(_pops(batch), _pops(alone), "Shan-Chen %s" % on)
