"""Graph replay of single-lattice steps (dlbc_gpu_step_graph).

The captured steps (32 per graph) are replayed several times with eager work between
the replays (field downloads, eager steps, odd step counts that flip the
buffer parity, parameter, schedule, launch and block-order changes that must
trigger a new capture), and every stage is compared bit for bit with the same
sequence stepped eagerly (dlbc_gpu_step, itself bit-identical to the oracle).
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


def _d2q9(n=64):
    return dict(d=2, q=9, gn=[n, n], components=1, tau=[1.0], fluidInit=["EqDistPerturb"],
                fluidDensities=[[1.0]], fluidPerturb=[1e-3], randomReseed=[True],
                randomSeed=[42])


def _spinodal(n, form="BDist2"):
    return dict(d=3, q=19, gn=list(n), components=2, tau=[1.0, 1.0], eqDistForm=form,
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistPerturb"] * 2, fluidDensities=[[0.7], [0.7]],
                fluidPerturb=[0.01, 0.01], randomReseed=[True, True], randomSeed=[42, 43])


def _walls(n, form="BDist2"):
    rng = np.random.default_rng(7)
    mask = (rng.random(n) < 0.12).astype(np.int32)
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    return dict(_spinodal(n, form), tau=[0.9, 1.1], gwc=[0.2, -0.1],
                globalAcc=[1e-5, 0.0, 0.0], fluidDensities=[[0.7], [0.3]],
                fluidPerturb=[0.05, 0.05], maskArray=mask)


def _d1q5_plates():
    return dict(d=1, q=5, gn=[64], components=2, tau=[1.0, 0.8], eqDistForm="ThirdOrder",
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistPerturbFrac"] * 2, fluidDensities=[[0.7], [0.5]],
                fluidPerturb=[0.2, 0.2], randomReseed=[True, True], randomSeed=[5, 6],
                maskInit="Plates", maskAxis="X", wallOffset=3, gwc=[0.1, -0.2],
                globalAcc=[1e-4])


def _three_components(n=(7, 6, 5)):
    rng = np.random.default_rng(33)
    mask = (rng.random(n) < 0.1).astype(np.int32)
    return dict(d=3, q=19, gn=list(n), components=3, tau=[0.9, 1.1, 1.2],
                eqDistForm="SecondOrder", gcc=[[0.5, 1.2, 0.8], [1.2, 0.0, 1.5],
                                               [0.8, 1.5, 0.3]],
                gwc=[0.1, -0.1, 0.05], enableShanChen=True, psiForm="Exponential",
                globalAcc=[1e-5, 0.0, 0.0], fluidInit=["EqDistPerturb"] * 3,
                fluidDensities=[[0.3], [0.5], [0.4]], fluidPerturb=[0.02] * 3,
                randomReseed=[True] * 3, randomSeed=[21, 22, 23], maskArray=mask)


# (id, parameters, storage, schedule)
CASES = [
    ("d1q5-plates", _d1q5_plates(), "f64", None),
    ("d3q19-3comp-twopass", _three_components(), "f64", None),
    ("d2q9-bgk", _d2q9(), "f64", None),
    ("d3q19-split", _spinodal((16, 20, 24)), "f64", None),
    ("d3q19-split-zrows", _spinodal((6, 7, 256)), "f64", None),
    ("d3q19-fused-walls", _walls((16, 16, 32)), "f64", "fused"),
    ("d3q19-twopass-walls", _walls((12, 10, 14), "SecondOrder"), "f64", "two_pass"),
    ("f32-split-walls", _walls((16, 20, 24)), "f32_shift", None),
    ("f16-fused", _spinodal((20, 16, 40)), "f16_shift", "fused"),
]


def _same(a, b, what):
    for f in range(a.nc):
        pa, pb = a.population(f), b.population(f)
        assert np.array_equal(pa.view(np.uint64), pb.view(np.uint64)), (what, f)


@pytest.mark.parametrize("name,p,storage,schedule", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_equal_eager_steps(D, name, p, storage, schedule):
    a = D.Lattice.from_params(p, storage=storage, schedule=schedule)
    b = D.Lattice.from_params(p, storage=storage, schedule=schedule)
    # 32 steps per graph: counts above, below and between multiples of it;
    # odd counts flip the parity
    for k, n in enumerate([40, 71, 2, 33, 64, 5]):
        a.step(n, graph=True)
        b.step(n)
        _same(a, b, "replay %d (%d steps)" % (k, n))
        for L in (a, b):  # eager work between the replays
            L.density(0)
    for L in (a, b):  # new parameters: the graphs are captured again
        if L.nc == 1:
            L.set_params([0.8])
        else:
            nc = L.nc
            gcc = [[0.0 if i == j else 3.0 + 0.5 * (i + j) for j in range(nc)]
                   for i in range(nc)]
            L.set_params([1.2 - 0.1 * i for i in range(nc)], gcc,
                         [0.1 * (-1) ** i for i in range(nc)] if p.get("gwc") else None, None,
                         None, "BDist2" if L.d > 1 else "ThirdOrder", "Exponential", True)
    a.step(66, graph=True)
    b.step(66)
    _same(a, b, "after set_params")
    if storage == "f64" and a.nc == 2:
        for L in (a, b):  # the launch shape and block order are baked in too
            L.set_launch(96)
            L.set_block_order(3)
        a.step(34, graph=True)
        b.step(34)
        _same(a, b, "after set_launch / set_block_order")
    a.step(3)  # eager steps on the graph's lattice, then replays again
    b.step(3)
    a.step(40, graph=True)
    b.step(40)
    _same(a, b, "eager then replay")


def test_graph_schedule_switch(D):
    """The same lattice replayed under one schedule, switched to another,
    replayed again: the stale graphs are not reused."""
    p = _walls((16, 16, 32))  # FP64 fused kernel: n_y % 8 == 0, n_z % 32 == 0
    a = D.Lattice.from_params(p)
    b = D.Lattice.from_params(p)
    for s in ["split", "fused", "two_pass", "split"]:
        a.set_schedule(s)
        b.set_schedule(s)
        a.step(34, graph=True)
        b.step(34)
        _same(a, b, s)


def test_graph_config1_d2q9_1000_steps_bitexact(D, oracle_mod):
    """D2Q9 single-component BGK 64x64, 1000 steps replayed from the graph,
    against the oracle (populations, densities and velocities bitwise)."""
    p = _d2q9()
    g = D.Lattice.from_params(p)
    o = oracle_mod.Oracle.from_params(p)
    g.step(1000, graph=True)
    o.run(1000)
    assert np.array_equal(g.population(0).view(np.uint64), o.population(0).view(np.uint64))
    assert np.array_equal(g.density(0).view(np.uint64), o.density(0).view(np.uint64))
    assert np.array_equal(g.velocity(0).view(np.uint64), o.velocity(0).view(np.uint64))


def test_graph_after_mask_change(D):
    """A new mask (solid densities recomputed before the replay) under the
    same graphs."""
    p = _walls((16, 20, 24))
    a = D.Lattice.from_params(p)
    b = D.Lattice.from_params(p)
    a.step(34, graph=True)
    b.step(34)
    for L in (a, b):
        L.init_mask("Walls", dim_has_wall=[[False, False], [True, True], [False, False]])
    a.step(36, graph=True)
    b.step(36)
    _same(a, b, "after init_mask")


@pytest.mark.parametrize("storage", ["f64", "f32_shift"])
def test_graph_after_restore(D, storage):
    """A restart in the middle of a run (the populations of another lattice
    uploaded, as from a checkpoint) with the graphs already captured: the
    replays read the new state, as eager steps do."""
    p = _walls((16, 16, 32))
    a = D.Lattice.from_params(p, storage=storage)
    b = D.Lattice.from_params(p, storage=storage)
    c = D.Lattice.from_params(p, storage=storage)
    a.step(40, graph=True)  # a holds captured graphs for both buffer parities
    a.step(33, graph=True)
    b.step(73)
    c.step(19)  # the state to restore
    state = [c.population(f) for f in range(2)]
    for L in (a, b):
        for f in range(2):
            L.upload_population(f, state[f])
    a.step(70, graph=True)
    b.step(70)
    _same(a, b, "after restore")


def test_mask_change_mid_run_matches_oracle(D, oracle_mod):
    """Walls added after 17 steps (an odd count: the two population buffers
    hold different states then): the new solid sites keep the populations
    they had, as the reference's advection copies a solid site's populations
    every step (advection.d:84-86).  Eager steps and graph replays both equal
    the oracle run through the same sequence, bit for bit."""
    p = _walls((16, 20, 24))
    o = oracle_mod.Oracle.from_params(p)
    a = D.Lattice.from_params(p)
    b = D.Lattice.from_params(p)
    a.step(17)
    b.step(17, graph=True)
    o.run(17)
    walls = [[False, False], [True, True], [False, False]]
    for L in (a, b):
        L.init_mask("Walls", dim_has_wall=walls)
    o.init_mask("Walls", 0, 0, walls, 1)
    o.exchange_halo()
    a.step(40)
    b.step(40, graph=True)
    o.run(40)
    for L, what in ((a, "eager"), (b, "graph")):
        for f in range(2):
            assert np.array_equal(L.population(f).view(np.uint64),
                                  o.population(f).view(np.uint64)), (what, f)


@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
def test_mask_change_mid_run_shifted_keeps_solid_state(D, storage):
    """Shifted storage: walls added after 17 steps (the two buffers then hold
    different states) keep, at every new solid site, the populations the site
    had when the walls came, in both buffers (an even and an odd number of
    steps later), and graph replays equal eager steps bit for bit."""
    p = _walls((16, 20, 24))
    a = D.Lattice.from_params(p, storage=storage)
    b = D.Lattice.from_params(p, storage=storage)
    a.step(17)
    b.step(17, graph=True)
    before = [a.population(f) for f in range(2)]
    walls = [[False, False], [True, True], [False, False]]
    old = a.mask()
    for L in (a, b):
        L.init_mask("Walls", dim_has_wall=walls)
    new = (a.mask() != 0) & (old == 0)
    assert new.any()
    for steps in (40, 1):
        a.step(steps)
        b.step(steps, graph=True)
        for f in range(2):
            got = a.population(f)[new]
            assert np.array_equal(got.view(np.uint64), before[f][new].view(np.uint64)), (steps, f)
        _same(a, b, "shifted after init_mask, %d steps" % steps)
