"""Seeded random sweep of the default D3Q19 two-component schedule (the split
passes) against the oracle, bit for bit: lattice shapes from 2 to 128 sites per
axis (z-rows that hold whole blocks, and ones that do not), solid fractions from
0 to 0.6 with and without solid slabs of whole cache lines, the three
equilibria, both psi forms, wall couplings and body forces of either sign."""
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


def _case(seed):
    rng = np.random.default_rng(seed)
    n = [int(rng.integers(2, 24)), int(rng.integers(2, 24)),
         int(rng.choice([rng.integers(2, 70), 32, 64, 128]))]
    mask = (rng.random(n) < float(rng.choice([0.0, 0.1, 0.3, 0.6]))).astype(np.int32)
    if n[2] > 20 and rng.random() < 0.5:
        z0 = int(rng.integers(0, n[2] - 16))
        mask[:, :, z0:z0 + 16] = 1
    g = float(rng.uniform(2.0, 4.0))
    p = dict(d=3, q=19, gn=n, components=2,
             tau=[float(rng.uniform(0.8, 1.3)), float(rng.uniform(0.8, 1.3))],
             eqDistForm=str(rng.choice(["SecondOrder", "ThirdOrder", "BDist2"])),
             gcc=[[0.0, g], [g, 0.0]],
             gwc=[float(rng.uniform(-0.2, 0.2)), float(rng.uniform(-0.2, 0.2))],
             enableShanChen=True, psiForm=str(rng.choice(["Exponential", "Linear"])),
             globalAcc=[1e-5, 0.0, 0.0], fluidInit=["EqDistPerturb"] * 2,
             fluidDensities=[[0.7], [0.3]], fluidPerturb=[0.05, 0.05],
             randomReseed=[True, True],
             randomSeed=[int(rng.integers(1, 1000)), int(rng.integers(1, 1000))],
             maskArray=mask)
    if rng.random() < 0.5:
        p["bodyForce"] = [[0.0, 0.0, 0.0], [float(rng.choice([-1e-5, 1e-5])), 0.0, 0.0]]
    return p, int(rng.integers(5, 20))


@pytest.mark.parametrize("seed", range(100, 116))
def test_random_lattices_bitexact(D, oracle_mod, seed):
    p, steps = _case(seed)
    g = D.Lattice.from_params(p)
    o = oracle_mod.Oracle.from_params(p)
    g.step(steps)
    o.run(steps)
    for f in range(2):
        assert np.array_equal(g.population(f).view(np.uint64),
                              o.population(f).view(np.uint64)), (p["gn"], p["eqDistForm"], f)


def _shifted_case(seed):
    """A random lattice for the shifted storages (n >= 4 on every axis)."""
    p, steps = _case(seed)
    rng = np.random.default_rng(seed + 7)
    n = [max(4, v) for v in p["gn"]]
    mask = (rng.random(n) < float(rng.choice([0.0, 0.1, 0.3]))).astype(np.int32)
    return dict(p, gn=n, maskArray=mask, rho0=[0.5, 0.5]), steps


@pytest.mark.parametrize("seed", range(200, 210))
@pytest.mark.parametrize("storage", ["f32_shift", "f16_shift"])
def test_random_lattices_shifted_split_equals_tiled(D, storage, seed):
    """The same random lattices in FP32- and FP16-shifted storage: the split
    passes (the default) and the tiled kernel give the same bits."""
    p, steps = _shifted_case(seed)
    a = D.Lattice.from_params(p, storage=storage)
    b = D.Lattice.from_params(p, storage=storage, schedule="fused")
    assert a.schedule == "split"
    a.step(steps)
    b.step(steps)
    for f in range(2):
        assert np.array_equal(a.population(f).view(np.uint64),
                              b.population(f).view(np.uint64)), (p["gn"], p["eqDistForm"], f)
