"""HIP path against the oracle, bit for bit (FP64), through the C ABI.

Every test builds the same seeded input for the GPU (dlbc_amd) and for the
oracle, advances both and compares populations with a bitwise equality: the
kernels evaluate the reference's expressions in the reference's order
without contraction, and exp is correctly rounded on both sides.
"""
import os

import numpy as np
import pytest

from cases import assert_close, case, expand, golden, tolerance

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import dlbc_amd
    import torch  # noqa: F401  (device visibility check)
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    return dlbc_amd


def _pair(D, oracle_mod, p, schedule=None):
    g = D.Lattice.from_params(p, schedule=schedule)
    o = oracle_mod.Oracle.from_params(p)
    return g, o


def _same_state(g, o, what):
    for f in range(o.nc):
        assert_close(g.population(f), o.population(f), 0.0, "%s population %d" % (what, f))


def test_psi_exp_bitexact(D, oracle_mod):
    """psi = 1 - exp(-rho) on the device equals the oracle's (both correctly
    rounded) on 2^20 random densities plus edge values."""
    L = D.Lattice(3, 19, 1, [4, 4, 4])
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0.0, 3.0, 1 << 19), rng.uniform(-1.0, 40.0, 1 << 18),
                        -rng.uniform(-745.0, 708.0, 1 << 18),
                        np.array([0.0, -0.0, 1e-300, 5e-324, 708.39, 708.5, 745.1, 745.2,
                                  746.0, -709.7, -710.0, np.inf, -np.inf])])
    y = L.eval_psi(1, x)
    ref = np.array([oracle_mod.psi_value(1, v) for v in x])
    bad = np.nonzero(y.view(np.uint64) != ref.view(np.uint64))[0]
    assert bad.size == 0, [(x[i], y[i], ref[i]) for i in bad[:5]]
    assert np.array_equal(L.eval_psi(0, x[:100]), x[:100])


@pytest.mark.parametrize("dq", [(3, 19), (2, 9), (1, 3)])
def test_advect_bounce_back_bitexact(D, oracle_mod, dq):
    """Streaming with bounce-back alone (advectFields) on random populations
    and a random mask: identical to the oracle bit for bit, and equal to a
    numpy restatement (roll, then bounce at solid upstream sites)."""
    d, q = dq
    n = {3: [9, 7, 12], 2: [17, 10], 1: [33]}[d]
    rng = np.random.default_rng(d)
    mask = (rng.random(n) < 0.2).astype(np.int32)
    pops = [rng.uniform(0.0, 1.0, tuple(n) + (q,)) for _ in range(2)]
    g = D.Lattice(d, q, 2, n)
    g.set_params([1.0, 1.0])
    g.upload_mask(mask)
    for f in range(2):
        g.upload_population(f, pops[f])
    o = oracle_mod.Oracle(d, q, 2, n)
    o.set_params([1.0, 1.0])
    o.set_mask(mask)
    for f in range(2):
        o.set_population(f, pops[f])
    o.exchange_halo()
    o.advect()
    g.advect()
    _same_state(g, o, "advect")
    vel, bounce, _, _ = oracle_mod.conn(d, q)
    for f in range(2):
        exp = np.empty_like(pops[f])
        for i in range(q):
            shifted = np.roll(pops[f][..., i], shift=tuple(vel[i]), axis=tuple(range(d)))
            solid_up = np.roll(mask, shift=tuple(vel[i]), axis=tuple(range(d))) == 1
            exp[..., i] = np.where(solid_up, pops[f][..., bounce[i]], shifted)
        exp[mask == 1] = pops[f][mask == 1]
        assert_close(g.population(f), exp, 0.0, "numpy streaming f%d" % f)


STEP_CASES = [
    ("bdist2_3d", {}, 10),
    ("bdist2_2d", {}, 10),
    ("bdist2_sc_3d", {"gn": [4, 4, 64]}, 100),
    ("bdist2_sc_2d", {"gn": [64, 4]}, 100),
    ("merging_droplets_2d", {}, 300),
    ("poiseuille_2d", {}, 1000),
    ("laplace_quasi_1d_d2q9", {}, 1000),
    ("laplace_quasi_1d_d1q3", {}, 1000),
    ("laplace_droplet_3d", {"gn": [24, 24, 24]}, 20),
    ("checkpoint_3d", {}, 5),
]


@pytest.mark.parametrize("name,over,steps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_reference_cases_bitexact(D, oracle_mod, name, over, steps):
    """The reference's runnable-test configurations: GPU == oracle bitwise."""
    p = case(name, **over)
    g, o = _pair(D, oracle_mod, p)
    g.step(steps)
    o.run(steps)
    _same_state(g, o, name)
    for f in range(p["components"]):
        assert_close(g.density(f), o.density(f), 0.0, "density")
        assert_close(g.velocity(f), o.velocity(f), 0.0, "velocity")
    if p["components"] == 2:
        assert_close(g.colour(), o.colour(), 0.0, "colour")


def _config1(n=64):
    """D2Q9 single-component BGK, 64x64 periodic, tau 1, rho 1 +- 1e-3 seeded."""
    return dict(d=2, q=9, gn=[n, n], components=1, tau=[1.0], fluidInit=["EqDistPerturb"],
                fluidDensities=[[1.0]], fluidPerturb=[1e-3], randomReseed=[True],
                randomSeed=[42])


def _config2(n, form="SecondOrder"):
    """D3Q19 two-component Shan-Chen spinodal decomposition, periodic,
    tau 1, g 3.6, psi = 1 - exp(-rho), rho 0.7 +- 0.01 seeded per fluid."""
    return dict(d=3, q=19, gn=[n, n, n], components=2, tau=[1.0, 1.0], eqDistForm=form,
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistPerturb"] * 2, fluidDensities=[[0.7], [0.7]],
                fluidPerturb=[0.01, 0.01], randomReseed=[True, True], randomSeed=[42, 43])


def test_config1_d2q9_1000_steps_bitexact(D, oracle_mod):
    p = _config1()
    g, o = _pair(D, oracle_mod, p)
    m0 = g.total_mass(0)
    g.step(1000)
    o.run(1000)
    _same_state(g, o, "config1")
    assert abs(g.total_mass(0) - m0) <= 1e-12 * m0


@pytest.mark.parametrize("form", ["SecondOrder", "BDist2", "ThirdOrder"])
def test_config2_spinodal_bitexact(D, oracle_mod, form):
    """Spinodal decomposition amplifies any 1-ulp difference; 200 steps at
    24^3 stay bit-identical."""
    p = _config2(24, form)
    g, o = _pair(D, oracle_mod, p)
    g.step(200)
    o.run(200)
    _same_state(g, o, "config2 " + form)


@pytest.mark.parametrize("form", ["SecondOrder", "BDist2"])
def test_config2_full_size_1000_steps_bitexact(D, oracle_mod, form):
    """The 128^3 spinodal workload at its full size and length, 1000 steps,
    with the reference's default SecondOrder equilibrium and with BDist2:
    bit-identical to the oracle.  (On MI355X the GPU takes 0.4-0.6 s
    including initialisation, the oracle 67-78 s on 16 cores.)"""
    p = _config2(128, form)
    g, o = _pair(D, oracle_mod, p)
    g.step(1000)
    o.run(1000)
    _same_state(g, o, "config2 128^3 " + form)


def _d1q5_case(form, walls):
    """D1Q5 (c_s^2 = 1, links of length 2, halo 4 in the reference):
    lb/laplace-quasi-1d-d1q5.in with a seeded perturbation, optionally with
    plates, a wall coupling and an acceleration."""
    p = dict(d=1, q=5, gn=[64], components=2, tau=[1.0, 0.8], eqDistForm=form,
             gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
             fluidInit=["EqDistPerturbFrac"] * 2, fluidDensities=[[0.7], [0.5]],
             fluidPerturb=[0.2, 0.2], randomReseed=[True, True], randomSeed=[5, 6])
    if walls:
        p.update(maskInit="Plates", maskAxis="X", wallOffset=3, gwc=[0.1, -0.2],
                 globalAcc=[1e-4])
    return p


@pytest.mark.parametrize("form", ["SecondOrder", "ThirdOrder", "BDist2"])
@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "plates"])
def test_d1q5_bitexact(D, oracle_mod, form, walls):
    """D1Q5 through the general two-pass kernels against the oracle, bit for
    bit, after 300 steps.  (The reference's own D1Q5 test is disabled and
    holds no data, so the oracle restatement is the pin.)"""
    p = _d1q5_case(form, walls)
    g, o = _pair(D, oracle_mod, p)
    g.step(300)
    o.run(300)
    _same_state(g, o, "D1Q5 %s walls=%s" % (form, walls))


def test_d1q5_laplace_case_bitexact(D, oracle_mod):
    """lb/laplace-quasi-1d-d1q5.in as written (ThirdOrder, a centred
    lamella of 0.7/0.5, 1000 steps)."""
    p = dict(d=1, q=5, gn=[64], components=2, tau=[1.0, 1.0], eqDistForm="ThirdOrder",
             gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
             fluidInit=["EqDistSphereFrac"] * 2, fluidDensities=[[0.7, 0.5], [0.5, 0.7]],
             initRadius=0.25, interfaceThickness=5.0)
    g, o = _pair(D, oracle_mod, p)
    g.step(1000)
    o.run(1000)
    _same_state(g, o, "D1Q5 laplace")


def test_quasi_1d_across_dimensions(D):
    """lb/laplace-quasi-1d/compare-lower-dimensions.py: the same x-lamellae
    (ThirdOrder, Shan-Chen, 1000 steps) on D3Q19 64x4x4, D2Q9 64x4 and D1Q3 64
    give density profiles along x equal to 1e-14 (the reference's bound)."""
    base = dict(components=2, tau=[1.0, 1.0], eqDistForm="ThirdOrder",
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistLamellaeFrac"] * 2,
                fluidDensities=[[0.5, 0.7, 0.5], [0.7, 0.5, 0.7]],
                lamellaeWidths=[0.25, 0.50, 0.25], interfaceThickness=5.0)
    prof = {}
    for d, q, gn in ((3, 19, [64, 4, 4]), (2, 9, [64, 4]), (1, 3, [64])):
        L = D.Lattice.from_params(dict(base, d=d, q=q, gn=gn))
        L.step(1000)
        prof[d] = [L.density(f).reshape(64, -1) for f in range(2)]
    for f in range(2):
        for d in (3, 2):
            # uniform across y (and z), and equal to the D1Q3 profile
            assert np.abs(prof[d][f] - prof[d][f][:, :1]).max() <= 1e-14
            assert np.abs(prof[d][f][:, 0] - prof[1][f][:, 0]).max() <= 1e-14


def test_walls_wetting_bodyforce_bitexact(D, oracle_mod):
    """Walls with a wall-fluid coupling (gwc), random solid sites and a body
    force on both components, which no reference fixture exercises: GPU ==
    oracle."""
    rng = np.random.default_rng(7)
    n = [12, 10, 16]
    mask = (rng.random(n) < 0.1).astype(np.int32)
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    p = dict(d=3, q=19, gn=n, components=2, tau=[0.8, 1.2], gcc=[[0.0, 3.6], [3.6, 0.0]],
             gwc=[0.3, -0.2], enableShanChen=True, psiForm="Exponential",
             globalAcc=[1e-5, 0.0, 0.0], fluidInit=["EqDistPerturb"] * 2,
             fluidDensities=[[0.7], [0.3]], fluidPerturb=[0.05, 0.05],
             randomReseed=[True, True], randomSeed=[5, 6], maskArray=mask,
             bodyForce=[[0.0, 3e-6, 0.0], [-2e-6, 0.0, 1e-6]])
    for form in ("SecondOrder", "BDist2"):
        p["eqDistForm"] = form
        g, o = _pair(D, oracle_mod, p)
        g.step(50)
        o.run(50)
        _same_state(g, o, "walls " + form)


def test_body_force_equals_global_acceleration(D):
    """On a uniform state a body force F_f = rho_f g and a global
    acceleration g give the same velocity shift (up to the rounding of
    (rho g)/rho), and a zero body force is bitwise a no-op."""
    acc = [2e-5, 0.0, -1e-5]
    p = dict(d=3, q=19, gn=[8, 8, 8], components=2, tau=[1.0, 0.8], eqDistForm="SecondOrder",
             fluidInit=["EqDist"] * 2, fluidDensities=[[0.7], [0.3]])
    a = D.Lattice.from_params(dict(p, globalAcc=acc))
    b = D.Lattice.from_params(dict(p, bodyForce=[[0.7 * v for v in acc], [0.3 * v for v in acc]]))
    c = D.Lattice.from_params(dict(p, bodyForce=[[0.0] * 3, [0.0] * 3]))
    z = D.Lattice.from_params(p)
    for L in (a, b, c, z):
        L.step(1)
    for f in range(2):
        assert_close(b.population(f), a.population(f), 1e-17, "body force vs globalAcc")
        assert not np.array_equal(a.population(f), z.population(f))
        assert_close(c.population(f), z.population(f), 0.0, "zero body force")


def test_mass_conservation_128(D):
    """Config 2 at its full size, 1000 steps: total mass per component is
    conserved to round-off (compensated device sum)."""
    p = _config2(128)
    g = D.Lattice.from_params(p)
    m0 = [g.total_mass(f, parts=True) for f in range(2)]
    g.step(1000)
    for f in range(2):
        m1 = g.total_mass(f, parts=True)
        assert abs((m1[0] - m0[f][0]) + (m1[1] - m0[f][1])) <= 1e-11 * m0[f][0]
    assert np.all(np.isfinite(g.density(0)))


def test_total_mass_matches_exact_sum(D, oracle_mod):
    """The device's double-double mass equals math.fsum of the per-site
    densities (the exact sum) rounded to double."""
    import math
    p = _config2(20)
    g, o = _pair(D, oracle_mod, p)
    g.step(10)
    o.run(10)
    for f in range(2):
        exact = math.fsum(o.density(f).ravel().tolist())
        assert g.total_mass(f) == exact
        mom = g.total_momentum(f)
        assert np.allclose(mom, o.total_momentum(f), rtol=0, atol=1e-13)


def test_laplace_droplet_3d_against_reference_t10000(D):
    """lb/laplace-droplet-3d on the GPU for the full 10000 steps against the
    reference's own t100 and t10000 data (central planes and exact totals).
    t100 holds the JSON's 1e-15.  At t10000 the deviation is 6.2e-15: D's exp
    evaluates in x87 extended precision and differs from the correctly
    rounded exp used here in rare double-rounding cases, and 10^4 steps
    accumulate those 1-ulp differences (observed 6.2e-15 on the densities,
    1.12e-14 on their difference); the bound used there is 2.5e-14."""
    import math
    gd, meta = golden("laplace_droplet_3d")
    g = D.Lattice.from_params(case("laplace_droplet_3d"))
    t = 0
    for step in (100, 10000):
        g.step(step - t)
        t = step
        for fname, arr in (("density_red", g.density(0)), ("density_blue", g.density(1)),
                           ("colour_red_blue", g.colour())):
            key = "%s_t%d" % (fname, step)
            info = meta["fields"][key]
            px, py, pz = info["planes"]
            tol = tolerance(meta, step) if step < 10000 else 2.5e-14
            assert_close(arr[px], gd[key + "_x"], tol, key)
            assert_close(arr[:, py], gd[key + "_y"], tol, key)
            assert_close(arr[:, :, pz], gd[key + "_z"], tol, key)
            assert abs(math.fsum(arr.ravel().tolist()) - info["fsum"]) <= tol * arr.size


def test_merging_droplets_3d_against_reference_t1000(D):
    import math
    gd, meta = golden("merging_droplets_3d")
    g = D.Lattice.from_params(case("merging_droplets_3d"))
    g.step(1000)
    for fname, arr in (("density_red", g.density(0)), ("density_blue", g.density(1)),
                       ("colour_red_blue", g.colour())):
        key = "%s_t1000" % fname
        info = meta["fields"][key]
        px, py, pz = info["planes"]
        assert_close(arr[px], gd[key + "_x"], 1e-14, key)
        assert_close(arr[:, py], gd[key + "_y"], 1e-14, key)
        assert_close(arr[:, :, pz], gd[key + "_z"], 1e-14, key)
        assert abs(math.fsum(arr.ravel().tolist()) - info["fsum"]) <= 1e-14 * arr.size


def test_bdist2_sc_3d_full_size_against_reference(D):
    """lb/bdist2-sc-3d at its real 64^3 size for 100 steps, against the
    reference's data and against LB3D's run of the same case."""
    gd, meta = golden("bdist2_sc_3d")
    lb3d, meta3 = golden("bdist2_sc_3d_lb3d")
    g = D.Lattice.from_params(case("bdist2_sc_3d"))
    g.step(100)
    for f, c in enumerate(("red", "blue")):
        key = "population_%s_t100" % c
        pop = g.population(f)
        assert_close(pop, expand(gd[key], meta["shapes"][key]), 1e-15, key)
        assert_close(pop, expand(lb3d[key], meta3["shapes"][key]), 1e-15, "LB3D " + key)


def test_errors_are_returned_not_fatal(D):
    """Invalid calls return error codes with a message (the library never
    exits): stepping before upload, a bad component, a bad mask value."""
    L = D.Lattice(3, 19, 2, [8, 8, 8])
    with pytest.raises(D.DlbcError, match="call out of order"):
        L.step(1)
    L.set_params([1.0, 1.0])
    with pytest.raises(D.DlbcError):
        L.upload_mask(np.full((8, 8, 8), 7, np.int32))
    with pytest.raises(D.DlbcError):
        L.population(5)
    with pytest.raises(D.DlbcError):
        D.Lattice(3, 19, 2, [1, 8, 8])


def _walls_case(n, form, seed=3):
    rng = np.random.default_rng(seed)
    mask = (rng.random(n) < 0.12).astype(np.int32)
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    return dict(d=3, q=19, gn=list(n), components=2, tau=[0.9, 1.1], eqDistForm=form,
                gcc=[[0.0, 3.6], [3.6, 0.0]], gwc=[0.2, -0.1], enableShanChen=True,
                psiForm="Exponential", globalAcc=[1e-5, 0.0, 0.0],
                fluidInit=["EqDistPerturb"] * 2, fluidDensities=[[0.7], [0.3]],
                fluidPerturb=[0.05, 0.05], randomReseed=[True, True], randomSeed=[11, 12],
                maskArray=mask)


@pytest.mark.parametrize("form", ["SecondOrder", "BDist2", "ThirdOrder"])
@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_fused_schedule_matches_oracle(D, oracle_mod, form, walls):
    """The single-pass D3Q19 kernel (LDS psi ring, x-marching tiles) against
    the oracle, bit for bit, periodic and with walls + random solids + gwc."""
    n = [10, 16, 64]
    p = _walls_case(n, form) if walls else dict(_config2(8, form), gn=n)
    g = D.Lattice.from_params(p, schedule="fused")
    assert g.schedule == "fused"
    o = oracle_mod.Oracle.from_params(p)
    g.step(40)
    o.run(40)
    _same_state(g, o, "fused %s walls=%s" % (form, walls))


def test_fused_equals_two_pass_with_body_force(D):
    """All three schedules give the same bits, body force included, over
    several chunks of x and tiles of (y, z)."""
    p = dict(_walls_case([24, 32, 64], "SecondOrder"),
             bodyForce=[[0.0, 0.0, 0.0], [1e-6, 0.0, 0.0]])
    a = D.Lattice.from_params(p, schedule="fused")
    b = D.Lattice.from_params(p, schedule="two_pass")
    c = D.Lattice.from_params(p, schedule="split")
    for L in (a, b, c):
        L.step(60)
    for f in range(2):
        assert_close(a.population(f), b.population(f), 0.0, "fused vs two-pass f%d" % f)
        assert_close(c.population(f), b.population(f), 0.0, "split vs two-pass f%d" % f)


@pytest.mark.parametrize("n", [[9, 13, 21], [5, 4, 128]], ids=["unaligned", "z-rows"])
@pytest.mark.parametrize("form", ["SecondOrder", "BDist2", "ThirdOrder"])
@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_split_schedule_matches_oracle(D, oracle_mod, form, walls, n):
    """The lane-per-component two-pass D3Q19 kernels against the oracle, bit
    for bit: on a lattice no tile divides (any shape is allowed), and on one
    whose z-rows hold whole blocks, where the collide pass stages the psi rows
    in LDS."""
    p = _walls_case(n, form) if walls else dict(_config2(8, form), gn=n)
    g = D.Lattice.from_params(p, schedule="split")
    assert g.schedule == "split"
    o = oracle_mod.Oracle.from_params(p)
    g.step(40)
    o.run(40)
    _same_state(g, o, "split %s walls=%s" % (form, walls))


def _solid_blocks_case(n, form="BDist2"):
    """Walls case plus solid boxes that cover whole cache lines along z (16
    FP64 sites a line) and lines of which only part is solid."""
    p = _walls_case(n, form)
    mask = p["maskArray"].copy()
    mask[1:4, 2:6, 16:48] = 1   # two whole 16-site lines per row
    mask[5:7, :, 70:75] = 1     # parts of lines
    return dict(p, maskArray=mask)


@pytest.mark.parametrize("n", [[8, 9, 128], [6, 7, 100]], ids=["z-rows", "unaligned"])
def test_split_whole_solid_lines_bitexact(D, oracle_mod, n):
    """The split collide pass leaves cache lines of solid sites only alone and
    copies the populations of the solid sites of mixed lines: after 30 steps
    every population, solid sites' included, equals the oracle's bit for bit."""
    p = _solid_blocks_case(n)
    g = D.Lattice.from_params(p, schedule="split")
    o = oracle_mod.Oracle.from_params(p)
    g.step(30)
    o.run(30)
    _same_state(g, o, "split, solid boxes")


# (z-rows per workgroup, populations held in registers, launch target):
# one row pulling both neighbouring rows, the same holding its populations,
# two rows, and the two row counts with one chunk of x per workgroup.
MARCH_SHAPES = [(1, 0, 0), (1, 1, 0), (2, 0, 0), (1, 0, 1), (2, 0, 1)]


@pytest.mark.parametrize("shape", MARCH_SHAPES, ids=["rows1", "rows1-hold", "rows2",
                                                      "rows1-long", "rows2-long"])
@pytest.mark.parametrize("form", ["SecondOrder", "BDist2", "ThirdOrder"])
@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_march_schedule_matches_oracle(D, oracle_mod, form, walls, shape):
    """The single-pass kernel over whole z-rows (LDS psi ring, x-marching)
    against the oracle, bit for bit: chunks of one plane (the default target
    on a small lattice) and one chunk per row tile, periodic and with walls,
    random solids and gwc."""
    n = [9, 6, 32]
    p = _walls_case(n, form) if walls else dict(_config2(8, form), gn=n)
    g = D.Lattice.from_params(p, schedule="march")
    g.set_march(shape[0], shape[1])
    if shape[2]:
        g.set_launch(shape[2])
    assert g.schedule == "march"
    o = oracle_mod.Oracle.from_params(p)
    g.step(40)
    o.run(40)
    _same_state(g, o, "march %s %s walls=%s" % (shape, form, walls))


@pytest.mark.parametrize("n", [[7, 5, 64], [6, 4, 256], [3, 3, 512], [5, 2, 32]],
                         ids=["nz64-odd-ny", "nz256", "nz512", "ny2"])
def test_march_equals_split_with_body_force(D, n):
    """The march schedule gives the split pair's bits on the other row
    lengths it serves (64, 256 and 512 sites, n_y odd and n_y = 2), with
    walls, random solids, gwc and the body force, in every launch shape."""
    p = dict(_walls_case(n, "BDist2"), bodyForce=[[0.0, 0.0, 0.0], [1e-6, 0.0, 0.0]])
    ref = D.Lattice.from_params(p, schedule="split")
    ref.step(25)
    want = [ref.population(f) for f in range(2)]
    for rows, hold in [(0, -1), (1, 0), (1, 1), (2, 0)]:
        g = D.Lattice.from_params(p, schedule="march")
        if rows == 2 and (n[1] % 2 or n[2] > 256):
            with pytest.raises(D.DlbcError):
                g.set_march(rows, hold)
            continue
        g.set_march(rows, hold)
        g.step(25)
        for f in range(2):
            assert_close(g.population(f), want[f], 0.0, "march %s %s f%d" % (n, (rows, hold), f))


def test_march_refuses_unserved_lattices(D):
    """Row lengths other than 32, 64, 256 and 512, shifted storage and
    D2Q9 are refused with DLBC_ERR_UNSUPPORTED, and the automatic schedule
    is unaffected."""
    for n, storage in (([8, 8, 48], "f64"), ([8, 8, 128], "f64"), ([8, 8, 64], "f32_shift")):
        L = D.Lattice(3, 19, 2, n, storage=storage)
        with pytest.raises(D.DlbcError):
            L.set_schedule("march")
        assert L.schedule != "march"
    L = D.Lattice(2, 9, 2, [32, 32])
    with pytest.raises(D.DlbcError):
        L.set_schedule("march")


@pytest.mark.parametrize("n", [[10, 24, 40], [9, 10, 256]], ids=["gather", "staged"])
@pytest.mark.parametrize("storage", ["f64", "f32_shift", "f16_shift"])
def test_split_block_order_does_not_change_results(D, storage, n):
    """The XCD block order only permutes which workgroup takes which sites:
    site order, the automatic order, and chunks of 1, 3 and 16 blocks per
    XCD turn give the same populations bit for bit, on lattices whose grid
    is not a whole number of turns (the tail keeps site order), with psi
    gathered per lane and staged in LDS."""
    p = _walls_case(n, "BDist2")
    pops = []
    for order in (0, -1, 1, 3, 16):
        L = D.Lattice.from_params(p, storage=storage, schedule="split")
        L.set_block_order(order)
        assert L.block_order == order or order == -1
        L.step(6)
        pops.append([L.population(f) for f in range(2)])
    for other in pops[1:]:
        for f in range(2):
            assert np.array_equal(other[f].view(np.uint64), pops[0][f].view(np.uint64))


@pytest.mark.parametrize("dq,n,nc", [((3, 19), [7, 6, 5], 3), ((2, 9), [12, 9], 4),
                                     ((1, 3), [31], 3), ((3, 19), [6, 5, 4], 4)])
@pytest.mark.parametrize("form", ["SecondOrder", "BDist2"])
def test_more_than_two_components_bitexact(D, oracle_mod, dq, n, nc, form):
    """Three and four components (the general two-pass schedule), with
    self- and cross-couplings, wall couplings and a random mask, against the
    oracle bit for bit after 25 steps."""
    d, q = dq
    rng = np.random.default_rng(nc * 10 + d)
    mask = (rng.random(n) < 0.1).astype(np.int32)
    g = rng.uniform(0.5, 2.0, (nc, nc))
    gcc = ((g + g.T) / 2.0).tolist()
    p = dict(d=d, q=q, gn=list(n), components=nc, tau=list(rng.uniform(0.8, 1.3, nc)),
             eqDistForm=form, gcc=gcc, gwc=list(rng.uniform(-0.2, 0.2, nc)),
             enableShanChen=True, psiForm="Exponential",
             globalAcc=[1e-5] + [0.0] * (d - 1), fluidInit=["EqDistPerturb"] * nc,
             fluidDensities=[[v] for v in rng.uniform(0.2, 0.8, nc)],
             fluidPerturb=[0.02] * nc, randomReseed=[True] * nc,
             randomSeed=list(range(21, 21 + nc)), maskArray=mask)
    gl, o = _pair(D, oracle_mod, p)
    gl.step(25)
    o.run(25)
    _same_state(gl, o, "C=%d D%dQ%d %s" % (nc, d, q, form))


@pytest.mark.parametrize("n", [[2, 2, 2], [2, 3, 5], [3, 2, 7]])
@pytest.mark.parametrize("schedule", ["two_pass", "split"])
def test_smallest_lattices_bitexact(D, oracle_mod, n, schedule):
    """Lattices two sites wide, where a site's +1 and -1 neighbours are the
    same site, against the oracle bit for bit."""
    p = _walls_case(n, "BDist2")
    p["maskArray"] = np.zeros(n, np.int32)
    p["maskArray"][0, 0, 0] = 1  # one solid site: bounce-back and wall force
    g = D.Lattice.from_params(p, schedule=schedule)
    o = oracle_mod.Oracle.from_params(p)
    g.step(10)
    o.run(10)
    _same_state(g, o, "n=%s %s" % (n, schedule))


def _trace_close(got, row, rel=6e-7):
    t, sigma, radius, p_in, p_out, deform = row
    for name, ref in (("sigma", sigma), ("radius", radius), ("in_pressure", p_in),
                      ("out_pressure", p_out)):
        assert abs(got[name] - ref) <= rel * abs(ref) + 1e-12, (t, name, got[name], ref)
    if np.isnan(deform):
        assert np.isnan(got["deformation"])
    else:  # round-off sized on both sides (a circular droplet)
        assert abs(got["deformation"]) <= 2e-8 and abs(deform) <= 1e-8, (t, got, deform)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_laplace_trace_against_reference(D, dim):
    """The reference's Laplace output (sigma, radius, pressures, deformation;
    every 100 steps to t = 10000, 7 significant digits) reproduced by the
    product's step and its device-side Laplace diagnostics."""
    import os
    from cases import GOLDEN
    g = np.load(os.path.join(GOLDEN, "laplace_trace_%s.npz" % dim))["trace"]
    L = D.Lattice.from_params(case("laplace_droplet_%s" % dim))
    t = 0
    for row in g:
        L.step(int(row[0]) - t)
        t = int(row[0])
        _trace_close(L.laplace(), row)


def test_laplace_equals_oracle(D, oracle_mod):
    """Device Laplace diagnostics == the oracle's restatement on the same
    (bitwise identical) state; only the total-mass summation differs
    (double-double on the device, sequential in the oracle)."""
    p = case("laplace_droplet_3d", gn=[32, 32, 32])
    g, o = _pair(D, oracle_mod, p)
    g.step(50)
    o.run(50)
    a, b = g.laplace(), o.laplace(p["gcc"], p["psiForm"])
    for k in ("sigma", "radius", "in_pressure", "out_pressure"):
        assert abs(a[k] - b[k]) <= 1e-13 * abs(b[k]), (k, a[k], b[k])


def test_split_wide_offsets_equal_fused(D):
    """A lattice whose population index space exceeds 2^29 (19 x 28.7 M
    sites) takes the split kernels' 64-bit addressing path; after three steps
    its fields equal the fused schedule's bit for bit."""
    n = [232, 352, 352]
    p = _walls_case(n, "BDist2")
    a = D.Lattice.from_params(p, schedule="split")
    b = D.Lattice.from_params(p, schedule="fused")
    a.step(3)
    b.step(3)
    for f in range(2):
        assert_close(a.density(f), b.density(f), 0.0, "density f%d" % f)
        assert_close(a.velocity(f), b.velocity(f), 0.0, "velocity f%d" % f)


def test_restart_from_hdf5_checkpoint(D, tmp_path):
    """Checkpoint and restart through the reference's HDF5 layout: the
    checkpoint-3d state at t = 0 written as cp-fluids0/1 and cp-mask files,
    restored into a fresh lattice, advanced 5 steps: the populations equal
    the reference's t = 5 output bit for bit; the output files written then
    read back equal to the device fields."""
    from dlbc_amd import h5io
    if not h5io.available():
        pytest.skip("libdlbc_h5.so not built (no HDF5 library)")
    g, _ = golden("checkpoint_3d")
    p = case("checkpoint_3d")
    sim = "checkpoint-3d"
    src = D.Lattice(3, 19, 2, p["gn"])
    src.set_params(p["tau"], p.get("gcc"), p.get("gwc"), p.get("globalAcc"), None,
                   p.get("eqDistForm", "SecondOrder"), p.get("psiForm", "Linear"),
                   p.get("enableShanChen", False))
    src.upload_mask(g["cp_mask_t0"])
    for f in range(2):
        src.upload_population(f, g["cp_fluids%d_t0" % f])
    h5io.checkpoint(src, str(tmp_path), sim, 0, g["cp_mask_t0"])
    L = D.Lattice(3, 19, 2, p["gn"])
    L.set_params(p["tau"], p.get("gcc"), p.get("gwc"), p.get("globalAcc"), None,
                 p.get("eqDistForm", "SecondOrder"), p.get("psiForm", "Linear"),
                 p.get("enableShanChen", False))
    assert h5io.restore(L, str(tmp_path), sim, 0) == 0
    L.step(5)
    for f, col in ((0, "red"), (1, "blue")):
        assert_close(L.population(f), g["population_%s_t5" % col], 0.0, "restart %s" % col)
    files = h5io.dump(L, str(tmp_path), sim, 5, fields=("density", "population", "colour"))
    back, _ = h5io.read_field(h5io.output_name(str(tmp_path), "population-red", sim, 5))
    assert_close(back, L.population(0), 0.0, "output population-red")
    back, _ = h5io.read_field(h5io.output_name(str(tmp_path), "colour-red-blue", sim, 5))
    assert_close(back, L.colour(), 0.0, "output colour")
    assert len(files) == 5


def test_bench_channel_workload_bitexact_at_64(D, oracle_mod):
    """The bench.py channel workload (walls at z = 0 and n - 1, seeded random
    solid spheres to a solid fraction of 0.2, red/blue lamellae along x,
    BDist2, g 3.6, body force 1e-6 along x on component 1) at 64^3 on the
    default schedule, 20 steps: bit-identical to the oracle, whose body force
    is the reference's pre-collision hook (lb/collision.d:29,68-70)."""
    import bench
    n = 64
    p = bench.channel_params(n)
    assert p["bodyForce"][1][0] == 1e-6
    p["maskArray"] = bench.channel_mask(n)
    g, o = _pair(D, oracle_mod, p)
    assert g.schedule == "split"
    g.step(20)
    o.run(20)
    _same_state(g, o, "channel 64^3")


def test_bench_channel_workload_full_size_bitexact(D, oracle_mod):
    """The bench.py channel workload at its full 256^3 (16.8 M sites, solid
    fraction 0.2, body force on component 1), 30 steps on the default
    schedule: bit-identical to the oracle (about 20 s of oracle time and
    about 20 GB of host memory for the two states)."""
    import bench
    n = 256
    p = bench.channel_params(n)
    p["maskArray"] = bench.channel_mask(n)
    g, o = _pair(D, oracle_mod, p)
    assert g.schedule == "split"
    g.step(30)
    o.run(30)
    for f in range(2):
        a, b = g.population(f), o.population(f)
        assert a.shape == b.shape
        bad = int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))
        assert bad == 0, "channel 256^3 population %d: %d values differ" % (f, bad)
        del a, b


@pytest.mark.parametrize("name", ["checkpoint_2d", "checkpoint_d1q3"])
def test_checkpoint_lowdim_restart(D, oracle_mod, name):
    """technical/checkpoint-2d (D2Q9) and checkpoint-1d (D1Q3): the t0 restart
    state uploaded and advanced 10 steps equals the oracle bit for bit, and the
    reference's t10 output (bit for bit for D1Q3; within 6 ulps for D2Q9,
    whose data an older revision wrote with D's x87 exp)."""
    from cases import CHECKPOINT_LOWDIM
    g, _ = golden(name)
    p = CHECKPOINT_LOWDIM[name]
    L = D.Lattice(p["d"], p["q"], 2, p["gn"])
    L.set_params(p["tau"], p.get("gcc"), None, None, None, p["eqDistForm"], p["psiForm"],
                 p["enableShanChen"])
    L.upload_mask(g["cp_mask_t0"])
    o = oracle_mod.Oracle.from_params(p)
    for f in range(2):
        L.upload_population(f, g["cp_fluids%d_t0" % f])
        o.set_population(f, g["cp_fluids%d_t0" % f])
    o.set_mask(g["cp_mask_t0"])
    o.exchange_halo()
    L.step(10)
    o.run(10)
    _same_state(L, o, name)
    tol = 2.5e-16 if name == "checkpoint_2d" else 0.0
    for f, c in enumerate(("red", "blue")):
        assert_close(L.population(f), g["population_%s_t10" % c], tol, "t10 " + c)
        assert_close(L.density(f), g["density_%s_t10" % c], 2 * tol, "t10 density " + c)
    assert_close(L.colour(), g["colour_t10"], 2 * tol, "t10 colour")
