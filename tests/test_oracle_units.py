"""Known-answer checks of the oracle's building blocks.

These restate the reference's D unittest blocks (cited per test) and add
checks that the mathematics fixes (closed forms, invariants, high-precision
references), each chosen so that a plausible slip (a wrong sign, a dropped
term, a transposed index, a sequential instead of pairwise sum) fails.
"""
import math
import random

import numpy as np
import pytest


def test_connectivity_tables(oracle_mod):
    """lb/connectivity.d:432-480: bounce symmetry, weights sum to one,
    c_s^2 = 1/3, plus the isotropy moments sum_i w_i c_ia c_ib = c_s^2 delta."""
    for d, q in ((3, 19), (2, 9), (1, 3), (1, 5)):
        vel, bounce, w, css = oracle_mod.conn(d, q)
        assert css == (1.0 if q == 5 else 1.0 / 3.0)  # connectivity.d:214-248
        assert abs(w.sum() - 1.0) < 1e-15
        for i in range(q):
            assert np.all(vel[bounce[i]] == -vel[i])
            assert bounce[bounce[i]] == i
        m2 = np.einsum("i,ia,ib->ab", w, vel, vel)
        assert np.allclose(m2, css * np.eye(d), atol=1e-15)
        assert np.allclose(np.einsum("i,ia->a", w, vel), 0.0, atol=1e-16)
    # D1Q5 (connectivity.d:279-283, 347-353): velocities 0, 1, -1, 2, -2 and the
    # fourth moment sum_i w_i c^4 = 3 c_s^4
    vel, _, w, _ = oracle_mod.conn(1, 5)
    assert vel[:, 0].tolist() == [0, 1, -1, 2, -2]
    assert w.tolist() == [0.5, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 12.0, 1.0 / 12.0]
    assert abs(np.sum(w * vel[:, 0] ** 4) - 3.0) < 1e-15
    vel, _, _, _ = oracle_mod.conn(3, 19)
    # D3Q19 order fixed by connectivity.d:401-418
    assert vel[7].tolist() == [1, 1, 0] and vel[12].tolist() == [1, 0, -1]
    assert vel[13].tolist() == [1, -1, 0] and vel[18].tolist() == [-1, 0, -1]


def test_density_known_values(oracle_mod):
    """lb/density.d:50-61."""
    pop = np.zeros(19)
    pop[7] = 0.1
    assert oracle_mod.pop_density(pop) == 0.1
    pop = np.zeros(19)
    pop[[0, 2, 7, 10, 14]] = 0.1
    assert oracle_mod.pop_density(pop) == 0.5


def test_density_is_phobos_pairwise(oracle_mod):
    """density() is std.algorithm.sum: pairwise by recursive halving.
    The explicit tree for Q=19 must match bit for bit, and inputs exist on
    which a sequential sum differs."""
    rng = np.random.default_rng(5)

    def tree19(p):
        return 0.0 + ((((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + (p[7] + p[8]))))
                      + (((p[9] + p[10]) + (p[11] + (p[12] + p[13])))
                         + ((p[14] + p[15]) + (p[16] + (p[17] + p[18])))))
    differs = 0
    for _ in range(2000):
        p = rng.uniform(0.0, 0.2, 19) * rng.choice([1e-3, 1.0, 1e3], 19)
        s = oracle_mod.pop_density(p)
        assert s == tree19(p)
        seq = 0.0
        for v in p:
            seq += v
        differs += seq != s
    assert differs > 100


def test_velocity_known_values(oracle_mod):
    """lb/velocity.d:59-74 through the SecondOrder equilibrium at dv = 0:
    the raw velocity of a population is recovered from its equilibrium."""
    pop = np.zeros(19)
    pop[7] = 0.1
    # c_7 = (1,1,0): raw velocity (1,1,0)
    vel, _, _, _ = oracle_mod.conn(3, 19)
    assert (pop @ vel / oracle_mod.pop_density(pop)).tolist() == [1.0, 1.0, 0.0]
    pop = np.zeros(19)
    pop[[0, 2, 7, 10, 14]] = 0.1
    u = pop @ vel / oracle_mod.pop_density(pop)
    assert np.allclose(u, [-0.2, -0.2, 0.2], atol=1e-16)


@pytest.mark.parametrize("form", [0, 1, 2])
@pytest.mark.parametrize("dq", [(3, 19), (2, 9)])
def test_eqdist_conservation(oracle_mod, form, dq):
    """lb/eqdist.d:317-360: mass conservation for all forms; the first moment
    of n_eq equals rho (u + dv) except for BDist2, whose first moment is
    rho dv (it ignores the population's own velocity)."""
    d, q = dq
    vel, _, w, _ = oracle_mod.conn(d, q)
    rng = np.random.default_rng(1)
    for _ in range(50):
        pop = rng.uniform(0.0, 1.0, q)
        dv = rng.uniform(-1e-2, 1e-2, d)
        rho = oracle_mod.pop_density(pop)
        eq = oracle_mod.eq_dist(d, q, form, pop, dv, rho)
        assert abs(eq.sum() - rho) < 1e-13 * rho
        u = pop @ vel / rho
        if form == 2 and d == 3:
            assert np.allclose(eq @ vel, rho * dv, atol=1e-14)
        elif form == 2:
            # the D2Q9 BDist2 constants leave an O(|dv|^3) first-moment term
            assert np.allclose(eq @ vel, rho * dv, atol=2 * rho * np.max(np.abs(dv)) ** 3)
        else:
            assert np.allclose(eq @ vel, rho * (u + dv), atol=1e-14)
    # unity equilibrium (eqDistUnity): the weights for SecondOrder/ThirdOrder
    pop0 = np.zeros(q)
    pop0[0] = 1.0
    eq = oracle_mod.eq_dist(d, q, form, pop0, np.zeros(d), 1.0)
    if form != 2:
        assert np.array_equal(eq, w)
    elif (d, q) == (2, 9):  # BDist2 D2Q9 constants 1/2, 1/12, 1/24 (eqdist.d:185-235)
        assert eq.tolist() == [0.5] + [1.0 / 12] * 4 + [1.0 / 24] * 4
    else:
        assert np.array_equal(eq, w)


def test_secondorder_closed_form(oracle_mod):
    """n_eq = rho w (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2) for D3Q19 (c_s^2 = 1/3)."""
    vel, _, w, _ = oracle_mod.conn(3, 19)
    rng = np.random.default_rng(2)
    pop = rng.uniform(0.01, 0.1, 19)
    dv = np.array([1e-3, -2e-3, 5e-4])
    rho = oracle_mod.pop_density(pop)
    u = pop @ vel / rho + dv
    cu = vel @ u
    ref = rho * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))
    assert np.allclose(oracle_mod.eq_dist(3, 19, 0, pop, dv, rho), ref, rtol=1e-14, atol=0)
    ref3 = ref + rho * w * (4.5 * cu ** 3 - 4.5 * (u @ u) * cu)
    assert np.allclose(oracle_mod.eq_dist(3, 19, 1, pop, dv, rho), ref3, rtol=1e-14, atol=0)


def test_d1q5_equilibria_closed_form(oracle_mod):
    """D1Q5 has c_s^2 = 1: n_eq = rho w (1 + c u + (c u)^2 / 2 - u^2 / 2), plus
    (c u)^3 / 6 - u^2 (c u) / 2 for ThirdOrder; both conserve mass and give
    the first moment rho (u + dv).  BDist2 has no D1Q5 form and returns the
    population (eqdist.d:269-271)."""
    vel, _, w, _ = oracle_mod.conn(1, 5)
    c = vel[:, 0].astype(float)
    rng = np.random.default_rng(3)
    for _ in range(20):
        pop = rng.uniform(0.01, 0.2, 5)
        dv = rng.uniform(-1e-2, 1e-2, 1)
        rho = oracle_mod.pop_density(pop)
        u = (pop @ c) / rho + dv[0]
        cu = c * u
        ref = rho * w * (1 + cu + 0.5 * cu ** 2 - 0.5 * u * u)
        eq2 = oracle_mod.eq_dist(1, 5, 0, pop, dv, rho)
        assert np.allclose(eq2, ref, rtol=1e-13, atol=0)
        ref3 = ref + rho * w * (cu ** 3 / 6.0 - 0.5 * u * u * cu)
        eq3 = oracle_mod.eq_dist(1, 5, 1, pop, dv, rho)
        assert np.allclose(eq3, ref3, rtol=1e-13, atol=0)
        for eq in (eq2, eq3):
            assert abs(eq.sum() - rho) < 1e-14 * rho
            assert abs(eq @ c - rho * u) < 1e-14
        assert np.array_equal(oracle_mod.eq_dist(1, 5, 2, pop, dv, rho), pop)


def test_psi_known_values(oracle_mod):
    """lb/force.d:370-384."""
    assert oracle_mod.psi_value(0, 0.0) == 0.0 and oracle_mod.psi_value(1, 0.0) == 0.0
    assert oracle_mod.psi_value(0, 0.7) == 0.7
    assert math.isclose(oracle_mod.psi_value(1, 0.7), 0.5034146962, rel_tol=1e-9)
    assert math.isclose(oracle_mod.psi_value(1, 1.0), 0.63212055882, rel_tol=1e-9)
    assert math.isclose(oracle_mod.psi_value(1, 2.0), 0.86466471676, rel_tol=1e-9)


def _mp_exp(x):
    """exp(x) at 256 bits, rounded once to double.  mpmath's float() rounds
    to 53 bits before scaling, which double-rounds subnormal results, so
    those are rounded here on the 2^-1074 grid directly."""
    import mpmath
    with mpmath.workprec(256):
        v = mpmath.exp(mpmath.mpf(x))
        if v < mpmath.mpf(2) ** -1022:
            return int(mpmath.nint(v * mpmath.mpf(2) ** 1074)) * 2.0 ** -1074
        return float(v)


def test_exp_correctly_rounded(oracle_mod):
    """or_exp is the correctly rounded exp: checked against a 256-bit mpmath
    evaluation on 20000 random arguments in the densities' range, on
    arguments spread over the whole double range, and at special values."""
    pytest.importorskip("mpmath")
    rnd = random.Random(11)
    xs = [-rnd.uniform(0.0, 3.0) for _ in range(15000)]
    xs += [rnd.uniform(-745.0, 709.7) for _ in range(4000)]
    xs += [rnd.uniform(-1e-6, 1e-6) for _ in range(1000)]
    bad = [(x, oracle_mod.exp(x), _mp_exp(x)) for x in xs if oracle_mod.exp(x) != _mp_exp(x)]
    assert not bad, bad[:5]
    assert oracle_mod.exp(0.0) == 1.0 and oracle_mod.exp(-0.0) == 1.0
    assert oracle_mod.exp(1.0) == _mp_exp(1.0)
    assert oracle_mod.exp(710.0) == math.inf and oracle_mod.exp(-746.0) == 0.0
    assert math.isnan(oracle_mod.exp(math.nan))
    # subnormal results
    for x in (-708.5, -720.0, -744.0):
        assert oracle_mod.exp(x) == _mp_exp(x)


def test_exp_quad_reference(oracle_mod):
    """The 113-bit path used for hard cases agrees with mpmath too."""
    pytest.importorskip("mpmath")
    rnd = random.Random(3)
    for _ in range(2000):
        x = rnd.uniform(-20.0, 20.0)
        assert oracle_mod.exp_quad(x) == _mp_exp(x)


def test_mt19937_golden_values(oracle_mod):
    """random.d:55-89: uniform(0.0, 2.0) golden values for seeds 0, 1, 42."""
    gold = {0: [1.0976270048640728305144875776022672653198242187500000,
                1.1856892330538686408658577420283108949661254882812500,
                1.4303787302762218658358506218064576387405395507812500],
            1: [0.8340439970684339066053780697984620928764343261718750,
                1.9943696167306901312343825338757596909999847412109375,
                1.4406489789114911292955412136507220566272735595703125],
            42: [0.7490802287936861869610538633423857390880584716796875,
                 1.5930859687722020989752991226851008832454681396484375,
                 1.9014286235676676195538448155275546014308929443359375]}
    for seed, vals in gold.items():
        r = oracle_mod.Rng(seed)
        assert [r.uniform(0.0, 2.0) for _ in range(3)] == vals
    # raw outputs equal numpy's MT19937 with init_genrand seeding
    r = oracle_mod.Rng(43)
    ref = np.random.RandomState(43).randint(0, 2 ** 32, size=2000, dtype=np.uint64)
    assert [r.next() for _ in range(2000)] == ref.tolist()


def test_mass_and_momentum_conservation(oracle_mod):
    """Periodic, no force, no walls: total mass is conserved to round-off and
    total momentum stays zero-change (BDist2 and SecondOrder)."""
    # SecondOrder conserves each component's momentum for any tau; BDist2
    # shares one weighted velocity and conserves the total for equal tau.
    for form, tau in (("SecondOrder", [0.8, 1.3]), ("BDist2", [0.9, 0.9])):
        p = dict(d=3, q=19, gn=[8, 6, 5], components=2, tau=tau, eqDistForm=form,
                 fluidInit=["EqDistPerturb"] * 2, fluidDensities=[[0.7], [0.3]],
                 fluidPerturb=[0.05, 0.05], randomReseed=[True, True], randomSeed=[1, 2])
        o = oracle_mod.Oracle.from_params(p)
        m0 = [o.total_mass(f) for f in range(2)]
        p0 = sum(o.total_momentum(f) for f in range(2))
        o.run(20)
        for f in range(2):
            assert abs(o.total_mass(f) - m0[f]) < 1e-12 * m0[f]
        p1 = sum(o.total_momentum(f) for f in range(2))
        assert np.allclose(p1, p0, atol=1e-13)


def test_bounce_back_wall_no_flux(oracle_mod):
    """Full-way bounce-back: with a closed box nothing leaves the fluid, so
    mass is conserved and the streaming of a population next to a wall
    returns its own opposite population."""
    p = dict(d=2, q=9, gn=[6, 5], components=1, tau=[1.0], maskInit="Box",
             fluidInit=["EqDistPerturb"], fluidDensities=[[1.0]], fluidPerturb=[0.1],
             randomReseed=[True], randomSeed=[7])
    o = oracle_mod.Oracle.from_params(p)
    m0 = o.total_mass(0)
    pre = o.population(0).copy()
    o.exchange_halo()
    o.advect()
    post = o.population(0)
    _, bounce, _, _ = oracle_mod.conn(2, 9)
    # site (1,1) is fluid; its neighbour in -x direction (0,1) is solid, so
    # population q=1 (c=(1,0)) comes from its own q=2 population
    assert post[1, 1, 1] == pre[1, 1, bounce[1]]
    o.run(30)
    assert abs(o.total_mass(0) - m0) < 1e-12 * m0


def _wall_box(oracle_mod, gcc, gwc, rho=0.6):
    """Uniform fluid at rest between solid planes z = 0 and z = n-1."""
    n = [5, 4, 7]
    mask = np.zeros(n, np.int32)
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    p = dict(d=3, q=19, gn=n, components=2, tau=[1.0, 1.0], eqDistForm="SecondOrder",
             gcc=gcc, gwc=gwc, enableShanChen=True, psiForm="Exponential",
             fluidInit=["EqDist"] * 2, fluidDensities=[[rho], [rho]], maskArray=mask)
    o = oracle_mod.Oracle.from_params(p)
    o.run(1)
    phys = (slice(2, -2),) * 3
    return [o.raw("force", f)[phys].copy() for f in range(2)], oracle_mod.psi_value(1, rho)


def test_shan_chen_force_solid_substitution(oracle_mod):
    """lb/force.d:253-283: a solid neighbour contributes the site's own psi.
    On a uniform state next to walls every force is then exactly zero; any
    other value at the solid sites (their density 1, or 0) leaves a force on
    the layers next to the walls."""
    F, _ = _wall_box(oracle_mod, [[0.0, 3.6], [3.6, 0.0]], [0.0, 0.0])
    for f in range(2):
        assert np.all(F[f][:, :, 1:-1] == 0.0)


def test_shan_chen_wall_force_closed_form(oracle_mod):
    """lb/force.d:284-307 with gcc = 0: the site next to the wall z = 0 has
    solid neighbours along c_z = -1 (one with w = 1/18, four with w = 1/36)
    whose density is 1 (solid sites hold [1, 0, ...], fields/init.d:469-477),
    so F = -sum_i w_i psi gwc c_i = (0, 0, +psi gwc / 6); the mirror image
    at z = n-2; zero in between."""
    gwc = [0.25, -0.4]
    F, psi = _wall_box(oracle_mod, [[0.0, 0.0], [0.0, 0.0]], gwc)
    for f in range(2):
        fz = psi * gwc[f] / 6.0
        assert np.allclose(F[f][:, :, 1, 2], fz, rtol=4e-16, atol=0)
        assert np.allclose(F[f][:, :, -2, 2], -fz, rtol=4e-16, atol=0)
        assert np.all(np.abs(F[f][:, :, 1:-1, :2]) <= 1e-18)
        assert np.all(F[f][:, :, 2:-2, :] == 0.0)


def test_quasi_1d_across_dimensions(oracle_mod):
    """lb/laplace-quasi-1d/compare-lower-dimensions.py on the oracle: the same
    x-lamellae (ThirdOrder, Shan-Chen, 1000 steps) on D3Q19 64x4x4, D2Q9 64x4
    and D1Q3 64 give density profiles equal to 1e-14.  A wrong weight, speed
    of sound or stencil term in one connectivity breaks this."""
    base = dict(components=2, tau=[1.0, 1.0], eqDistForm="ThirdOrder",
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistLamellaeFrac"] * 2,
                fluidDensities=[[0.5, 0.7, 0.5], [0.7, 0.5, 0.7]],
                lamellaeWidths=[0.25, 0.50, 0.25], interfaceThickness=5.0)
    prof = {}
    for d, q, gn in ((3, 19, [64, 4, 4]), (2, 9, [64, 4]), (1, 3, [64])):
        o = oracle_mod.Oracle.from_params(dict(base, d=d, q=q, gn=gn))
        o.run(1000)
        prof[d] = [o.density(f).reshape(64, -1) for f in range(2)]
    for f in range(2):
        for d in (3, 2):
            assert np.abs(prof[d][f] - prof[d][f][:, :1]).max() == 0.0
            assert np.abs(prof[d][f][:, 0] - prof[1][f][:, 0]).max() <= 1e-14



@pytest.mark.parametrize("form", ["SecondOrder", "BDist2"])
def test_body_force_momentum_closed_form(oracle_mod, form):
    """The body-force pre-collision hook (F_f += b_f after distributeForce,
    lb/collision.d:29,68-70) through the velocity shift u + tau F / rho: with
    omega tau = 1 the collision gives the momentum j + F per site, so on a
    periodic lattice each step adds exactly N b_f to the total momentum of
    component f, on the forced component and axis only.  A force acting on
    the wrong component, axis or with the wrong sign, or divided by rho
    twice, fails here."""
    n = [6, 5, 4]
    N = 6 * 5 * 4
    # BDist2 replaces the component's own velocity by the weighted velocity
    # u_w = sum_f (j_f / tau_f) / sum_f rho_f (collision.d:156-196): with
    # tau = 1 the components share momentum through it and their total grows
    # by N (b_0 + b_1) per step; SecondOrder keeps each component's own
    tau = [1.0, 1.0] if form == "BDist2" else [1.0, 0.8]
    p = dict(d=3, q=19, gn=n, components=2, tau=tau, eqDistForm=form,
             fluidInit=["EqDistPerturb"] * 2, fluidDensities=[[0.9], [0.4]],
             fluidPerturb=[0.01, 0.01], randomReseed=[True, True], randomSeed=[3, 4],
             bodyForce=[[0.0, 0.0, 0.0], [2e-5, -1e-5, 0.0]])
    o = oracle_mod.Oracle.from_params(p)
    j0 = [o.total_momentum(f) for f in range(2)]
    steps = 7
    o.run(steps)
    j1 = [o.total_momentum(f) for f in range(2)]
    want = np.array([2e-5, -1e-5, 0.0]) * N * steps
    if form == "BDist2":
        got = (j1[0] + j1[1]) - (j0[0] + j0[1])
        assert np.allclose(got, want, rtol=0, atol=1e-15), got
    else:
        assert np.allclose(j1[1] - j0[1], want, rtol=0, atol=1e-15), j1[1] - j0[1]
        assert np.allclose(j1[0] - j0[0], 0.0, rtol=0, atol=1e-15), j1[0] - j0[0]
    # and without the hook nothing moves
    q = oracle_mod.Oracle.from_params(dict(p, bodyForce=None))
    q.run(steps)
    drift = (q.total_momentum(0) + q.total_momentum(1)) - (j0[0] + j0[1])
    assert np.allclose(drift, 0.0, rtol=0, atol=1e-15), drift


def test_body_force_poiseuille_closed_form(oracle_mod):
    """A body force density b on a component of density rho between plates
    drives the same Poiseuille profile as the global acceleration b / rho:
    v(x) = (b/rho) / (2 nu) (x - 1/2)(L - (x - 1/2)) at steady state, up to
    the constant offset of the velocity-shift forcing (lb/poiseuille-2d, with
    its globalAcc 1e-4 given as forces 0.7e-4 and 1e-4 on its two
    components of density 0.7 and 1.0)."""
    from cases import case
    p = case("poiseuille_2d", globalAcc=[0.0, 0.0], bodyForce=[[0.0, 0.7e-4], [0.0, 1e-4]])
    o = oracle_mod.Oracle.from_params(p)
    o.run(30000)
    x = np.arange(1, 31) - 0.5
    nu = (1.0 - 0.5) / 3.0
    an = 1e-4 / (2 * nu) * x * (30 - x)
    for f in range(2):
        v = o.velocity(f)[1:31, 0, 1]
        off = v - an
        assert np.ptp(off) < 1e-9, ("not the Poiseuille parabola", f, np.ptp(off))
        assert abs(off[0]) < 2e-4
        assert np.max(np.abs(o.velocity(f)[1:31, :, 0])) < 1e-12  # none across the plates


@pytest.mark.parametrize("tau", [0.6, 0.8, 1.0, 1.25])
def test_bdist2_total_momentum_factor(oracle_mod, tau):
    """BDist2 relaxes every component towards the weighted velocity
    u_w = sum_f (j_f / tau_f) / sum_f rho_f (lb/collision.d:156-196), and its
    equilibrium carries rho_f u_w.  With equal tau the collision therefore
    maps the total momentum J to (1 - omega) J + omega (omega J) =
    (1 - omega + omega^2) J: conserved at tau = 1 only, amplified by a factor
    above one for every tau < 1.  Checked on a uniform two-component state
    moving at a small velocity, no forces, one step (streaming moves a
    uniform state into itself)."""
    n = [4, 4, 4]
    o = oracle_mod.Oracle(3, 19, 2, n)
    o.set_params([tau, tau], eq_dist_form=2)
    vel, _, _, _ = oracle_mod.conn(3, 19)
    rhos, u0 = (0.7, 0.3), np.array([1e-3, -2e-3, 5e-4])
    for f, rho in enumerate(rhos):
        # SecondOrder equilibrium at velocity u0 as the moving initial state
        rest = np.zeros(19)
        rest[0] = rho
        eq = oracle_mod.eq_dist(3, 19, 0, rest, u0, rho)
        o.set_population(f, np.broadcast_to(eq, tuple(n) + (19,)).copy())
    o.exchange_halo()
    j0 = o.total_momentum(0) + o.total_momentum(1)
    o.run(1)
    j1 = o.total_momentum(0) + o.total_momentum(1)
    om = 1.0 / tau
    assert np.allclose(j1, (1.0 - om + om * om) * j0, rtol=1e-12, atol=0), (j1, j0)


@pytest.mark.parametrize("g", [3.0, 4.5])
def test_ensemble_droplets_diverge_below_tau_one(oracle_mod, g):
    """The batched-ensemble droplets (EqDistSphereFrac 0.7/0.04, radius 0.25 L,
    BDist2) at 32^3: with tau = 0.6, the low end of the ensemble's specified
    range, the state is no longer finite within 500 steps for both ends of
    the coupling range (the momentum factor of the test above); with tau = 1
    the same droplet stays finite.  This is why the ensemble draws tau from
    [1.0, 1.5]."""
    import bench
    base = dict(bench.ensemble_params(1)[0], gn=[32, 32, 32], initRadius=0.25,
                gcc=[[0.0, g], [g, 0.0]])
    o = oracle_mod.Oracle.from_params(dict(base, tau=[0.6, 0.6]))
    diverged_at = None
    for t in range(25, 501, 25):
        o.run(25)
        if not np.all(np.isfinite(o.density(0))):
            diverged_at = t
            break
    assert diverged_at is not None and diverged_at <= 500
    s = oracle_mod.Oracle.from_params(dict(base, tau=[1.0, 1.0]))
    s.run(200)
    assert np.all(np.isfinite(s.density(0))) and np.all(np.isfinite(s.density(1)))
