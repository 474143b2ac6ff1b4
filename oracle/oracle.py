"""ctypes front end of the C oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: the checker of the HIP product path.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product package ``dlbc_amd`` never does.

The functions mirror the reference's D interface (``src/main.d:76-200``):
``Oracle.from_params`` performs the start-up sequence (``initRNG``,
``initMask``, ``initFluids``, ``exchangeHalo``) for a parameter dictionary that
restates a reference ``.in`` file, and ``run`` executes the ``runTimeloop``
body.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

SECOND_ORDER, THIRD_ORDER, BDIST2 = 0, 1, 2
PSI_LINEAR, PSI_EXPONENTIAL = 0, 1
EQ_FORMS = {"SecondOrder": 0, "ThirdOrder": 1, "BDist2": 2}
PSI_FORMS = {"Linear": 0, "Exponential": 1}
FLUID_INIT = {n: i for i, n in enumerate([
    "None", "File", "Const", "EqDist", "EqDistPerturb", "EqDistPerturbFrac",
    "ConstRandom", "EqDistRandom", "EqDistSphere", "EqDistSphereFrac", "EqDistCylinder",
    "EqDistCylinderFrac", "EqDistTwoSpheres", "EqDistTwoSpheresFrac", "EqDistLamellae",
    "EqDistLamellaeFrac"])}
MASK_INIT = {"None": 0, "File": 1, "Plates": 2, "Tube": 3, "Box": 4, "Walls": 5}
AXIS = {"X": 0, "Y": 1, "Z": 2}
TIMER_NAMES = ["haloExchange", "lb.advection", "force.addSC", "force.distr",
               "lb.coll.prep", "lb.coll.coll", "force.reset"]

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class MT19937(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int), ("front", C.c_uint32)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "or_create": (vp, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "or_destroy": (None, [vp]),
            "or_set_params": (C.c_int, [vp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]),
            "or_set_body_force": (C.c_int, [vp, C.POINTER(C.c_double)]),
            "or_sites_h": (C.c_size_t, [vp]),
            "or_fluid": (C.POINTER(C.c_double), [vp, C.c_int]),
            "or_mask": (C.POINTER(C.c_int32), [vp]),
            "or_density": (C.POINTER(C.c_double), [vp, C.c_int]),
            "or_psi": (C.POINTER(C.c_double), [vp, C.c_int]),
            "or_force": (C.POINTER(C.c_double), [vp, C.c_int]),
            "or_exchange_halo": (None, [vp]),
            "or_advect_fields": (None, [vp]),
            "or_reset_force": (None, [vp]),
            "or_add_shan_chen_force": (None, [vp]),
            "or_distribute_force": (None, [vp]),
            "or_prepare_to_collide": (None, [vp]),
            "or_collide_fields": (None, [vp]),
            "or_precalculate_densities": (None, [vp]),
            "or_run_timesteps": (None, [vp, C.c_int]),
            "or_timers": (None, [vp, _dp]),
            "or_local_total_mass": (C.c_double, [vp, C.c_int]),
            "or_set_threads": (None, [C.c_int]),
            "or_get_threads": (C.c_int, []),
            "or_local_total_momentum": (None, [vp, C.c_int, _dp]),
            "or_velocity_field": (None, [vp, C.c_int, _dp]),
            "or_density_field_phys": (None, [vp, C.c_int, _dp]),
            "or_colour_field": (None, [vp, C.c_int, C.c_int, _dp]),
            "or_get_population": (None, [vp, C.c_int, _dp]),
            "or_set_population": (None, [vp, C.c_int, _dp]),
            "or_get_mask": (None, [vp, _ip]),
            "or_set_mask": (None, [vp, _ip]),
            "or_conn": (C.c_int, [C.c_int, C.c_int, _ip, _ip, _dp, C.POINTER(C.c_double)]),
            "or_pop_density": (C.c_double, [_dp, C.c_int]),
            "or_psi_value": (C.c_double, [C.c_int, C.c_double]),
            "or_exp": (C.c_double, [C.c_double]),
            "or_exp_quad": (C.c_double, [C.c_double]),
            "or_exp_fallbacks": (C.c_int, []),
            "or_eq_dist": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, _dp]),
            "or_mt_seed": (None, [C.POINTER(MT19937), C.c_uint32]),
            "or_mt_next": (C.c_uint32, [C.POINTER(MT19937)]),
            "or_uniform": (C.c_double, [C.POINTER(MT19937), C.c_double, C.c_double]),
            "or_seed_global_rng": (None, [C.c_uint32]),
            "or_init_fluid": (C.c_int, [vp, C.c_int, C.c_int, _dp, C.c_int, C.c_double, C.c_int,
                                        C.c_uint32, C.c_double, _dp, _dp, C.c_int, C.c_double,
                                        _dp]),
            "or_init_mask": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _ip, C.c_int]),
            "or_init_eq_dist_wall": (None, [vp, C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def conn(d, q):
    vel = np.zeros(q * d, np.int32)
    bounce = np.zeros(q, np.int32)
    w = np.zeros(q)
    css = C.c_double()
    if lib().or_conn(d, q, vel, bounce, w, C.byref(css)) != 0:
        raise ValueError("unsupported connectivity D%dQ%d" % (d, q))
    return vel.reshape(q, d), bounce, w, css.value


def exp(x):
    return lib().or_exp(float(x))


def exp_quad(x):
    return lib().or_exp_quad(float(x))


def pop_density(pop):
    pop = np.ascontiguousarray(pop, dtype=np.float64)
    return lib().or_pop_density(pop, pop.size)


def psi_value(form, rho):
    return lib().or_psi_value(int(form), float(rho))


def eq_dist(d, q, form, pop, dv, rho0):
    out = np.zeros(q)
    lib().or_eq_dist(d, q, form, np.ascontiguousarray(pop, np.float64),
                     np.ascontiguousarray(dv, np.float64), float(rho0), out)
    return out


class Rng:
    """std.random.Mt19937 with uniform!"[)"."""

    def __init__(self, seed):
        self.s = MT19937()
        lib().or_mt_seed(C.byref(self.s), seed & 0xFFFFFFFF)

    def next(self):
        return lib().or_mt_next(C.byref(self.s))

    def uniform(self, a, b):
        return lib().or_uniform(C.byref(self.s), a, b)


class Oracle:
    """One rank's lattice (``Lattice!conn``, src/dlbc/lattice.d:43-223)."""

    def __init__(self, d, q, ncomp, n):
        self.d, self.q, self.nc = d, q, ncomp
        self.n = tuple(int(v) for v in n)
        arr = (C.c_int * d)(*self.n)
        self.h = lib().or_create(d, q, ncomp, arr)
        if not self.h:
            raise ValueError("oracle: unsupported lattice")
        self.sites = int(np.prod(self.n))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_destroy(self.h)
            self.h = None

    # -- parameters -------------------------------------------------------
    def set_params(self, tau, gcc=None, gwc=None, global_acc=None, eq_dist_form=0,
                   psi_form=0, enable_shan_chen=False):
        nc, d = self.nc, self.d
        tau = np.asarray(tau, np.float64).reshape(nc)
        gcc = np.zeros((nc, nc)) if gcc is None else np.asarray(gcc, np.float64).reshape(nc, nc)
        gwc = np.zeros(nc) if gwc is None else np.asarray(gwc, np.float64).reshape(nc)
        ga = np.zeros(d) if global_acc is None else np.asarray(global_acc, np.float64).reshape(d)
        rc = lib().or_set_params(self.h, np.ascontiguousarray(tau), np.ascontiguousarray(gcc),
                                 np.ascontiguousarray(gwc), np.ascontiguousarray(ga),
                                 int(eq_dist_form), int(psi_form), int(bool(enable_shan_chen)))
        if rc != 0:
            raise ValueError("oracle: bad parameters")

    def set_body_force(self, body):
        """Register the body-force pre-collision hook: F_f += body[f] after
        distributeForce (lb/collision.d:29,68-70).  None removes it."""
        if body is None:
            lib().or_set_body_force(self.h, None)
            return
        b = np.ascontiguousarray(np.asarray(body, np.float64).reshape(self.nc * self.d))
        lib().or_set_body_force(self.h, b.ctypes.data_as(C.POINTER(C.c_double)))

    # -- passes (src/main.d:174-200) ------------------------------------------
    def exchange_halo(self):
        lib().or_exchange_halo(self.h)

    def run(self, nsteps):
        lib().or_run_timesteps(self.h, int(nsteps))

    def advect(self):
        lib().or_advect_fields(self.h)

    def timers(self):
        out = np.zeros(7)
        lib().or_timers(self.h, out)
        return dict(zip(TIMER_NAMES, out))

    # -- state -------------------------------------------------------------
    def population(self, f):
        out = np.zeros(self.sites * self.q)
        lib().or_get_population(self.h, f, out)
        return out.reshape(self.n + (self.q,))

    def set_population(self, f, pop):
        pop = np.ascontiguousarray(pop, np.float64).reshape(-1)
        assert pop.size == self.sites * self.q
        lib().or_set_population(self.h, f, pop)

    def mask(self):
        out = np.zeros(self.sites, np.int32)
        lib().or_get_mask(self.h, out)
        return out.reshape(self.n)

    def set_mask(self, m):
        m = np.ascontiguousarray(m, np.int32).reshape(-1)
        assert m.size == self.sites
        lib().or_set_mask(self.h, m)

    def raw(self, what, f=0):
        """Halo-padded storage as a numpy view (fluid, density, psi, force, mask);
        the halo is 2, and 4 for D1Q5 (src/dlbc/lattice.d:96-103)."""
        h = 4 if (self.d, self.q) == (1, 5) else 2
        nh = tuple(v + 2 * h for v in self.n)
        if what == "mask":
            p = lib().or_mask(self.h)
            return np.ctypeslib.as_array(p, shape=nh)
        p = getattr(lib(), "or_" + what)(self.h, f)
        extra = {"fluid": (self.q,), "density": (), "psi": (), "force": (self.d,)}[what]
        return np.ctypeslib.as_array(p, shape=nh + extra)

    def density(self, f):
        out = np.zeros(self.sites)
        lib().or_density_field_phys(self.h, f, out)
        return out.reshape(self.n)

    def velocity(self, f):
        out = np.zeros(self.sites * self.d)
        lib().or_velocity_field(self.h, f, out)
        return out.reshape(self.n + (self.d,))

    def colour(self, f1=0, f2=1):
        out = np.zeros(self.sites)
        lib().or_colour_field(self.h, f1, f2, out)
        return out.reshape(self.n)

    def total_mass(self, f):
        return lib().or_local_total_mass(self.h, f)

    def total_momentum(self, f):
        out = np.zeros(self.d)
        lib().or_local_total_momentum(self.h, f, out)
        return out

    # -- diagnostics ---------------------------------------------------------
    def laplace(self, gcc, psi_form):
        """dumpLaplace (src/dlbc/lb/laplace.d:55-218) for two components:
        volume-averaged densities of the 3^d cubes at int(gn/2) and at 1
        (volumeAveragedDensity, laplace.d:236-281), their pressures
        (pressurePsi, lb/density.d:316-326), the radius from the excess mass of
        component 0, sigma from Laplace's law, and for d = 2 the deformation
        from the inertia tensor of the sites with rho_0 > 0.35 (d != 2: NaN)."""
        import math
        d, n = self.d, self.n
        dens = [self.density(f) for f in range(2)]
        mask = self.mask()
        if isinstance(psi_form, str):
            psi_form = PSI_FORMS[psi_form]
        css = 1.0 / 3.0

        def cube(offset):
            lo = [max(0, o - 1) for o in offset]
            hi = [min(n[vd] - 1, offset[vd] + 1) for vd in range(d)]
            out = []
            for f in range(2):
                tot, cnt = 0.0, 0
                for idx in np.ndindex(*[hi[vd] - lo[vd] + 1 for vd in range(d)]):
                    site = tuple(lo[vd] + idx[vd] for vd in range(d))
                    if mask[site] == 0:
                        tot += float(dens[f][site])
                        cnt += 1
                out.append(tot / cnt)
            return out

        def pressure(den):
            pr = 0.0
            for i in range(2):
                pr += den[i]
                for j in range(2):
                    pr += 0.5 * gcc[i][j] * psi_value(psi_form, den[i]) * \
                        psi_value(psi_form, den[j])
            return pr * css

        in_den = cube([int(v * 0.5) for v in n])
        out_den = cube([1] * d)
        in_p, out_p = pressure(in_den), pressure(out_den)
        eff = self.total_mass(0) - float(np.prod(n)) * out_den[0]
        if d == 3:
            r = math.pow(eff / (4.0 / 3.0 * math.pi * (in_den[0] - out_den[0])), 1.0 / 3.0)
            sigma = r * (in_p - out_p) / 2.0
        elif d == 2:
            r = math.pow(eff / (math.pi * (in_den[0] - out_den[0])), 1.0 / 2.0)
            sigma = r * (in_p - out_p)
        else:
            r = eff / (in_den[0] - out_den[0])
            sigma = r * (in_p - out_p)
        deform = float("nan")
        if d == 2:  # laplace.d:118-199
            cutoff = 0.35
            rx = ry = mass = 0.0
            for x in range(n[0]):
                for y in range(n[1]):
                    den = float(dens[0][x, y])
                    if mask[x, y] == 0 and den > cutoff:
                        mass += den
                        rx += den * x
                        ry += den * y
            rx /= mass
            ry /= mass
            ixx = ixy = iyy = 0.0
            for x in range(n[0]):
                for y in range(n[1]):
                    den = float(dens[0][x, y])
                    if mask[x, y] == 0 and den > cutoff:
                        xx, yy = x - rx, y - ry
                        ixx += den * (yy * yy)
                        ixy += den * (xx * yy)
                        iyy += den * (xx * xx)
            tr = ixx + iyy
            det = ixx * iyy - ixy * ixy
            disc = 0.25 * tr * tr - det
            if disc < 0.0:
                deform = 0.0
            else:
                ev1 = 0.5 * tr + math.sqrt(disc)
                ev2 = 0.5 * tr - math.sqrt(disc)
                lmaj = math.sqrt(5.0 * max(ev1, ev2) / mass)
                bmin = math.sqrt(5.0 * min(ev1, ev2) / mass)
                deform = (lmaj - bmin) / (lmaj + bmin)
                if math.isnan(deform):
                    deform = 0.0
        return {"sigma": sigma, "radius": r, "in_pressure": in_p, "out_pressure": out_p,
                "deformation": deform}

    # -- initialisation ----------------------------------------------------
    def init_mask(self, kind="None", axis=0, wall_offset=0, dim_has_wall=None,
                  wall_thickness=1):
        dhw = np.zeros(2 * self.d, np.int32)
        if dim_has_wall is not None:
            dhw[:] = np.asarray(dim_has_wall, np.int32).reshape(-1)
        if isinstance(kind, str):
            kind = MASK_INIT[kind]
        if isinstance(axis, str):
            axis = AXIS[axis]
        if lib().or_init_mask(self.h, kind, axis, wall_offset, dhw, wall_thickness) != 0:
            raise ValueError("oracle: bad mask init")

    def init_fluid(self, f, kind, densities, perturb=0.0, reseed=False, seed=0, radius=0.0,
                   offset=None, separation=None, axis=0, thickness=0.0, widths=None):
        if isinstance(kind, str):
            kind = FLUID_INIT[kind]
        if isinstance(axis, str):
            axis = AXIS[axis]
        dens = np.ascontiguousarray(np.atleast_1d(densities), np.float64)
        off = np.zeros(3) if offset is None else np.ascontiguousarray(
            np.concatenate([np.asarray(offset, np.float64), np.zeros(3)])[:3])
        sep = np.zeros(3) if separation is None else np.ascontiguousarray(
            np.concatenate([np.asarray(separation, np.float64), np.zeros(3)])[:3])
        wid = np.zeros(max(1, dens.size)) if widths is None else np.ascontiguousarray(
            np.asarray(widths, np.float64))
        rc = lib().or_init_fluid(self.h, f, kind, dens, dens.size, float(perturb), int(reseed),
                                 seed & 0xFFFFFFFF, float(radius), off, sep, int(axis),
                                 float(thickness), wid)
        if rc != 0:
            raise ValueError("oracle: bad fluid init")

    def init_eq_dist_wall(self, f):
        lib().or_init_eq_dist_wall(self.h, f)

    @classmethod
    def from_params(cls, p):
        """Start-up sequence of src/main.d:76-147 for a parameter dictionary.

        Keys restate reference parameters: d, q, gn, components, tau, gcc, gwc,
        globalAcc, eqDistForm, psiForm, enableShanChen, seed, maskInit, initAxis,
        wallOffset, dimensionHasWall, wallThickness, fluidInit, fluidDensities,
        fluidPerturb, randomReseed, randomSeed, initRadius, initOffset,
        initSeparation, interfaceThickness, lamellaeWidths.
        """
        d, q, nc = p["d"], p["q"], p.get("components", 1)
        o = cls(d, q, nc, p["gn"])
        o.set_params(p.get("tau", [1.0] * nc), p.get("gcc"), p.get("gwc"), p.get("globalAcc"),
                     EQ_FORMS[p.get("eqDistForm", "SecondOrder")],
                     PSI_FORMS[p.get("psiForm", "Linear")], p.get("enableShanChen", False))
        if p.get("bodyForce") is not None:  # a preCollisionHooks plugin
            o.set_body_force(p["bodyForce"])
        lib().or_seed_global_rng(p.get("seed", 0) & 0xFFFFFFFF)  # initRNG, rank 0
        o.init_mask(p.get("maskInit", "None"), p.get("maskAxis", "X"),
                    p.get("wallOffset", 0), p.get("dimensionHasWall"),
                    p.get("wallThickness", 1))
        if "maskArray" in p:  # MaskInit.File: physical sites read, halo left None
            o.set_mask(p["maskArray"])
        for f in range(nc):  # initFluids, lb/init.d:77-134
            o.init_fluid(f, p["fluidInit"][f], p["fluidDensities"][f],
                         p.get("fluidPerturb", [0.0] * nc)[f],
                         p.get("randomReseed", [False] * nc)[f],
                         p.get("randomSeed", [0] * nc)[f], p.get("initRadius", 0.0),
                         p.get("initOffset"), p.get("initSeparation"),
                         p.get("initAxis", "X"), p.get("interfaceThickness", 0.0),
                         p.get("lamellaeWidths"))
            o.init_eq_dist_wall(f)
        o.exchange_halo()
        return o
