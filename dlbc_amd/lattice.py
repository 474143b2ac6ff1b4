"""ctypes front end of libdlbc_hip.so (include/dlbc_hip.h)."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libdlbc_hip.so")
_lib = None

EQ_FORMS = {"SecondOrder": 0, "ThirdOrder": 1, "BDist2": 2}
PSI_FORMS = {"Linear": 0, "Exponential": 1}
FLUID_INIT = {n: i for i, n in enumerate([
    "None", "File", "Const", "EqDist", "EqDistPerturb", "EqDistPerturbFrac",
    "ConstRandom", "EqDistRandom", "EqDistSphere", "EqDistSphereFrac", "EqDistCylinder",
    "EqDistCylinderFrac", "EqDistTwoSpheres", "EqDistTwoSpheresFrac", "EqDistLamellae",
    "EqDistLamellaeFrac"])}
MASK_INIT = {"None": 0, "File": 1, "Plates": 2, "Tube": 3, "Box": 4, "Walls": 5}
AXIS = {"X": 0, "Y": 1, "Z": 2}
STORAGE = {"f64": 0, "f32_shift": 1, "f16_shift": 2}
SCHEDULES = {"auto": 0, "two_pass": 1, "fused": 2, "split": 3, "march": 4}
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of memory",
          -4: "unsupported configuration", -5: "call out of order"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class DlbcError(RuntimeError):
    pass


class _Rng(C.Structure):
    _fields_ = [("state", C.c_uint32 * 627)]


def lib_path():
    return _LIB


def lib():
    """Load libdlbc_hip.so (built by dlbc_amd/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise DlbcError("libdlbc_hip.so is not built (run python dlbc_amd/build.py)")
    L = C.CDLL(_LIB)
    vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    opt_dp = C.POINTER(C.c_double)
    sig = {
        "dlbc_gpu_create": (i, [C.POINTER(vp), i, i, i, _i64p, i, i]),
        "dlbc_gpu_set_params": (i, [vp, opt_dp, opt_dp, opt_dp, opt_dp, opt_dp, i, i, i, opt_dp]),
        "dlbc_gpu_upload_populations": (i, [vp, i, _dp, i]),
        "dlbc_gpu_upload_mask": (i, [vp, _ip, i]),
        "dlbc_gpu_download_mask": (i, [vp, _ip, i]),
        "dlbc_gpu_step": (i, [vp, i]),
        "dlbc_gpu_step_graph": (i, [vp, i]),
        "dlbc_gpu_advect": (i, [vp]),
        "dlbc_gpu_step_timed": (i, [vp, i, C.POINTER(d), opt_dp, C.POINTER(i)]),
        "dlbc_gpu_kernel_name": (C.c_char_p, [vp, i]),
        "dlbc_gpu_synchronize": (i, [vp]),
        "dlbc_gpu_set_schedule": (i, [vp, i]),
        "dlbc_gpu_get_schedule": (i, [vp]),
        "dlbc_gpu_set_launch": (i, [vp, i]),
        "dlbc_gpu_set_block_order": (i, [vp, i]),
        "dlbc_gpu_set_march": (i, [vp, i, i]),
        "dlbc_gpu_get_block_order": (i, [vp]),
        "dlbc_gpu_laplace": (i, [vp, _dp]),
        "dlbc_batch_create": (i, [C.POINTER(vp), C.POINTER(vp), i]),
        "dlbc_batch_set_config": (i, [vp, i, i, i, i]),
        "dlbc_batch_get_config": (i, [vp, C.POINTER(i)]),
        "dlbc_batch_autotune": (i, [vp, C.POINTER(i), C.POINTER(d)]),
        "dlbc_batch_step": (i, [vp, i, i]),
        "dlbc_batch_step_looped": (i, [vp, i]),
        "dlbc_batch_synchronize": (i, [vp]),
        "dlbc_batch_last_error": (C.c_char_p, [vp]),
        "dlbc_batch_destroy": (None, [vp]),
        "dlbc_gpu_download_populations": (i, [vp, i, _dp, i]),
        "dlbc_gpu_density": (i, [vp, i, _dp, i]),
        "dlbc_gpu_velocity": (i, [vp, i, _dp, i]),
        "dlbc_gpu_colour": (i, [vp, i, i, _dp, i]),
        "dlbc_gpu_total_mass": (i, [vp, i, _dp]),
        "dlbc_gpu_total_mass_noalloc": (i, [vp, i, _dp]),
        "dlbc_gpu_total_momentum": (i, [vp, i, _dp]),
        "dlbc_gpu_eval_psi": (i, [vp, i, _dp, _dp, i64]),
        "dlbc_gpu_last_error": (C.c_char_p, [vp]),
        "dlbc_selftest_division": (i64, [i64, C.c_uint64]),
        "dlbc_selftest_site_division": (i64, [i64, C.c_uint64]),
        "dlbc_gpu_destroy": (None, [vp]),
        "dlbc_gpu_device_bytes": (i64, [vp]),
        "dlbc_rng_seed": (None, [C.POINTER(_Rng), C.c_uint32]),
        "dlbc_rng_next": (C.c_uint32, [C.POINTER(_Rng)]),
        "dlbc_rng_uniform": (d, [C.POINTER(_Rng), d, d]),
        "dlbc_init_fluid": (i, [_dp, i, i, _i64p, i, i, i, _dp, i, d, C.POINTER(_Rng), d,
                                _dp, _dp, i, d, _dp]),
        "dlbc_init_mask": (i, [_ip, i, _i64p, i, i, i, i, _ip, i]),
        "dlbc_init_eq_dist_wall": (i, [_dp, _ip, i, i, _i64p, i]),
        "dlbc_gpu_init_fluid": (i, [vp, i, i, i, i, _dp, i, d, C.POINTER(_Rng), d, _dp, _dp, i,
                                    d, _dp]),
        "dlbc_gpu_init_mask": (i, [vp, i, i, i, _ip, i]),
        "dlbc_gpu_init_eq_dist_wall": (i, [vp, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _opt(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return a.ctypes.data_as(C.POINTER(C.c_double)), a


class Rng:
    """std.random.Mt19937 (src/dlbc/random.d:24) with uniform!"[)"."""

    def __init__(self, seed):
        self.s = _Rng()
        lib().dlbc_rng_seed(C.byref(self.s), int(seed) & 0xFFFFFFFF)

    def next(self):
        return lib().dlbc_rng_next(C.byref(self.s))

    def uniform(self, a, b):
        return lib().dlbc_rng_uniform(C.byref(self.s), float(a), float(b))


def reference_halo(d, q):
    """The halo of the reference's fields (src/dlbc/lattice.d:96-103): 2, and
    4 for D1Q5, whose +-2 links and the Shan-Chen stencil on them reach four
    sites out.  It fixes the order of the random draws of the initialisation."""
    return 4 if (d == 1 and q == 5) else 2


def _padded_shape(n, halo):
    return tuple(int(v) + 2 * halo for v in n)


def _fluid_args(kind, densities, eq_dist_form, perturb, rng, radius, offset, separation, axis,
                thickness, widths):
    """The argument list dlbc_init_fluid and dlbc_gpu_init_fluid share, from
    kind onwards (eq_dist_form first)."""
    if isinstance(kind, str):
        kind = FLUID_INIT[kind]
    if isinstance(axis, str):
        axis = AXIS[axis]
    if isinstance(eq_dist_form, str):
        eq_dist_form = EQ_FORMS[eq_dist_form]
    dens = np.ascontiguousarray(np.atleast_1d(densities), np.float64)
    off = np.zeros(3) if offset is None else np.ascontiguousarray(
        np.concatenate([np.asarray(offset, np.float64), np.zeros(3)])[:3])
    sep = np.zeros(3) if separation is None else np.ascontiguousarray(
        np.concatenate([np.asarray(separation, np.float64), np.zeros(3)])[:3])
    wid = np.zeros(max(1, dens.size)) if widths is None else np.ascontiguousarray(
        np.asarray(widths, np.float64))
    r = rng.s if rng is not None else _Rng()
    return (int(eq_dist_form), int(kind), dens, dens.size, float(perturb), C.byref(r),
            float(radius), off, sep, int(axis), float(thickness), wid)


# D's default-initialised double, the quiet NaN 0x7FFC000000000000: a field
# starts out so, and FluidInit.None leaves it (lb/init-2d reference data).
D_DOUBLE_INIT = np.array([0x7FFC000000000000], np.uint64).view(np.float64)[0]


def init_fluid(d, q, n, kind, densities, eq_dist_form=0, perturb=0.0, rng=None, radius=0.0,
               offset=None, separation=None, axis=0, thickness=0.0, widths=None, halo=2):
    """initFluid (src/dlbc/lb/init.d:263-339) into a new halo-padded array."""
    out = np.full(_padded_shape(n, halo) + (q,), D_DOUBLE_INIT)
    a = _fluid_args(kind, densities, eq_dist_form, perturb, rng, radius, offset, separation, axis,
                    thickness, widths)
    rc = lib().dlbc_init_fluid(out.reshape(-1), d, q, np.asarray(n, np.int64), halo, *a)
    if rc != 0:
        raise DlbcError("dlbc_init_fluid: %s" % ERRORS.get(rc, rc))
    return out


def _mask_args(d, kind, axis, dim_has_wall):
    if isinstance(kind, str):
        kind = MASK_INIT[kind]
    if isinstance(axis, str):
        axis = AXIS[axis]
    dhw = np.zeros(2 * d, np.int32)
    if dim_has_wall is not None:
        dhw[:] = np.asarray(dim_has_wall, np.int32).reshape(-1)
    return int(kind), int(axis), dhw


def init_mask(d, n, kind="None", axis=0, wall_offset=0, dim_has_wall=None, wall_thickness=1,
              halo=2):
    """initMask (src/dlbc/lb/mask.d:134-172) into a new halo-padded int32 array."""
    kind, axis, dhw = _mask_args(d, kind, axis, dim_has_wall)
    out = np.zeros(_padded_shape(n, halo), np.int32)
    rc = lib().dlbc_init_mask(out.reshape(-1), d, np.asarray(n, np.int64), halo, kind, axis,
                              int(wall_offset), dhw, int(wall_thickness))
    if rc != 0:
        raise DlbcError("dlbc_init_mask: %s" % ERRORS.get(rc, rc))
    return out


def init_eq_dist_wall(pop, mask, d, q, n, halo=2):
    rc = lib().dlbc_init_eq_dist_wall(pop.reshape(-1), np.ascontiguousarray(mask).reshape(-1), d,
                                      q, np.asarray(n, np.int64), halo)
    if rc != 0:
        raise DlbcError("dlbc_init_eq_dist_wall failed")


class Lattice:
    """A lattice advanced on the GPU through the C ABI."""

    def __init__(self, d, q, ncomp, n, storage="f64", device=0):
        self.d, self.q, self.nc = d, q, ncomp
        self.n = tuple(int(v) for v in n)
        self.sites = int(np.prod(self.n))
        h = C.c_void_p()
        rc = lib().dlbc_gpu_create(C.byref(h), d, q, ncomp, np.asarray(self.n, np.int64),
                                   STORAGE[storage] if isinstance(storage, str) else storage,
                                   device)
        if rc != 0:
            raise DlbcError("dlbc_gpu_create: %s" % ERRORS.get(rc, rc))
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().dlbc_gpu_last_error(self.h)
            raise DlbcError("%s: %s (%s)" % (what, ERRORS.get(rc, rc),
                                             msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "h", None):
            lib().dlbc_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, tau, gcc=None, gwc=None, global_acc=None, body_force=None,
                   eq_dist_form=0, psi_form=0, enable_shan_chen=False, rho0=None):
        if isinstance(eq_dist_form, str):
            eq_dist_form = EQ_FORMS[eq_dist_form]
        if isinstance(psi_form, str):
            psi_form = PSI_FORMS[psi_form]
        keep = []

        def ptr(a):
            r = _opt(a)
            if r is None:
                return None
            keep.append(r[1])
            return r[0]
        rc = lib().dlbc_gpu_set_params(self.h, ptr(tau), ptr(gcc), ptr(gwc), ptr(global_acc),
                                       ptr(body_force), int(eq_dist_form), int(psi_form),
                                       int(bool(enable_shan_chen)), ptr(rho0))
        self._check(rc, "set_params")

    def upload_mask(self, mask, halo=0):
        m = np.ascontiguousarray(mask, np.int32)
        assert m.size == int(np.prod(_padded_shape(self.n, halo)))
        self._check(lib().dlbc_gpu_upload_mask(self.h, m.reshape(-1), halo), "upload_mask")

    def upload_population(self, f, pop, halo=0):
        p = np.ascontiguousarray(pop, np.float64)
        assert p.size == int(np.prod(_padded_shape(self.n, halo))) * self.q
        self._check(lib().dlbc_gpu_upload_populations(self.h, f, p.reshape(-1), halo),
                    "upload_populations")

    def init_fluid(self, f, kind, densities, eq_dist_form=0, perturb=0.0, rng=None, radius=0.0,
                   offset=None, separation=None, axis=0, thickness=0.0, widths=None, halo=2):
        """initFluid (src/dlbc/lb/init.d:263-339) of component f computed on
        the device: the populations of init_fluid() + upload_population(),
        without the host array.  Random kinds draw from rng in the order of
        the halo-padded field (halo 2 in the reference)."""
        a = _fluid_args(kind, densities, eq_dist_form, perturb, rng, radius, offset, separation,
                        axis, thickness, widths)
        self._check(lib().dlbc_gpu_init_fluid(self.h, int(f), int(halo), *a), "init_fluid")

    def init_mask(self, kind="None", axis=0, wall_offset=0, dim_has_wall=None,
                  wall_thickness=1):
        """initMask (src/dlbc/lb/mask.d:134-172) computed on the device."""
        kind, axis, dhw = _mask_args(self.d, kind, axis, dim_has_wall)
        self._check(lib().dlbc_gpu_init_mask(self.h, kind, axis, int(wall_offset), dhw,
                                             int(wall_thickness)), "init_mask")

    def init_eq_dist_wall(self, f):
        """initEqDistWall (src/dlbc/fields/init.d:469-477) on the device."""
        self._check(lib().dlbc_gpu_init_eq_dist_wall(self.h, int(f)), "init_eq_dist_wall")

    def step(self, nsteps=1, graph=False):
        """nsteps time steps (src/main.d:174-200); graph=True replays them
        from a captured two-step HIP graph (same bits, less launch latency)."""
        fn = lib().dlbc_gpu_step_graph if graph else lib().dlbc_gpu_step
        self._check(fn(self.h, int(nsteps)), "step")

    def advect(self):
        """advectFields alone (streaming with bounce-back)."""
        self._check(lib().dlbc_gpu_advect(self.h), "advect")

    def step_timed(self, nsteps):
        el = C.c_double()
        km = (C.c_double * 8)()
        nk = C.c_int()
        self._check(lib().dlbc_gpu_step_timed(self.h, int(nsteps), C.byref(el), km, C.byref(nk)),
                    "step_timed")
        names = [lib().dlbc_gpu_kernel_name(self.h, k).decode() for k in range(nk.value)]
        return el.value, dict(zip(names, [km[k] for k in range(nk.value)]))

    def synchronize(self):
        self._check(lib().dlbc_gpu_synchronize(self.h), "synchronize")

    def set_schedule(self, s):
        self._check(lib().dlbc_gpu_set_schedule(self.h, SCHEDULES[s] if isinstance(s, str) else s),
                    "set_schedule")

    def set_launch(self, min_workgroups):
        """Minimum workgroups per fused-kernel launch (0: default)."""
        self._check(lib().dlbc_gpu_set_launch(self.h, int(min_workgroups)), "set_launch")

    def set_march(self, rows=0, hold=-1):
        """March schedule: z-rows per workgroup (1 or 2; 0 automatic) and
        whether populations stay in registers between pull and collision
        (1 yes, 0 pulled again, -1 automatic)."""
        self._check(lib().dlbc_gpu_set_march(self.h, int(rows), int(hold)), "set_march")

    def set_block_order(self, chunk):
        """Split schedule: K > 0 blocks per XCD turn, 0 site order, -1 automatic."""
        self._check(lib().dlbc_gpu_set_block_order(self.h, int(chunk)), "set_block_order")

    @property
    def block_order(self):
        return lib().dlbc_gpu_get_block_order(self.h)

    @property
    def schedule(self):
        v = lib().dlbc_gpu_get_schedule(self.h)
        return {0: "auto", 1: "two_pass", 2: "fused", 3: "split", 4: "march"}[v]

    def population(self, f, halo=0):
        out = np.zeros(_padded_shape(self.n, halo) + (self.q,))
        self._check(lib().dlbc_gpu_download_populations(self.h, f, out.reshape(-1), halo),
                    "download_populations")
        return out

    def mask(self, halo=0):
        """The mask field (int32, 0 fluid, 1 solid) as the reference writes it."""
        out = np.zeros(_padded_shape(self.n, halo), np.int32)
        self._check(lib().dlbc_gpu_download_mask(self.h, out.reshape(-1), halo), "download_mask")
        return out

    def density(self, f):
        out = np.zeros(self.n)
        self._check(lib().dlbc_gpu_density(self.h, f, out.reshape(-1), 0), "density")
        return out

    def velocity(self, f):
        out = np.zeros(self.n + (self.d,))
        self._check(lib().dlbc_gpu_velocity(self.h, f, out.reshape(-1), 0), "velocity")
        return out

    def colour(self, f1=0, f2=1):
        out = np.zeros(self.n)
        self._check(lib().dlbc_gpu_colour(self.h, f1, f2, out.reshape(-1), 0), "colour")
        return out

    def total_mass(self, f, parts=False):
        out = np.zeros(2)
        self._check(lib().dlbc_gpu_total_mass(self.h, f, out), "total_mass")
        return (out[0], out[1]) if parts else out[0]

    def total_momentum(self, f):
        out = np.zeros(2 * self.d)
        self._check(lib().dlbc_gpu_total_momentum(self.h, f, out), "total_momentum")
        return out[:self.d]

    def eval_psi(self, form, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.zeros_like(x)
        self._check(lib().dlbc_gpu_eval_psi(self.h, int(form), x, y, x.size), "eval_psi")
        return y

    def laplace(self):
        """Laplace diagnostics (src/dlbc/lb/laplace.d:55-218) of a centred
        two-component droplet: sigma, radius, in/out pressure, deformation."""
        out = np.zeros(5)
        self._check(lib().dlbc_gpu_laplace(self.h, out), "laplace")
        return dict(zip(("sigma", "radius", "in_pressure", "out_pressure", "deformation"),
                        out.tolist()))

    def device_bytes(self):
        return lib().dlbc_gpu_device_bytes(self.h)

    @classmethod
    def from_params(cls, p, storage="f64", device=0, schedule=None, device_init=True):
        """Start-up sequence of src/main.d:76-147 for a parameter dictionary
        (same keys as the oracle's Oracle.from_params).  device_init: compute
        the initial mask and populations on the device (the same bits as the
        host initialisation and upload, which device_init=False uses)."""
        d, q, nc = p["d"], p["q"], p.get("components", 1)
        n = p["gn"]
        h = reference_halo(d, q)  # the draw order of the random initialisations
        L = cls(d, q, nc, n, storage=storage, device=device)
        L.set_params(p.get("tau", [1.0] * nc), p.get("gcc"), p.get("gwc"), p.get("globalAcc"),
                     p.get("bodyForce"), p.get("eqDistForm", "SecondOrder"),
                     p.get("psiForm", "Linear"), p.get("enableShanChen", False), p.get("rho0"))
        if schedule:
            L.set_schedule(schedule)
        if device_init:
            if "maskArray" in p:  # MaskInit.File: physical sites from the file
                L.upload_mask(np.asarray(p["maskArray"], np.int32))
            else:
                L.init_mask(p.get("maskInit", "None"), p.get("maskAxis", "X"),
                            p.get("wallOffset", 0), p.get("dimensionHasWall"),
                            p.get("wallThickness", 1))
            grng = Rng(p.get("seed", 0))
            for f in range(nc):
                rng = Rng(p["randomSeed"][f]) if p.get("randomReseed", [False] * nc)[f] else grng
                L.init_fluid(f, p["fluidInit"][f], p["fluidDensities"][f],
                             p.get("eqDistForm", "SecondOrder"),
                             p.get("fluidPerturb", [0.0] * nc)[f], rng, p.get("initRadius", 0.0),
                             p.get("initOffset"), p.get("initSeparation"),
                             p.get("initAxis", "X"), p.get("interfaceThickness", 0.0),
                             p.get("lamellaeWidths"), halo=h)
                L.init_eq_dist_wall(f)
            return L
        mask = init_mask(d, n, p.get("maskInit", "None"), p.get("maskAxis", "X"),
                         p.get("wallOffset", 0), p.get("dimensionHasWall"),
                         p.get("wallThickness", 1), halo=h)
        if "maskArray" in p:  # MaskInit.File: physical sites from the file, halo None
            mask[...] = 0
            mask[tuple(slice(h, h + v) for v in n)] = np.asarray(p["maskArray"], np.int32)
        grng = Rng(p.get("seed", 0))  # initRNG on rank 0
        eqf = p.get("eqDistForm", "SecondOrder")
        pops = []
        for f in range(nc):
            rng = Rng(p["randomSeed"][f]) if p.get("randomReseed", [False] * nc)[f] else grng
            pop = init_fluid(d, q, n, p["fluidInit"][f], p["fluidDensities"][f], eqf,
                             p.get("fluidPerturb", [0.0] * nc)[f], rng, p.get("initRadius", 0.0),
                             p.get("initOffset"), p.get("initSeparation"), p.get("initAxis", "X"),
                             p.get("interfaceThickness", 0.0), p.get("lamellaeWidths"), halo=h)
            init_eq_dist_wall(pop, mask, d, q, n, halo=h)
            pops.append(pop)
        L.upload_mask(mask, halo=h)
        for f in range(nc):
            L.upload_population(f, pops[f], halo=h)
        return L


class Batch:
    """Independent D3Q19 two-component lattices (shifted storage, any size)
    advanced together: one persistent launch per step (dlbc_batch_*)."""

    def __init__(self, lattices):
        self.members = list(lattices)  # keep the contexts alive
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = C.c_void_p()
        rc = lib().dlbc_batch_create(C.byref(h), arr, len(self.members))
        if rc:
            raise DlbcError("dlbc_batch_create: %s" % ERRORS.get(rc, rc))
        self.h = h

    def _check(self, rc, what):
        if rc:
            raise DlbcError("%s: %s (%s)" % (what, ERRORS.get(rc, rc),
                                             lib().dlbc_batch_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            lib().dlbc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, tile_y, chunk, workgroups, kernel=0):
        """kernel: 0 / 1 the tiled persistent kernel (with / without solids),
        2 the split pair, 3 the split pair as one persistent launch."""
        self._check(lib().dlbc_batch_set_config(self.h, tile_y, chunk, workgroups, kernel),
                    "set_config")

    @property
    def config(self):
        out = (C.c_int * 4)()
        self._check(lib().dlbc_batch_get_config(self.h, out), "get_config")
        return tuple(out)

    def autotune(self):
        out = (C.c_int * 4)()
        ms = C.c_double()
        self._check(lib().dlbc_batch_autotune(self.h, out, C.byref(ms)), "autotune")
        return tuple(out), ms.value

    def step(self, nsteps=1, graph=True):
        self._check(lib().dlbc_batch_step(self.h, int(nsteps), int(bool(graph))), "step")

    def step_looped(self, nsteps=1):
        self._check(lib().dlbc_batch_step_looped(self.h, int(nsteps)), "step_looped")

    def synchronize(self):
        self._check(lib().dlbc_batch_synchronize(self.h), "synchronize")
