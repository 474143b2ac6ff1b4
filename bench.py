#!/usr/bin/env python3
"""Throughput of DLBC's lattice Boltzmann step on one MI355X.

Workload (default): D3Q19 two-component Shan-Chen channel, 256^3, solid walls
at z = 0 and 255 plus seeded random solid spheres (radius 4..12, solid
fraction 0.2) with bounce-back, body force 1e-6 along x on component 1,
red/blue halves (EqDistLamellae along x: red 1.0/0.05, blue 0.05/1.0),
tau 1, g = 3.6, psi = 1 - exp(-rho), BDist2 equilibrium, FP64.  (With the
SecondOrder or ThirdOrder equilibrium this workload diverges within 100
steps in the reference's own scheme; BDist2 is the form of all the
reference's Shan-Chen tests.)  The same workload is also timed with the
FP32- and FP16-shifted population storages.

Prints one JSON line: MLUPS (sites counted once, as the reference's LUPS),
the roofline of the dominant kernel (algorithmic bytes per launch over its
HIP-event-measured average launch time) and the CPU baseline (the oracle,
a C restatement of the reference's AoS passes, on the host cores for a
bounded number of steps of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
C, Q = 2, 19


def shift_rho0(storage, light=0.05, dense=1.0):
    """Reference density of the shifted storages for components with two bulk
    phases (light, dense).  FP16: the harmonic mean 2 ab / (a + b), at which
    the stored deviation |rho - rho0| is the same fraction of rho in both
    phases (0.905 for 0.05 / 1.0): with 11 bits per deviation the light
    phase's populations are what limits FP16, and the midpoint leaves their
    deviations at 9.5 times their density (measured on the 128^3 droplet
    after 1000 steps: density error 0.009 against 0.113, radius 0.4 %
    against 2.9 %).  FP32: the midpoint (a + b) / 2, the smallest deviations
    overall; its error is set by the dense phase's arithmetic (the 256^3
    channel after 2000 steps: largest velocity error 2.4e-5 of the largest
    speed with the midpoint, 1.1e-4 with the harmonic mean)."""
    if storage == "f16_shift":
        return 2.0 * light * dense / (light + dense)
    return 0.5 * (light + dense)


RHO0 = None  # --rho0: one reference density for every shifted leg (None: shift_rho0)


def rho0_for(storage):
    if RHO0 is not None:
        return list(RHO0)
    return [shift_rho0(storage)] * 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def channel_mask(n, fraction=0.2, seed=20261202):
    """Walls at z = 0, n-1 and random solid spheres (integer centres, radius
    U{4..12}, periodic in x and y) until the solid fraction is reached."""
    rng = np.random.default_rng(seed)
    m = np.zeros((n, n, n), np.int32)
    m[:, :, 0] = 1
    m[:, :, n - 1] = 1
    total = n ** 3
    solid = int(m.sum())
    while solid < fraction * total:
        r = int(rng.integers(4, 13))
        cx, cy, cz = (int(v) for v in rng.integers(0, n, 3))
        d = np.arange(-r, r + 1)
        dx, dy, dz = np.meshgrid(d, d, d, indexing="ij")
        ball = (dx ** 2 + dy ** 2 + dz ** 2 <= r * r).astype(np.int32)
        xs = (cx + d) % n
        ys = (cy + d) % n
        zs = cz + d
        zok = (zs >= 0) & (zs < n)
        idx = np.ix_(xs, ys, zs[zok])
        sub = m[idx]
        new = ball[:, :, zok] | sub
        solid += int(new.sum() - sub.sum())
        m[idx] = new
    return m


def channel_params(n):
    return dict(d=3, q=19, gn=[n, n, n], components=2, tau=[1.0, 1.0],
                eqDistForm="BDist2", gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True,
                psiForm="Exponential", bodyForce=[[0.0, 0.0, 0.0], [1e-6, 0.0, 0.0]],
                fluidInit=["EqDistLamellae"] * 2, fluidDensities=[[1.0, 0.05], [0.05, 1.0]],
                lamellaeWidths=[n // 2, n // 2], initAxis="X")


def build_inputs(D, p, mask):
    n = p["gn"]
    pops = []
    for f in range(2):
        pop = D.init_fluid(3, Q, n, p["fluidInit"][f], p["fluidDensities"][f], p["eqDistForm"],
                           0.0, None, 0.0, None, None, p["initAxis"], 0.0, p["lamellaeWidths"],
                           halo=0)
        D.init_eq_dist_wall(pop, mask, 3, Q, n, halo=0)
        pops.append(pop)
    return pops


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernels_channel256_traffic.txt")


def pmc_traffic(storage, kernel, n):
    """HBM bytes per launch of a step kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a
    copy kernel of known size); None where that summary has no figure."""
    if n != 256 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 7 and parts[0] == "traffic" and parts[1] == storage \
                    and parts[2] == kernel:
                return int(parts[6])
    return None


BLOCK_ORDER = -1  # --block-order: split-schedule XCD block order (-1: the library's choice)
MARCH = (0, -1)  # --march: rows per workgroup and hold of the march schedule
LAUNCH = 0  # --launch: minimum workgroups per fused / march launch (0: default)


def make_lattice(D, p, mask, pops, schedule, storage="f64", rho0=None):
    L = D.Lattice(3, Q, 2, p["gn"], storage=storage)
    L.set_params(p["tau"], p["gcc"], None, None, p["bodyForce"], p["eqDistForm"], p["psiForm"],
                 True, rho0)
    if schedule != "auto":
        L.set_schedule(schedule)
    if L.schedule == "march":
        L.set_march(*MARCH)
    L.set_launch(LAUNCH)
    L.set_block_order(BLOCK_ORDER)
    L.upload_mask(mask)
    for f in range(2):
        L.upload_population(f, pops[f])
    return L


def alg_bytes(fluid, sites, elem):
    """Populations read and written once per fluid site, one mask byte per site."""
    return fluid * 2 * C * Q * elem + sites


def kernel_bytes(name, fluid, sites, elem):
    """Algorithmic bytes of one launch of a step kernel: the populations it
    must read and write once, psi written by a psi pass (FP64, or FP32 for
    shifted storage), one mask byte per site for a kernel that streams."""
    pops = fluid * C * Q * elem
    if name.startswith("psi"):
        return pops + fluid * C * (8 if elem == 8 else 4)  # read populations, write psi
    return 2 * pops + sites


def time_steps(L, steps, warmup):
    """Wall time of `steps` steps bracketed by device synchronisations."""
    import torch
    L.step(warmup)
    L.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.step(steps)
    L.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def storage_leg(D, p, mask, pops, storage, steps, warmup, fluid):
    """MLUPS and roofline of the same workload with shifted storage."""
    rho0 = rho0_for(storage)
    L = make_lattice(D, p, mask, pops, "auto", storage, rho0)
    m0 = [L.total_mass(f) for f in range(2)]
    dt = time_steps(L, steps, warmup)
    sites = int(np.prod(p["gn"]))
    el, kern = L.step_timed(max(5, steps // 2))
    dom = max(kern, key=kern.get)
    elem = {"f32_shift": 4, "f16_shift": 2}[storage]
    b = alg_bytes(fluid, sites, elem)
    drift = max(abs(L.total_mass(f) - m0[f]) / m0[f] for f in range(2))
    out = {"value": round(sites * steps / dt / 1e6, 2), "unit": "MLUPS",
           "ms_per_step": round(1e3 * dt / steps, 4), "kernel": dom,
           "launch_ms": round(kern[dom], 4), "bytes_per_launch": b,
           "achieved": round(b / (kern[dom] * 1e-3) / 1e9, 1),
           "frac": round(b / (kern[dom] * 1e-3) / HBM_PEAK, 4),
           "traffic": pmc_traffic(storage, dom, p["gn"][0]), "rho0": rho0,
           "step_frac_per_site": round(sites * steps / dt * (2 * C * Q * elem + 1) / HBM_PEAK, 4),
           "mass_rel_drift": float("%.3g" % drift)}
    del L
    return out


def interface_displacement(phi, phi_ref, fluid):
    """Displacement of the colour interface (phi = 0) from its position in a
    reference field, in lattice units, to first order: at every fluid site
    where the reference colour changes sign towards a fluid face neighbour,
    |phi - phi_ref| over the reference colour gradient there (central
    differences, taken only where both neighbours along every axis are
    fluid).  Returns the largest and the mean displacement, the number of
    interface sites they cover, and the number of fluid sites whose colour
    has the other sign than in the reference."""
    iface = np.zeros(phi_ref.shape, bool)
    inner = fluid.copy()
    grad2 = np.zeros(phi_ref.shape)
    sref = np.sign(phi_ref)
    for ax in range(phi_ref.ndim):
        up, dn = np.roll(phi_ref, -1, ax), np.roll(phi_ref, 1, ax)
        fu, fd = np.roll(fluid, -1, ax), np.roll(fluid, 1, ax)
        inner &= fu & fd
        iface |= fluid & ((fu & (np.sign(up) != sref)) | (fd & (np.sign(dn) != sref)))
        grad2 += (0.5 * (up - dn)) ** 2
    sel = iface & inner
    disp = np.abs(phi[sel] - phi_ref[sel]) / np.maximum(np.sqrt(grad2[sel]), 1e-300)
    flips = int(np.count_nonzero(fluid & (np.sign(phi) != sref)))
    if disp.size == 0:
        return 0.0, 0.0, 0, flips
    return float(disp.max()), float(disp.mean()), int(disp.size), flips


def storage_accuracy(D, p, mask, pops, storages, steps):
    """Error of the shifted storages against the FP64 path on the same
    workload after `steps` steps: density, velocity (max norm relative to the
    largest FP64 speed, and relative 2-norm over all fluid sites), mass and
    the position of the red/blue interface (interface_displacement)."""
    import torch
    fluid = mask == 0

    def run(storage):
        L = make_lattice(D, p, mask, pops, "auto", storage,
                         None if storage == "f64" else rho0_for(storage))
        m0 = [L.total_mass(f) for f in range(2)]
        L.step(steps)
        out = {"rho": [L.density(f) for f in range(2)],
               "u": [L.velocity(f) for f in range(2)],
               "colour": L.colour(),
               "drift": max(abs(L.total_mass(f) - m0[f]) / m0[f] for f in range(2))}
        del L
        torch.cuda.empty_cache()
        return out

    ref = run("f64")
    res = {"steps": steps}
    for st in storages:
        r = run(st)
        du = [r["u"][f][fluid] - ref["u"][f][fluid] for f in range(2)]
        uref = [ref["u"][f][fluid] for f in range(2)]
        umax = max(float(np.abs(u).max()) for u in uref)
        res[st] = {
            "max_abs_drho": max(float(np.abs(r["rho"][f] - ref["rho"][f])[fluid].max())
                                for f in range(2)),
            "max_abs_du": max(float(np.abs(d).max()) for d in du),
            "max_u": umax,
            "max_abs_du_over_max_u": max(float(np.abs(d).max()) for d in du) / umax,
            "rel_l2_du": max(float(np.linalg.norm(du[f]) / np.linalg.norm(uref[f]))
                             for f in range(2)),
            "mass_rel_drift": r["drift"], "f64_mass_rel_drift": ref["drift"],
            "rho0": rho0_for(st)}
        dmax, dmean, nif, flips = interface_displacement(r["colour"], ref["colour"], fluid)
        res[st].update({"interface_disp_max": dmax, "interface_disp_mean": dmean,
                        "interface_sites": nif, "colour_sign_flips": flips})
        log("accuracy %s after %d steps: %s" % (st, steps, res[st]))
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, mask, pops, seconds=10.0, max_steps=60):
    """The oracle (C restatement of the reference's passes) on the host cores:
    whole steps of the workload until about `seconds` of wall time."""
    from oracle import oracle as OR
    OR.build()
    n = p["gn"]
    o = OR.Oracle(3, Q, 2, n)
    o.set_params(p["tau"], p["gcc"], None, None, 2, 1, True)
    o.set_body_force(p.get("bodyForce"))  # the reference's pre-collision hook plugin
    o.set_mask(mask)
    for f in range(2):
        o.set_population(f, pops[f])
    o.exchange_halo()
    o.run(1)  # warm-up (first touch of the AoS fields)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and (steps < 2 or time.perf_counter() - t0 < seconds):
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    threads = OR.lib().or_get_threads()
    # one step on one core, for the single-core rate
    OR.lib().or_set_threads(1)
    t1 = time.perf_counter()
    o.run(1)
    dt1 = time.perf_counter() - t1
    OR.lib().or_set_threads(threads)
    return {"value": round(float(np.prod(n)) * steps / dt / 1e6, 3), "unit": "MLUPS",
            "cores": threads, "kind": "port", "seconds": round(dt, 2), "cpu": cpu_model(),
            "nproc": os.cpu_count(),
            "single_core_value": round(float(np.prod(n)) / dt1 / 1e6, 3),
            "sample": "%d steps of the same %s workload (AoS + halo 2, reference pass "
                      "order, OpenMP), the body force added as the reference's "
                      "pre-collision hook" % (steps, "x".join(map(str, n)))}


def droplet_params(n, radius):
    """Centred droplet: red 1.0 inside / 0.05 outside, blue swapped (sharp
    interface), periodic, tau 1, g 3.6, psi = 1 - exp(-rho), BDist2."""
    return dict(d=3, q=19, gn=[n, n, n], components=2, tau=[1.0, 1.0], eqDistForm="BDist2",
                gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                fluidInit=["EqDistSphere"] * 2, fluidDensities=[[1.0, 0.05], [0.05, 1.0]],
                initRadius=float(radius))


def droplet_workload(D, n, steps, storages=("f16_shift", "f32_shift")):
    """Shifted storages against FP64 on the droplet: throughput of each, and
    after `steps` steps the largest density and velocity differences, the
    droplet radius of the Laplace diagnostics, the interface displacement
    (interface_displacement) and the mass drift."""
    import torch
    radius = 96.0 * n / 512.0
    p = droplet_params(n, radius)
    init_s = {}

    def init_args(f):
        return ("EqDistSphere", p["fluidDensities"][f], "BDist2", 0.0, None, radius)

    def run(storage):
        L = D.Lattice(3, Q, 2, p["gn"], storage=storage)
        L.set_params(p["tau"], p["gcc"], None, None, None, p["eqDistForm"], p["psiForm"], True,
                     rho0_for(storage))
        if storage == "f64":
            # the host initialisation and upload, timed once beside the device one
            t = time.perf_counter()
            L.upload_mask(np.zeros((n, n, n), np.int32))
            for f in range(2):
                L.upload_population(f, D.init_fluid(3, Q, p["gn"], *init_args(f), halo=0))
            L.synchronize()
            init_s["host_init_and_upload"] = round(time.perf_counter() - t, 2)
        t = time.perf_counter()
        L.init_mask("None")
        for f in range(2):
            L.init_fluid(f, *init_args(f), halo=0)
        L.synchronize()
        init_s["device_init_" + storage] = round(time.perf_counter() - t, 2)
        log("%s initialised: %s" % (storage, init_s))
        m0 = [L.total_mass(f) for f in range(2)]
        lap0 = L.laplace()
        L.synchronize()
        torch.cuda.synchronize()
        t = time.perf_counter()
        L.step(steps)
        L.synchronize()
        dt = time.perf_counter() - t
        # whole-step algorithmic bytes: populations read and written once per
        # site (no solids, so no mask byte), over the nominal HBM peak
        elem = {"f64": 8, "f32_shift": 4, "f16_shift": 2}[storage]
        step_GBps = n ** 3 * 2 * C * Q * elem * steps / dt / 1e9
        out = {"mlups": round(n ** 3 * steps / dt / 1e6, 1), "seconds": round(dt, 2),
               "step_alg_GBps": round(step_GBps, 1),
               "step_frac": round(step_GBps * 1e9 / HBM_PEAK, 4),
               "radius_t0": lap0["radius"], "laplace": L.laplace(),
               "mass_rel_drift": max(abs(L.total_mass(f) - m0[f]) / m0[f] for f in range(2))}
        log("%s: %s" % (storage, out))
        return L, out

    ref, res = run("f64")
    fields = {"rho": [ref.density(f) for f in range(2)], "u": [ref.velocity(f) for f in range(2)],
              "colour": ref.colour()}
    result = {"f64": res}
    del ref
    torch.cuda.empty_cache()
    for st in storages:
        L, res = run(st)
        res["max_abs_drho"] = max(float(np.abs(L.density(f) - fields["rho"][f]).max())
                                  for f in range(2))
        res["max_abs_du"] = max(float(np.abs(L.velocity(f) - fields["u"][f]).max())
                                for f in range(2))
        umax = max(float(np.abs(fields["u"][f]).max()) for f in range(2))
        res["max_abs_du_over_max_u"] = res["max_abs_du"] / umax
        res["radius_rel_diff"] = (res["laplace"]["radius"] / result["f64"]["laplace"]["radius"]
                                  - 1.0)
        dmax, dmean, nif, flips = interface_displacement(L.colour(), fields["colour"],
                                                         np.ones((n, n, n), bool))
        res.update({"interface_disp_max": dmax, "interface_disp_mean": dmean,
                    "interface_sites": nif, "colour_sign_flips": flips})
        result[st] = res
        del L
        torch.cuda.empty_cache()
    print(json.dumps({"workload": "D3Q19 two-component Shan-Chen droplet %d^3, radius %.0f, "
                                  "1.0/0.05, BDist2, tau 1, g 3.6, %d steps" % (n, radius, steps),
                      "init_seconds": init_s, "results": result}), flush=True)


def slab_workload(D, n, planes, steps, warmup):
    """One GPU's share of the channel when x is cut into slabs: the first
    `planes` x-planes of the n^3 workload, periodic in x (the wrap stands in
    for the face exchange). Throughput per storage, and the bytes a face
    exchange would move per step (the populations crossing the face plus one
    psi plane, both components)."""
    import torch
    p = channel_params(n)
    mask = channel_mask(n)
    pops = build_inputs(D, p, mask)
    mask = np.ascontiguousarray(mask[:planes])
    pops = [np.ascontiguousarray(pp[:planes]) for pp in pops]
    ps = dict(p, gn=[planes, n, n])
    sites = planes * n * n
    out = {}
    for st in ("f64", "f32_shift", "f16_shift"):
        L = make_lattice(D, ps, mask, pops, "auto", st, None if st == "f64" else rho0_for(st))
        dt = time_steps(L, steps, warmup)
        el, kern = L.step_timed(max(5, steps // 2))
        elem = {"f64": 8, "f32_shift": 4, "f16_shift": 2}[st]
        out[st] = {"mlups": round(sites * steps / dt / 1e6, 1),
                   "ms_per_step": round(1e3 * dt / steps, 4),
                   "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
                   "face_exchange_bytes": 2 * n * n * (5 * C * elem + C * 8)}
        log("slab %s: %s" % (st, out[st]))
        del L
        torch.cuda.empty_cache()
    print(json.dumps({"workload": "%dx%dx%d slab of the %d^3 channel (one GPU's share of %d)"
                                  % (planes, n, n, n, n // planes),
                      "steps": steps, "results": out}), flush=True)


def ensemble_params(count=256, seed=20261202):
    """Independent droplet lattices: cubic L ~ U{32..96}, tau ~ U[1.0, 1.5]
    (both components), g ~ U[3.0, 4.5], radius ~ U[0.2, 0.35] * L
    (EqDistSphereFrac, 0.7/0.04, interface width 5), BDist2."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(32, 97))
        tau = float(rng.uniform(1.0, 1.5))
        g = float(rng.uniform(3.0, 4.5))
        frac = float(rng.uniform(0.2, 0.35))
        out.append(dict(d=3, q=19, gn=[n, n, n], components=2, tau=[tau, tau],
                        eqDistForm="BDist2", gcc=[[0.0, g], [g, 0.0]], enableShanChen=True,
                        psiForm="Exponential", fluidInit=["EqDistSphereFrac"] * 2,
                        fluidDensities=[[0.7, 0.04], [0.04, 0.7]], initRadius=frac,
                        interfaceThickness=5.0, rho0=[0.37, 0.37]))
    return out


def ensemble_workload(D, count, steps, storage="f32_shift"):
    """The ensemble advanced by the autotuned batch kernel (graph replay
    where it applies), by one launch per lattice and step, and by the split
    pair as two launches and as one persistent launch per step; MLUPS counts
    sum(L^3) per step."""
    import torch
    params = ensemble_params(count)
    elem_bytes = {"f64": 8, "f32_shift": 4, "f16_shift": 2}[storage]
    t0 = time.perf_counter()
    lats = [D.Lattice.from_params(p, storage=storage) for p in params]
    sites = sum(int(np.prod(p["gn"])) for p in params)
    log("ensemble: %d lattices, %d sites, inputs %.1f s"
        % (count, sites, time.perf_counter() - t0))
    B = D.Batch(lats)
    cfg, ms = B.autotune()
    log("autotuned (tile rows, chunk, workgroups) = %s, %.3f ms per launch" % (cfg, ms))
    B.step(2)
    B.synchronize()

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        B.synchronize()
        return time.perf_counter() - t
    dt_graph = timed(lambda: B.step(steps, graph=True))
    dt_eager = timed(lambda: B.step(steps, graph=False))
    dt_loop = timed(lambda: B.step_looped(steps))
    # the split pair in both of its forms, whatever the autotuner picked: two
    # launches per step (graph replay) and one persistent cooperative launch
    forms = {}
    for kn, name in ((2, "split_pair_graph_mlups"), (3, "split_persistent_mlups")):
        B.set_config(cfg[0], cfg[1], cfg[2], kn)
        B.step(2)
        forms[name] = round(sites * steps / timed(lambda: B.step(steps, graph=True)) / 1e6, 1)
    B.set_config(*cfg)
    finite = all(np.isfinite(L.total_mass(0)) for L in lats)
    res = {"lattices": count, "sites": sites, "steps": steps, "storage": storage,
           "config": {"tile_rows": cfg[0], "chunk": cfg[1], "workgroups": cfg[2],
                      "kernel": ["tiled, solid-capable", "tiled, no-solid", "split pair",
                                 "split pair, one persistent launch"][cfg[3]]},
           "autotune_launch_ms": round(ms, 4),
           "batched_graph_mlups": round(sites * steps / dt_graph / 1e6, 1),
           # whole-step algorithmic bytes (populations read and written once per
           # site; these lattices have no solids) over the nominal HBM peak
           "batched_graph_step_frac": round(sites * steps * 2 * C * Q * elem_bytes / dt_graph
                                            / HBM_PEAK, 4),
           "batched_eager_mlups": round(sites * steps / dt_eager / 1e6, 1),
           "looped_mlups": round(sites * steps / dt_loop / 1e6, 1),
           "speedup_vs_looped": round(dt_loop / dt_graph, 3), **forms,
           "finite": bool(finite)}
    print(json.dumps({"workload": "ensemble of %d D3Q19 two-component droplet lattices, "
                                  "L ~ U{32..96}" % count, "results": res}), flush=True)


def small_workload(D, reps=3):
    """Small lattices, whose steps are bound by launch latency: eager steps
    against graph replay (dlbc_gpu_step_graph, same bits), best of `reps`
    timings each.  D2Q9 64^2 single-component BGK (rho 1 +- 1e-3, seed 42) and
    256^2; D3Q19 two-component Shan-Chen spinodal decomposition (rho 0.7 +-
    0.01, g 3.6, BDist2) at 32^3, 64^3 and 128^3, and at 128^3 with the
    SecondOrder equilibrium for 1000 steps."""
    def d2q9(n):
        return dict(d=2, q=9, gn=[n, n], components=1, tau=[1.0], fluidInit=["EqDistPerturb"],
                    fluidDensities=[[1.0]], fluidPerturb=[1e-3], randomReseed=[True],
                    randomSeed=[42])

    def spinodal(n, form="BDist2"):
        return dict(d=3, q=19, gn=[n, n, n], components=2, tau=[1.0, 1.0], eqDistForm=form,
                    gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True, psiForm="Exponential",
                    fluidInit=["EqDistPerturb"] * 2, fluidDensities=[[0.7], [0.7]],
                    fluidPerturb=[0.01, 0.01], randomReseed=[True, True], randomSeed=[42, 43])
    cases = [("D2Q9 64^2 BGK", d2q9(64), 1000), ("D2Q9 256^2 BGK", d2q9(256), 1000),
             ("D3Q19 32^3 two-fluid SC", spinodal(32), 1000),
             ("D3Q19 64^3 two-fluid SC", spinodal(64), 500),
             ("D3Q19 128^3 two-fluid SC", spinodal(128), 200),
             ("D3Q19 128^3 two-fluid SC, SecondOrder (the spinodal workload as specified)",
              spinodal(128, "SecondOrder"), 1000)]
    out = []
    for name, p, steps in cases:
        L = D.Lattice.from_params(p)
        sites = int(np.prod(p["gn"]))
        res = {"lattice": name, "steps": steps}
        for graph in (False, True):
            L.step(2 * 32, graph=graph)  # warm-up (and the captures)
            L.synchronize()
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                L.step(steps, graph=graph)
                L.synchronize()
                best = min(best, time.perf_counter() - t0)
            key = "graph" if graph else "eager"
            res[key + "_mlups"] = round(sites * steps / best / 1e6, 1)
            res[key + "_us_per_step"] = round(best / steps * 1e6, 2)
        out.append(res)
        log("%s: eager %.1f, graph %.1f MLUPS" % (name, res["eager_mlups"], res["graph_mlups"]))
        del L
    print(json.dumps({"workload": "small lattices, eager steps against graph replay",
                      "results": out}), flush=True)


def copy_bandwidth(gib=2.0, reps=5):
    """A plain streaming kernel of known traffic (torch.neg over FP64 of
    `gib` GiB: every byte read once and written once), for scaling the
    FETCH_SIZE / WRITE_SIZE counters and as the measured copy bandwidth."""
    import torch
    n = int(gib * (1 << 30)) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del x, y
    torch.cuda.empty_cache()
    return {"copy_kernel": "torch.neg fp64", "bytes_read": 8 * n, "bytes_written": 8 * n,
            "ms": round(dt * 1e3, 4), "GBps_read_plus_write": round(16 * n / dt / 1e9, 1)}


def copy_calibration():
    print(json.dumps(copy_bandwidth()), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--schedule", default="auto",
                    choices=["auto", "two_pass", "fused", "split", "march"])
    ap.add_argument("--march", default="0,-1",
                    help="march schedule launch shape: rows per workgroup, hold (0,-1: automatic)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rho0", type=float, default=None,
                    help="reference density of every shifted storage (default: FP16 the "
                         "harmonic mean of the two bulk densities, %.4f; FP32 their midpoint, "
                         "%.3f)" % (shift_rho0("f16_shift"), shift_rho0("f32_shift")))
    ap.add_argument("--launch", type=int, default=0,
                    help="minimum workgroups per launch of the fused and march kernels "
                         "(0: the library's default)")
    ap.add_argument("--block-order", type=int, default=-1,
                    help="split schedule: blocks per XCD turn (0 site order, -1 automatic)")
    ap.add_argument("--no-storage-legs", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=2000,
                    help="steps after which the shifted storages are compared with FP64 "
                         "(0: no comparison)")
    ap.add_argument("--workload", default="channel",
                    choices=["channel", "droplet", "ensemble", "slab", "small"],
                    help="channel: the MLUPS line; droplet: shifted-storage errors against "
                         "FP64 on a centred droplet (radius 96 at 512^3, scaled with --size); "
                         "ensemble: 256 uneven lattices, batched vs looped; slab: one GPU's "
                         "x-slab of the channel; small: small lattices, eager steps against "
                         "graph replay")
    ap.add_argument("--lattices", type=int, default=256, help="ensemble size")
    ap.add_argument("--slab-planes", type=int, default=32,
                    help="slab workload: x-planes of one GPU's share")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain FP64 streaming kernel of known size and exit")
    args = ap.parse_args()
    global BLOCK_ORDER, MARCH, LAUNCH, RHO0
    if args.rho0 is not None:
        RHO0 = [args.rho0, args.rho0]
    BLOCK_ORDER = args.block_order
    LAUNCH = args.launch
    MARCH = tuple(int(v) for v in args.march.split(","))
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import dlbc_amd as D
    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "droplet":
        droplet_workload(D, args.size, args.steps)
        return
    if args.workload == "ensemble":
        ensemble_workload(D, args.lattices, args.steps)
        return
    if args.workload == "slab":
        slab_workload(D, args.size, args.slab_planes, args.steps, args.warmup)
        return
    if args.workload == "small":
        small_workload(D)
        return
    n = args.size
    p = channel_params(n)
    t0 = time.perf_counter()
    mask = channel_mask(n)
    pops = build_inputs(D, p, mask)
    log("inputs %.1f s, solid fraction %.4f" % (time.perf_counter() - t0, mask.mean()))
    L = make_lattice(D, p, mask, pops, args.schedule)
    log("uploaded; schedule %s; block order %d; device bytes %.2f GB"
        % (L.schedule, L.block_order, L.device_bytes() / 1e9))
    m0 = [L.total_mass(f) for f in range(2)]
    dt = time_steps(L, args.steps, args.warmup)
    sites = n ** 3
    mlups = sites * args.steps / dt / 1e6
    log("timed %d steps: %.3f ms/step, %.1f MLUPS" % (args.steps, 1e3 * dt / args.steps, mlups))
    # per-kernel launch durations (HIP events on the library stream)
    el, kern = L.step_timed(max(5, args.steps // 2))
    dom = max(kern, key=kern.get)
    fluid = int(sites - mask.sum())
    b_alg = kernel_bytes(dom, fluid, sites, 8)
    achieved = b_alg / (kern[dom] * 1e-3)
    per_kernel = {k: {"launch_ms": round(v, 4), "bytes": kernel_bytes(k, fluid, sites, 8),
                      "GBps": round(kernel_bytes(k, fluid, sites, 8) / (v * 1e-3) / 1e9, 1)}
                  for k, v in kern.items()}
    log("kernels (ms/launch): %s" % kern)
    m1 = [L.total_mass(f) for f in range(2)]
    drift = max(abs(m1[f] - m0[f]) / m0[f] for f in range(2))
    if not all(np.isfinite(m1)):
        sys.exit("bench.py: the state is not finite after the timed steps")
    result = {
        "metric": "MLUPS + % HBM peak, D3Q19 2-fluid SC 256\u00b3: FP64 vs FP32/FP16-shifted "
                  "storage",
        "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "D3Q19 two-component Shan-Chen channel %d^3, walls z=0,%d + "
                               "random spheres r 4-12 (solid fraction %.3f), red/blue halves "
                               "1.0/0.05, body force 1e-6 x on component 1, BDist2, tau 1, "
                               "g 3.6, FP64" % (n, n - 1, float(mask.mean())),
                   "schedule": L.schedule, "block_order": L.block_order, "sites": sites,
                   "fluid_sites": fluid,
                   "mass_rel_drift": float("%.3g" % drift)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": pmc_traffic("f64", dom, n),
                     "bytes_per_launch": b_alg, "launch_ms": round(kern[dom], 4),
                     "step_frac": round(mlups * 1e6 * (alg_bytes(fluid, sites, 8) / sites)
                                        / HBM_PEAK, 4),
                     # the metric's own "% HBM peak": MLUPS x 609 B per site
                     # (2 C Q 8 B + the mask byte, every site counted)
                     "step_frac_per_site": round(mlups * 1e6 * (2 * C * Q * 8 + 1) / HBM_PEAK, 4),
                     "kernels": per_kernel},
    }
    del L
    torch.cuda.empty_cache()
    # the practical ceiling of a streaming kernel on this GPU, beside the nominal peak
    cb = copy_bandwidth()
    result["roofline"]["copy_GBps"] = cb["GBps_read_plus_write"]
    result["roofline"]["frac_of_copy"] = round(achieved / 1e9 / cb["GBps_read_plus_write"], 4)
    if not args.no_storage_legs:
        result["storage_legs"] = {}
        for st in ("f32_shift", "f16_shift"):
            leg = storage_leg(D, p, mask, pops, st, args.steps, args.warmup, fluid)
            log("%s: %s" % (st, leg))
            result["storage_legs"][st] = leg
    if not args.no_storage_legs and args.accuracy_steps > 0:
        result["storage_accuracy"] = storage_accuracy(D, p, mask, pops, ("f32_shift", "f16_shift"),
                                                      args.accuracy_steps)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(p, mask, pops)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
