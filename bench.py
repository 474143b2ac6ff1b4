#!/usr/bin/env python3
"""Throughput of DLBC's lattice Boltzmann step on one MI355X.

Workload (default): D3Q19 two-component Shan-Chen channel, 256^3, solid walls
at z = 0 and 255 plus seeded random solid spheres (radius 4..12, solid
fraction 0.2) with bounce-back, body force 1e-6 along x on component 1,
red/blue halves (EqDistLamellae along x: red 1.0/0.05, blue 0.05/1.0),
tau 1, g = 3.6, psi = 1 - exp(-rho), SecondOrder equilibrium, FP64.

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
                eqDistForm="SecondOrder", gcc=[[0.0, 3.6], [3.6, 0.0]], enableShanChen=True,
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


def make_lattice(D, p, mask, pops, schedule):
    L = D.Lattice(3, Q, 2, p["gn"])
    L.set_params(p["tau"], p["gcc"], None, None, p["bodyForce"], p["eqDistForm"], p["psiForm"],
                 True)
    if schedule != "auto":
        L.set_schedule(schedule)
    L.upload_mask(mask)
    for f in range(2):
        L.upload_population(f, pops[f])
    return L


def cpu_baseline(p, mask, pops, steps=2):
    """The oracle (C restatement of the reference's passes) on the host cores."""
    from oracle import oracle as OR
    OR.build()
    n = p["gn"]
    o = OR.Oracle(3, Q, 2, n)
    o.set_params(p["tau"], p["gcc"], None, None, 0, 1, True)
    o.set_mask(mask)
    for f in range(2):
        o.set_population(f, pops[f])
    o.exchange_halo()
    o.run(1)  # warm-up (first touch of the AoS fields)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(float(np.prod(n)) * steps / dt / 1e6, 3), "unit": "MLUPS",
            "cores": threads, "kind": "port",
            "sample": "%d steps of the same %s workload (AoS + halo 2, reference pass "
                      "order, OpenMP); the body force, which the reference adds from a "
                      "plugin hook, is left out of this leg" % (steps, "x".join(map(str, n)))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--schedule", default="auto", choices=["auto", "two_pass", "fused"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import dlbc_amd as D
    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    n = args.size
    p = channel_params(n)
    t0 = time.perf_counter()
    mask = channel_mask(n)
    pops = build_inputs(D, p, mask)
    log("inputs %.1f s, solid fraction %.4f" % (time.perf_counter() - t0, mask.mean()))
    L = make_lattice(D, p, mask, pops, args.schedule)
    log("uploaded; schedule %s; device bytes %.2f GB" % (L.schedule, L.device_bytes() / 1e9))
    L.step(args.warmup)
    L.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.step(args.steps)
    L.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sites = n ** 3
    mlups = sites * args.steps / dt / 1e6
    log("timed %d steps: %.3f ms/step, %.1f MLUPS" % (args.steps, 1e3 * dt / args.steps, mlups))
    # per-kernel launch durations (HIP events on the library stream)
    el, kern = L.step_timed(max(5, args.steps // 2))
    dom = max(kern, key=kern.get)
    fluid = int(sites - mask.sum())
    b_alg = fluid * 2 * C * Q * 8 + sites * 1  # populations read+written once, mask byte
    achieved = b_alg / (kern[dom] * 1e-3)
    log("kernels (ms/launch): %s" % kern)
    result = {
        "metric": "MLUPS, D3Q19 2-fluid SC 256^3 (FP64 plain storage)",
        "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "D3Q19 two-component Shan-Chen channel %d^3, walls z=0,%d + "
                               "random spheres r 4-12 (solid fraction %.3f), body force 1e-6 x "
                               "on component 1, SecondOrder, tau 1, g 3.6, FP64"
                               % (n, n - 1, float(mask.mean())),
                   "schedule": L.schedule, "sites": sites, "fluid_sites": fluid},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "bytes_per_launch": b_alg, "launch_ms": round(kern[dom], 4),
                     "step_frac": round(mlups * 1e6 * (b_alg / sites) / HBM_PEAK, 4)},
    }
    if not args.no_cpu_baseline:
        del L
        torch.cuda.empty_cache()
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(p, mask, pops)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
