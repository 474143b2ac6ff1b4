"""Host-side pieces of bench.py (no GPU): the one-GPU refusal, the synthetic
inputs of the workloads, the algorithmic byte counts and the committed PMC
traffic figures it reports beside them."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_channel_mask_walls_spheres_and_fraction():
    n = 48
    m = bench.channel_mask(n)
    assert m.shape == (n, n, n) and m.dtype == np.int32
    assert m[:, :, 0].all() and m[:, :, n - 1].all()  # the two walls
    frac = m.mean()
    assert 0.2 <= frac < 0.2 + 4 * np.pi * 12 ** 3 / 3 / n ** 3  # stops at the first sphere past 0.2
    assert np.array_equal(m, bench.channel_mask(n))  # seeded


def test_algorithmic_bytes():
    fluid, sites = 1000, 1200
    # two components, 19 populations, read and write, plus a mask byte per site
    assert bench.alg_bytes(fluid, sites, 8) == 1000 * 2 * 2 * 19 * 8 + 1200
    assert bench.kernel_bytes("stream_force_collide_split", fluid, sites, 4) == \
        1000 * 2 * 2 * 19 * 4 + 1200
    # a psi pass reads the populations once and writes psi (FP64, or FP32 when shifted)
    assert bench.kernel_bytes("psi_split", fluid, sites, 8) == 1000 * 2 * 19 * 8 + 1000 * 2 * 8
    assert bench.kernel_bytes("psi_split_shifted", fluid, sites, 2) == \
        1000 * 2 * 19 * 2 + 1000 * 2 * 4


def test_pmc_traffic_figures_are_committed_for_the_channel():
    for storage, kernel, elem in [("f64", "stream_force_collide_split", 8),
                                  ("f64", "psi_split", 8),
                                  ("f32_shift", "stream_force_collide_split_shifted", 4),
                                  ("f16_shift", "stream_force_collide_split_shifted", 2)]:
        t = bench.pmc_traffic(storage, kernel, 256)
        assert isinstance(t, int) and t > 0
        # measured bytes lie above the algorithmic ones (line waste at solids,
        # neighbour psi) but within a factor of two
        fluid = 13418099  # fluid sites of the 256^3 channel mask
        alg = bench.kernel_bytes(kernel, fluid, 256 ** 3, elem)
        assert 1.0 < t / alg < 2.0, (storage, kernel, t / alg)
    assert bench.pmc_traffic("f64", "stream_force_collide_split", 128) is None


def test_channel_mask_fluid_sites_256():
    m = bench.channel_mask(256)
    assert int((m == 0).sum()) == 13418099


def test_ensemble_parameters_in_range_and_seeded():
    ps = bench.ensemble_params(64)
    assert ps == bench.ensemble_params(64)
    for p in ps:
        n = p["gn"][0]
        assert p["gn"] == [n, n, n] and 32 <= n <= 96
        assert 1.0 <= p["tau"][0] <= 1.5 and p["tau"][0] == p["tau"][1]
        g = p["gcc"][0][1]
        assert 3.0 <= g <= 4.5 and p["gcc"][1][0] == g
        assert 0.2 <= p["initRadius"] <= 0.35


def test_interface_displacement_of_a_shifted_profile():
    """A colour profile cos(k (x - a)) shifted by delta: the first-order
    displacement at the interface sites is delta k / sin(k) (the central
    difference's factor), the two crossings give 2 x 2 x 64 interface sites,
    and a shift by one site turns one plane per crossing to the other sign."""
    n = 64
    x = np.arange(n, dtype=float)[:, None, None] * np.ones((1, 8, 8))
    k = 2 * np.pi / n
    ref = np.cos(k * (x - 0.3))
    fluid = np.ones(ref.shape, bool)
    delta = 1e-3
    dmax, dmean, nif, flips = bench.interface_displacement(np.cos(k * (x - 0.3 - delta)), ref,
                                                           fluid)
    assert nif == 2 * 2 * 64 and flips == 0
    want = delta * k / np.sin(k)
    assert abs(dmax - want) < 0.01 * want and abs(dmean - want) < 0.01 * want
    *_, flips = bench.interface_displacement(np.cos(k * (x - 1.3)), ref, fluid)
    assert flips == 2 * 64
    # a solid plane at x = 17 hides the first crossing (16.3): only the
    # second one's sites count
    fluid[17] = False
    dmax, _, nif, flips = bench.interface_displacement(ref, ref, fluid)
    assert nif == 2 * 64 and dmax == 0.0 and flips == 0
