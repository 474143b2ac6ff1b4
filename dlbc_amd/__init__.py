"""dlbc_amd -- DLBC's lattice Boltzmann time step on AMD Instinct MI355X.

The compute path is libdlbc_hip.so (hand-written HIP kernels for gfx950,
C ABI in include/dlbc_hip.h).  This module is the Python front end used by
tests and the benchmark; it mirrors the reference's start-up sequence and
time loop (src/main.d:76-200):

    L = Lattice.from_params(params)   # initRNG, initMask, initFluids
    L.step(n)                         # runTimeloop body x n
    L.population(f), L.density(f), L.velocity(f), L.colour(), L.total_mass(f)

There is no CPU fallback: every call goes through the HIP library, and a
missing library raises.
"""
from .lattice import (  # noqa: F401
    Batch, DlbcError, Lattice, Rng, init_fluid, init_mask, init_eq_dist_wall, lib, lib_path,
    EQ_FORMS, PSI_FORMS, FLUID_INIT, MASK_INIT, AXIS, STORAGE, SCHEDULES)
