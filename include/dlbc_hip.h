/*
 * dlbc_hip.h -- C ABI of the MI355X lattice Boltzmann step library.
 *
 * The library advances DLBC's per-timestep lattice Boltzmann path on one
 * AMD Instinct MI355X (gfx950): pull streaming with mask bounce-back, the
 * Shan-Chen psi/force with its velocity shift, and BGK collision with the
 * SecondOrder, ThirdOrder and BDist2 equilibria.  It replaces the body of
 * the reference's time loop (src/main.d:174-200):
 *
 *   L.exchangeHalo();      src/dlbc/lattice.d:200-214 -> periodic wrap in-kernel
 *   L.advectFields();      src/dlbc/lb/advection.d:35-89
 *   L.resetForce();        src/dlbc/lb/force.d:165-170
 *   L.addShanChenForce();  src/dlbc/lb/force.d:223-309
 *   L.distributeForce();   src/dlbc/lb/force.d:179-197
 *   L.prepareToCollide();  src/dlbc/lb/collision.d:41-51,156-196
 *   L.collideFields(tau, globalAcc);  src/dlbc/lb/collision.d:61-119
 *
 * with dlbc_gpu_step().  Host arrays use the reference's own layout: row-major
 * [n0+2h][n1+2h][n2+2h][k] with the component index innermost (halo h = 2 in
 * the reference, src/dlbc/lattice.d:96-103; h = 0 for physical-only arrays).
 *
 * Conventions
 *  - Return value 0 is success; a negative DLBC_ERR_* code is a failure whose
 *    message dlbc_gpu_last_error() returns.  The library never exits.
 *  - The caller owns every host buffer and no host pointer is kept after a
 *    call returns.  The library owns its device buffers and one HIP stream.
 *  - A context is used from one thread at a time.  dlbc_gpu_step() is
 *    stream-ordered and returns before the GPU finishes; every download
 *    synchronises.
 */
#ifndef DLBC_HIP_H
#define DLBC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dlbc_gpu_ctx dlbc_gpu_ctx;

enum {
  DLBC_OK = 0,
  DLBC_ERR_ARG = -1,         /* invalid argument */
  DLBC_ERR_HIP = -2,         /* HIP runtime error */
  DLBC_ERR_NOMEM = -3,       /* device or host allocation failed */
  DLBC_ERR_UNSUPPORTED = -4, /* configuration the library does not provide */
  DLBC_ERR_STATE = -5        /* call out of order (e.g. step before upload) */
};

/* Population storage on the device. */
enum {
  DLBC_STORAGE_F64 = 0,       /* plain FP64 populations                       */
  DLBC_STORAGE_F32_SHIFT = 1, /* f_i - w_i rho0 in FP32, FP32 arithmetic      */
  DLBC_STORAGE_F16_SHIFT = 2  /* f_i - w_i rho0 in FP16, FP32 arithmetic      */
};

/* EqDistForm (src/dlbc/lb/eqdist.d:23-36) and PsiForm (src/dlbc/lb/force.d:48-57). */
enum { DLBC_EQ_SECOND_ORDER = 0, DLBC_EQ_THIRD_ORDER = 1, DLBC_EQ_BDIST2 = 2 };
enum { DLBC_PSI_LINEAR = 0, DLBC_PSI_EXPONENTIAL = 1 };

/* Kernel schedule selection for dlbc_gpu_set_schedule(). */
enum {
  DLBC_SCHED_AUTO = 0,     /* fastest available for the configuration        */
  DLBC_SCHED_TWO_PASS = 1, /* psi pass + stream/force/collide pass (any DdQq)  */
  DLBC_SCHED_FUSED = 2,    /* single-pass D3Q19 kernel with an LDS psi ring   */
  DLBC_SCHED_SPLIT = 3,    /* D3Q19 two-component two-pass, lane per component */
  DLBC_SCHED_MARCH = 4     /* D3Q19 two-component FP64 single pass over whole z-rows */
};

/* Create a context for one lattice on HIP device `device`.
 * d, q: connectivity (3,19), (2,9), (1,3) or (1,5) -- the reference's
 * population connectivities d3q19, d2q9, d1q3 and d1q5
 * (src/dlbc/lb/connectivity.d:21-32).
 * components: number of fluid components C (1..4), lb.components.
 * n: physical lattice lengths n[0..d) (lattice.gn on one rank), each >= 2,
 * and >= 4 for D1Q5, whose reference halo is 4 (src/dlbc/lattice.d:96-103).
 * Replaces the Lattice constructor and field allocation
 * (src/dlbc/lattice.d:166-181, src/dlbc/lb/init.d:60-73, lb/force.d:67-107). */
int dlbc_gpu_create(dlbc_gpu_ctx **out, int d, int q, int components, const int64_t *n,
                    int storage, int device);

/* Physical parameters (src/dlbc/lb/lb.d:39, lb/force.d:26-43, lb/eqdist.d:18).
 * tau[C]       lb.tau
 * gcc[C*C]     lb.force.gcc, row-major (may be NULL when SC is off)
 * gwc[C]       lb.force.gwc (may be NULL: zero)
 * global_acc[d] lb.force.globalAcc (may be NULL: zero)
 * body_force[C*d] per-component force density added to F_f before the
 *              velocity shift (no reference counterpart: the reference adds
 *              such a force from a preCollisionHooks plugin); NULL for none
 * rho0[C]      reference densities of the shifted storages (ignored for F64).
 *              NULL keeps the current values (1.0 until first given).  Once a
 *              component's populations are on the device its rho0 is part of
 *              their encoding: a different value is refused with
 *              DLBC_ERR_STATE (set rho0 before upload or device init).
 * May be called again between steps to change tau, g, forces or forms. */
int dlbc_gpu_set_params(dlbc_gpu_ctx *ctx, const double *tau, const double *gcc,
                        const double *gwc, const double *global_acc,
                        const double *body_force, int eq_dist_form, int psi_form,
                        int enable_shan_chen, const double *rho0);

/* Upload component f from a host array [n0+2h]..[q] (halo h >= 0; only the
 * physical sites are read).  Equivalent of initFluids / readCheckpoint
 * followed by exchangeHalo (src/main.d:136-147).  Until a component is
 * uploaded or initialised its populations hold D's default double, the quiet
 * NaN 0x7FFC000000000000, as the reference's fields do (FluidInit.None). */
int dlbc_gpu_upload_populations(dlbc_gpu_ctx *ctx, int f, const double *host, int halo);

/* Upload the mask [n0+2h].. int32, 0 = Mask.None (fluid), 1 = Mask.Solid
 * (src/dlbc/lb/mask.d:84-93).  Solid sites keep their populations forever
 * (advection copies them, collision skips them), also when the mask changes
 * between steps. */
int dlbc_gpu_upload_mask(dlbc_gpu_ctx *ctx, const int32_t *host, int halo);

/* Download the mask into a host array [n0+2h].. int32 (physical sites
 * written, halo untouched): the mask field the reference writes as output and
 * checkpoint (src/dlbc/lb/io.d:56-101 maskFreq, src/dlbc/io/checkpoint.d). */
int dlbc_gpu_download_mask(dlbc_gpu_ctx *ctx, int32_t *host, int halo);

/* Advance nsteps time steps: nsteps x the runTimeloop body
 * (src/main.d:174-200).  The state afterwards is post-collision, as the
 * reference's at the end of a step. */
int dlbc_gpu_step(dlbc_gpu_ctx *ctx, int nsteps);

/* As dlbc_gpu_step, with the steps replayed from a captured HIP graph of 32
 * steps (one graph per buffer parity; the remainder of nsteps runs eagerly).
 * Same results bit for bit; for small lattices, whose steps are bound by
 * launch latency (D2Q9 64x64: 3.9 instead of 4.7 us per step).  The graphs are
 * captured on first use and again when the schedule, launch shape, block
 * order, parameters or solid state changed since.
 * Replaces the same runTimeloop body as dlbc_gpu_step (src/main.d:174-200). */
int dlbc_gpu_step_graph(dlbc_gpu_ctx *ctx, int nsteps);

/* The streaming phase alone: advectFields (src/dlbc/lb/advection.d:35-89),
 * pull streaming with mask bounce-back, after the periodic halo exchange.
 * For callers that run the reference's phases individually (e.g. with
 * their own hooks between them); dlbc_gpu_step already includes it. */
int dlbc_gpu_advect(dlbc_gpu_ctx *ctx);

/* As dlbc_gpu_step, and report the elapsed GPU time of the nsteps (HIP
 * events on the library stream) and, per kernel of the schedule, the average
 * duration of one launch (kernel_ms[i], i < *nkernels, at most 8 entries;
 * events bracket every launch, so this mode adds a little overhead). */
int dlbc_gpu_step_timed(dlbc_gpu_ctx *ctx, int nsteps, double *elapsed_ms, double *kernel_ms,
                        int *nkernels);
/* Name of kernel slot i of the current schedule (static string). */
const char *dlbc_gpu_kernel_name(const dlbc_gpu_ctx *ctx, int i);

int dlbc_gpu_synchronize(dlbc_gpu_ctx *ctx);
int dlbc_gpu_set_schedule(dlbc_gpu_ctx *ctx, int schedule);
int dlbc_gpu_get_schedule(const dlbc_gpu_ctx *ctx);
/* Launch size of the fused kernels: the x axis is cut into chunks until a
 * launch holds at least min_workgroups workgroups of 8 x 32 sites (0: the
 * default, one workgroup per CU for FP64 storage and two for the shifted
 * storages).  No counterpart in the reference (its loops are per rank). */
int dlbc_gpu_set_launch(dlbc_gpu_ctx *ctx, int min_workgroups);
/* Block order of the split schedule: chunk K > 0 gives the GPU's XCDs K
 * consecutive blocks of sites each in turn (block b runs on XCD b mod 8), so
 * neighbouring rows share an XCD's L2; 0 keeps site order; -1 lets the
 * library choose.  The results do not depend on it.  No counterpart in the
 * reference. */
int dlbc_gpu_set_block_order(dlbc_gpu_ctx *ctx, int chunk);
/* Launch shape of the march schedule: rows = z-rows per workgroup (1, or 2
 * for n_z <= 256 and n_y even; 0 the library's choice); hold = 1 keeps each
 * site's pulled populations in registers from its psi pull to its collision
 * (one row per workgroup), 0 pulls them again for the collision (-1 the
 * library's choice).  Results do not depend on it.  No counterpart in the
 * reference. */
int dlbc_gpu_set_march(dlbc_gpu_ctx *ctx, int rows, int hold);
int dlbc_gpu_get_block_order(const dlbc_gpu_ctx *ctx);

/* Download component f into a host array [n0+2h]..[q]: physical sites are
 * written, halo sites are left untouched (the caller's exchangeHalo fills
 * them).  For shifted storage the populations are w_i rho0 + stored deviation. */
int dlbc_gpu_download_populations(dlbc_gpu_ctx *ctx, int f, double *host, int halo);

/* Output fields of src/dlbc/lb/io.d:56-101 into host arrays [n0+2h].. :
 * density (densityField, lb/density.d:89-98; all sites incl. solid),
 * raw velocity [..][d] (rawVelocityField, lb/velocity.d:106-117; 0 on solid),
 * colour rho_f1 - rho_f2 (colourField, lb/density.d:284-295; 0 on solid). */
int dlbc_gpu_density(dlbc_gpu_ctx *ctx, int f, double *host, int halo);
int dlbc_gpu_velocity(dlbc_gpu_ctx *ctx, int f, double *host, int halo);
int dlbc_gpu_colour(dlbc_gpu_ctx *ctx, int f1, int f2, double *host, int halo);

/* Reductions over fluid sites (localTotalMass, lb/density.d:136-147;
 * localTotalMomentum, lb/momentum.d:131-147).  Summed in a fixed order in
 * double-double (compensated) arithmetic; out[0] is the sum rounded to
 * double, out[1] (mass) / out[d..2d) (momentum) the low-order part. */
int dlbc_gpu_total_mass(dlbc_gpu_ctx *ctx, int f, double *out /*2*/);
int dlbc_gpu_total_momentum(dlbc_gpu_ctx *ctx, int f, double *out /*2d*/);

/* dlbc_gpu_total_mass without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (stepping
 * loops that read the mass every few steps, harnesses that forbid hipMalloc).
 * Same kernels and the same bits as dlbc_gpu_total_mass, but the block
 * partials live in one fixed 64 KiB array per GPU (held until the process
 * ends) instead of context scratch, so it serves lattices of at most
 * 4095 * 65536 = 268,369,920 sites and returns DLBC_ERR_UNSUPPORTED above
 * that.  Every context on a GPU shares that GPU's array: calls that overlap
 * on one GPU (from several threads or streams) race.  Like
 * dlbc_gpu_total_mass it waits for the context's stream. */
int dlbc_gpu_total_mass_noalloc(dlbc_gpu_ctx *ctx, int f, double *out /*2*/);

/* Laplace diagnostics of a centred two-component droplet, as dumpLaplace
 * (src/dlbc/lb/laplace.d:55-218): out[0] sigma, out[1] measured radius,
 * out[2] pressure inside, out[3] pressure outside (volume-averaged densities
 * of the 3^d cubes at int(n/2) and at 1, pressurePsi of lb/density.d:316-326),
 * out[4] droplet deformation for d = 2 (NaN otherwise).  The reference's
 * relative change of sigma is left to the caller. */
int dlbc_gpu_laplace(dlbc_gpu_ctx *ctx, double *out /*5*/);

/* Evaluate psi (form) on n densities on the device (parity tests of the
 * correctly rounded exp).  x and y are host arrays. */
int dlbc_gpu_eval_psi(dlbc_gpu_ctx *ctx, int psi_form, const double *x, double *y, int64_t n);

/* Self-check of the exact quotient by the constant c_s^2 products that the
 * equilibria use in place of IEEE division (same code on host and device):
 * compares it with IEEE division on n pseudo-random operands, half of them
 * next to exact multiples of a divisor, and returns the number of mismatches
 * (0 expected).  Host only, no GPU needed. */
int64_t dlbc_selftest_division(int64_t n, uint64_t seed);

/* Self-check of the site-coordinate division the kernels use (a multiply and
 * a shift by per-lattice constants in place of 32-bit integer division):
 * compares it with integer division on n pseudo-random (dividend, divisor)
 * pairs over the index range of a lattice, half of them next to exact
 * multiples, and returns the number of mismatches (0 expected).  Host only,
 * no GPU needed; no counterpart in the reference. */
int64_t dlbc_selftest_site_division(int64_t n, uint64_t seed);

const char *dlbc_gpu_last_error(const dlbc_gpu_ctx *ctx);
void dlbc_gpu_destroy(dlbc_gpu_ctx *ctx);

/* Bytes of device memory held by the context. */
int64_t dlbc_gpu_device_bytes(const dlbc_gpu_ctx *ctx);

/* ---------------------------------------------------------------------
 * Host-side initial conditions (no GPU needed): the reference's init shapes
 * (src/dlbc/fields/init.d, src/dlbc/lb/init.d:263-339) written into a
 * halo-padded host array [n0+2h]..[q] exactly as the reference fills its
 * fields, including one random draw per stored halo-padded site in row-major
 * order for the seeded kinds (std.random.Mt19937, uniform!"[)").
 * --------------------------------------------------------------------- */

/* ------------------------------------------------------------------------
 * Batches: many independent lattices advanced by one persistent launch per
 * step.  No counterpart in the reference (one lattice per run); each member
 * is an ordinary context (its own parameters, uploads, fields, reductions)
 * created with DLBC_STORAGE_F32_SHIFT or F16_SHIFT, D3Q19, two components, of
 * any size n >= 4.  The caller keeps the members alive while the batch
 * exists.  Results are the same bits as stepping each member alone.
 * ------------------------------------------------------------------------ */
typedef struct dlbc_batch dlbc_batch;

int dlbc_batch_create(dlbc_batch **out, dlbc_gpu_ctx *const *lattices, int n);
/* Tile height (4 or 8 rows of 32 sites), x-chunk length (planes), number of
 * workgroups of the persistent launch and kernel: 0 the tiled kernel
 * (solid-capable, 128 VGPRs), 1 the tiled kernel for batches without solid
 * sites (more loads in flight), 2 the split pair (a psi launch and a
 * stream/force/collide launch per step over all members; tile, chunk and
 * workgroups do not apply), 3 the split pair as one persistent cooperative
 * launch per step (a grid barrier between the passes; every workgroup
 * resident; not captured into graphs).  Defaults 8, whole lattice, 512, 0. */
int dlbc_batch_set_config(dlbc_batch *b, int tile_y, int chunk, int workgroups, int kernel);
int dlbc_batch_get_config(const dlbc_batch *b, int *out /*4*/);
/* Time every (tile height, chunk, workgroups, kernel) candidate on the
 * current state (reads it, writes only the buffer the next step overwrites)
 * and keep the fastest; out receives it, best_ms its launch time. */
int dlbc_batch_autotune(dlbc_batch *b, int *out /*4*/, double *best_ms);
/* nsteps steps, one launch per step (two for kernel 2); with use_graph
 * pairs of steps are replayed from a captured HIP graph (not for kernel 3). */
int dlbc_batch_step(dlbc_batch *b, int nsteps, int use_graph);
/* The same steps as one launch per member and step (the per-lattice loop). */
int dlbc_batch_step_looped(dlbc_batch *b, int nsteps);
int dlbc_batch_synchronize(dlbc_batch *b);
const char *dlbc_batch_last_error(const dlbc_batch *b);
void dlbc_batch_destroy(dlbc_batch *b);

/* FluidInit (src/dlbc/lb/init.d:139-260). */
enum {
  DLBC_INIT_NONE = 0, DLBC_INIT_FILE, DLBC_INIT_CONST, DLBC_INIT_EQDIST,
  DLBC_INIT_EQDIST_PERTURB, DLBC_INIT_EQDIST_PERTURB_FRAC, DLBC_INIT_CONST_RANDOM,
  DLBC_INIT_EQDIST_RANDOM, DLBC_INIT_EQDIST_SPHERE, DLBC_INIT_EQDIST_SPHERE_FRAC,
  DLBC_INIT_EQDIST_CYLINDER, DLBC_INIT_EQDIST_CYLINDER_FRAC, DLBC_INIT_EQDIST_TWO_SPHERES,
  DLBC_INIT_EQDIST_TWO_SPHERES_FRAC, DLBC_INIT_EQDIST_LAMELLAE, DLBC_INIT_EQDIST_LAMELLAE_FRAC
};

/* Opaque Mersenne Twister state (std.random.Mt19937, src/dlbc/random.d:24). */
typedef struct dlbc_rng {
  uint32_t state[627];
} dlbc_rng;
void dlbc_rng_seed(dlbc_rng *rng, uint32_t seed);
uint32_t dlbc_rng_next(dlbc_rng *rng);
double dlbc_rng_uniform(dlbc_rng *rng, double a, double b);

/* initFluid (src/dlbc/lb/init.d:263-339) for one component into host[], a
 * halo-padded array with halo h.  eq_dist_form selects eqDistUnity
 * (src/dlbc/lb/eqdist.d:310-315).  rng is used by the random kinds. */
int dlbc_init_fluid(double *host, int d, int q, const int64_t *n, int halo, int eq_dist_form,
                    int kind, const double *dens, int ndens, double perturb, dlbc_rng *rng,
                    double radius, const double *offset, const double *separation, int axis,
                    double thickness, const double *widths);
/* initMask (src/dlbc/lb/mask.d:134-172): kind 0 None, 2 Plates, 3 Tube,
 * 4 Box, 5 Walls; dim_has_wall[2d] as lb.mask.dimensionHasWall. */
int dlbc_init_mask(int32_t *host, int d, const int64_t *n, int halo, int kind, int axis,
                   int wall_offset, const int *dim_has_wall, int wall_thickness);
/* initEqDistWall (src/dlbc/fields/init.d:469-477): solid sites get
 * [1, 0, ..., 0]. */
int dlbc_init_eq_dist_wall(double *host, const int32_t *mask, int d, int q, const int64_t *n,
                           int halo);

/* Device initialisation: the same initial state written straight into the
 * device, without the halo-padded host array (8 q B per site) and its upload.
 *
 * dlbc_gpu_init_fluid: initFluid (src/dlbc/lb/init.d:263-339) of component f;
 * the device populations equal dlbc_init_fluid() + dlbc_gpu_upload_populations()
 * bit for bit.  halo is the halo of the reference's field (2,
 * src/dlbc/lattice.d:96-103): the random kinds draw from rng in its
 * halo-padded row-major order (per population for ConstRandom), as
 * dlbc_init_fluid does, and discard the halo's draws.  Shifted storage takes
 * rho0 from dlbc_gpu_set_params(), like an upload.  None and File write
 * nothing. */
int dlbc_gpu_init_fluid(dlbc_gpu_ctx *ctx, int f, int halo, int eq_dist_form, int kind,
                        const double *dens, int ndens, double perturb, dlbc_rng *rng,
                        double radius, const double *offset, const double *separation, int axis,
                        double thickness, const double *widths);
/* initMask (src/dlbc/lb/mask.d:134-172, shapes src/dlbc/fields/init.d:377-467) on
 * the device: equal to dlbc_init_mask() + dlbc_gpu_upload_mask(). */
int dlbc_gpu_init_mask(dlbc_gpu_ctx *ctx, int kind, int axis, int wall_offset,
                       const int *dim_has_wall, int wall_thickness);
/* initEqDistWall (src/dlbc/fields/init.d:469-477) on the device: the solid sites
 * of component f get [1, 0, ..., 0]. */
int dlbc_gpu_init_eq_dist_wall(dlbc_gpu_ctx *ctx, int f);

#ifdef __cplusplus
}
#endif
#endif /* DLBC_HIP_H */
