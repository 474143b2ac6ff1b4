/*
 * oracle.h -- CPU restatement of DLBC's per-timestep lattice Boltzmann path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product (dlbc_amd/) never links or calls it.
 *
 * Layout follows the reference: every field is stored array-of-structures,
 * row-major over the halo-padded local lattice [n0+2h][n1+2h][n2+2h][k]
 * (x slowest, component index innermost), halo h = 2, and 4 for D1Q5
 * (src/dlbc/lattice.d:96-103, src/dlbc/fields/field.d:35-215).
 * One MPI rank is modelled; the halo exchange is the periodic self-exchange
 * of src/dlbc/fields/parallel.d:40-153.
 *
 * Parity unpinned: bitwise equality of psi = 1 - exp(-rho) with the
 * reference's D std.math.exp.  or_exp is correctly rounded (checked against
 * __float128 expq and mpmath); the reference's golden data pin psi and the
 * trajectories built on it only to their 1e-15 .. 1e-14 tolerances.
 * SecondOrder + Shan-Chen trajectories, C > 2 and D1Q5 trajectories have no
 * reference golden data: their building blocks are pinned one by one
 * (tests/test_oracle_*.py).
 */
#ifndef DLBC_ORACLE_H
#define DLBC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* EqDistForm (src/dlbc/lb/eqdist.d:23-36), PsiForm (src/dlbc/lb/force.d:48-57). */
enum { OR_SECOND_ORDER = 0, OR_THIRD_ORDER = 1, OR_BDIST2 = 2 };
enum { OR_PSI_LINEAR = 0, OR_PSI_EXPONENTIAL = 1 };

typedef struct or_lattice or_lattice;

/* Lattice with connectivity DdQq (D3Q19, D2Q9, D1Q3), C components,
 * physical lengths n[0..d).  Returns NULL on unsupported input. */
or_lattice *or_create(int d, int q, int ncomp, const int *n);
/* D's default-initialised double (a NaN), the populations' value before any init */
double or_d_double_init(void);
void or_destroy(or_lattice *L);

int or_set_params(or_lattice *L, const double *tau, const double *gcc,
                  const double *gwc, const double *global_acc, int eq_dist_form,
                  int psi_form, int enable_shan_chen);

/* Body force of a preCollisionHooks plugin (lb/collision.d:29,68-70):
 * F_f += body[f*d + k] after distributeForce, before the collision.  NULL or
 * all zero: no hook. */
int or_set_body_force(or_lattice *L, const double *body);
void or_pre_collision_hooks(or_lattice *L);

/* Raw storage (halo-padded, AoS).  Sizes: or_sites_h(L) sites. */
size_t or_sites_h(const or_lattice *L);
double *or_fluid(or_lattice *L, int f);      /* [sitesH][q]   */
int32_t *or_mask(or_lattice *L);             /* [sitesH]      */
double *or_density(or_lattice *L, int f);    /* [sitesH]      */
double *or_psi(or_lattice *L, int f);        /* [sitesH]      */
double *or_force(or_lattice *L, int f);      /* [sitesH][d]   */
double *or_weighted_velocity(or_lattice *L); /* [sitesH][d]   */

/* The passes of one time step, in the order of src/main.d:174-200. */
void or_exchange_halo(or_lattice *L);
void or_advect_fields(or_lattice *L);
void or_reset_force(or_lattice *L);
void or_add_shan_chen_force(or_lattice *L);
void or_distribute_force(or_lattice *L);
void or_prepare_to_collide(or_lattice *L);
void or_collide_fields(or_lattice *L);
void or_precalculate_densities(or_lattice *L);
/* One full step (exchange, advect, forces, collide) repeated nsteps times. */
void or_run_timesteps(or_lattice *L, int nsteps);
/* Accumulated wall-clock seconds per pass (7 entries, reference timer names
 * haloExchange, lb.advection, force.addSC, force.distr, lb.coll.prep,
 * lb.coll.coll, force.reset). */
void or_timers(const or_lattice *L, double *out7);

/* Diagnostics (src/dlbc/lb/density.d:136-147, momentum.d:131-147,
 * velocity.d:106-117, density.d:267-297). Physical sites only, row-major. */
double or_local_total_mass(const or_lattice *L, int f);
void or_local_total_momentum(const or_lattice *L, int f, double *out);
void or_velocity_field(const or_lattice *L, int f, double *out);   /* [sites][d] */
void or_density_field_phys(const or_lattice *L, int f, double *out); /* [sites] */
void or_colour_field(const or_lattice *L, int f1, int f2, double *out);

/* Copy between halo-padded storage and a physical-only AoS array. */
void or_get_population(const or_lattice *L, int f, double *out);  /* [sites][q] */
void or_set_population(or_lattice *L, int f, const double *in);
void or_get_mask(const or_lattice *L, int32_t *out);
void or_set_mask(or_lattice *L, const int32_t *in);

/* Connectivity tables (src/dlbc/lb/connectivity.d). */
int or_conn(int d, int q, int *vel /*q*d*/, int *bounce /*q*/, double *w /*q*/,
            double *css);

/* Scalar primitives. */
double or_pop_density(const double *pop, int q);              /* Phobos sum */
double or_psi_value(int psi_form, double rho);
double or_exp(double x);      /* correctly rounded exp (fast path + quad) */
double or_exp_quad(double x); /* (double) expq(x): 113-bit reference      */
int or_exp_fallbacks(void);   /* number of fast-path rejections so far   */
void or_eq_dist(int d, int q, int form, const double *pop, const double *dv,
                double rho0, double *out);

/* Mersenne Twister as std.random.Mt19937 and uniform!"[)"(a, b). */
typedef struct { uint32_t mt[624]; int mti; uint32_t front; } or_mt19937;
void or_mt_seed(or_mt19937 *g, uint32_t seed);
uint32_t or_mt_next(or_mt19937 *g);                 /* returns front, advances */
double or_uniform(or_mt19937 *g, double a, double b);

/* Fluid initialisation (src/dlbc/lb/init.d:139-339, src/dlbc/fields/init.d).
 * kind follows enum FluidInit (init.d:139-260).  Parameters:
 *   dens[ndens]   fluidDensities[f]
 *   perturb       fluidPerturb[f]
 *   seed          RNG seed used when reseed != 0 (randomReseed[f])
 *   radius        initRadius;  offset[d] initOffset;  separation[d] initSeparation
 *   axis          initAxis;    thickness interfaceThickness
 *   widths[ndens] lamellaeWidths
 * The global generator (used when reseed == 0) is or_global_rng(). */
enum {
  OR_INIT_NONE = 0, OR_INIT_FILE, OR_INIT_CONST, OR_INIT_EQDIST, OR_INIT_EQDIST_PERTURB,
  OR_INIT_EQDIST_PERTURB_FRAC, OR_INIT_CONST_RANDOM, OR_INIT_EQDIST_RANDOM,
  OR_INIT_EQDIST_SPHERE, OR_INIT_EQDIST_SPHERE_FRAC, OR_INIT_EQDIST_CYLINDER,
  OR_INIT_EQDIST_CYLINDER_FRAC, OR_INIT_EQDIST_TWO_SPHERES, OR_INIT_EQDIST_TWO_SPHERES_FRAC,
  OR_INIT_EQDIST_LAMELLAE, OR_INIT_EQDIST_LAMELLAE_FRAC
};
or_mt19937 *or_global_rng(void);
void or_seed_global_rng(uint32_t seed); /* initRNG, random.d:43-53 (rank 0) */
int or_init_fluid(or_lattice *L, int f, int kind, const double *dens, int ndens,
                  double perturb, int reseed, uint32_t seed, double radius,
                  const double *offset, const double *separation, int axis,
                  double thickness, const double *widths);
/* MaskInit (src/dlbc/lb/mask.d:54-79): 0 None, 2 Plates, 3 Tube, 4 Box, 5 Walls. */
int or_init_mask(or_lattice *L, int kind, int axis, int wall_offset,
                 const int *dim_has_wall /*d*2*/, int wall_thickness);
void or_init_eq_dist_wall(or_lattice *L, int f);

/* Threads of the OpenMP loops (timing only; results do not depend on it). */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
