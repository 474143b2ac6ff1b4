/*
 * oracle.c -- CPU restatement of DLBC's per-timestep lattice Boltzmann path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker of the HIP product
 * path, never part of it.
 *
 * Every function below restates one reference function; the citation names
 * the reference file:line it follows.  Floating-point expressions keep the
 * reference's left-to-right evaluation order and are compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 */
#include "oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXQ 19
#define OR_MAXC 8

struct or_lattice {
  int d, q, nc, h;
  int n[3], nh[3];           /* physical and halo-padded lengths; unused dims = 1 */
  size_t sites, sites_h;
  int vel[OR_MAXQ][3];
  int bounce[OR_MAXQ];
  double w[OR_MAXQ];
  double css;
  double *fluids[OR_MAXC];
  double *advection;
  int32_t *mask;
  double *density[OR_MAXC];
  double *psi[OR_MAXC];
  double *force[OR_MAXC];
  double *force_distributed;
  double *wvel;
  /* parameters */
  double tau[OR_MAXC];
  double gcc[OR_MAXC][OR_MAXC];
  double gwc[OR_MAXC];
  double global_acc[3];
  int eq_dist_form, psi_form, enable_sc;
  double body[OR_MAXC][3]; /* force density of a pre-collision hook plugin */
  int has_body;
  double timers[7];
  int density_fresh; /* Field.isFresh of L.density (lb/density.d:344-362) */
};

/* ------------------------------------------------------------------------ */
/* Connectivity: src/dlbc/lb/connectivity.d:198-418                         */
/* ------------------------------------------------------------------------ */

/* D3Q19 order, connectivity.d:401-418 (D3Q7 prefix 382-392). */
static const int D3Q19_VEL[19][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0}, {0, 1, 1},  {0, -1, 1},  {1, 0, 1},   {1, 0, -1}, {1, -1, 0},
    {-1, -1, 0}, {0, 1, -1}, {0, -1, -1}, {-1, 0, 1}, {-1, 0, -1}};
/* D2Q9 order, connectivity.d:356-374 (D2Q5 prefix 344-354). */
static const int D2Q9_VEL[9][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                    {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
/* D1Q3 order, connectivity.d:318-324; D1Q5 appends +2 and -2 (347-353). */
static const int D1Q3_VEL[3][1] = {{0}, {1}, {-1}};
static const int D1Q5_VEL[5][1] = {{0}, {1}, {-1}, {2}, {-2}};

int or_conn(int d, int q, int *vel, int *bounce, double *w, double *css) {
  int i, j, k;
  if (d == 3 && q == 19) {
    for (i = 0; i < 19; i++)
      for (j = 0; j < 3; j++) vel[i * 3 + j] = D3Q19_VEL[i][j];
    /* weights, connectivity.d:252-256 */
    w[0] = 1.0 / 3.0;
    for (i = 1; i < 7; i++) w[i] = 1.0 / 18.0;
    for (i = 7; i < 19; i++) w[i] = 1.0 / 36.0;
  } else if (d == 2 && q == 9) {
    for (i = 0; i < 9; i++)
      for (j = 0; j < 2; j++) vel[i * 2 + j] = D2Q9_VEL[i][j];
    /* connectivity.d:267-271 */
    w[0] = 4.0 / 9.0;
    for (i = 1; i < 5; i++) w[i] = 1.0 / 9.0;
    for (i = 5; i < 9; i++) w[i] = 1.0 / 36.0;
  } else if (d == 1 && q == 3) {
    for (i = 0; i < 3; i++) vel[i] = D1Q3_VEL[i][0];
    /* connectivity.d:285-288 */
    w[0] = 4.0 / 6.0;
    w[1] = 1.0 / 6.0;
    w[2] = 1.0 / 6.0;
  } else if (d == 1 && q == 5) {
    for (i = 0; i < 5; i++) vel[i] = D1Q5_VEL[i][0];
    /* connectivity.d:280-284 */
    w[0] = 6.0 / 12.0;
    w[1] = w[2] = 2.0 / 12.0;
    w[3] = w[4] = 1.0 / 12.0;
  } else {
    return -1;
  }
  /* c_s^2, connectivity.d:214-248: 1 for D1Q5, 1/3 for the others */
  *css = (d == 1 && q == 5) ? 1.0 : 1.0 / 3.0;
  /* generateBounce, connectivity.d:198-212: first j with c_i + c_j == 0 */
  for (i = 0; i < q; i++) {
    for (j = 0; j < q; j++) {
      int any = 0;
      for (k = 0; k < d; k++)
        if (vel[i * d + k] + vel[j * d + k] != 0) any = 1;
      if (!any) {
        bounce[i] = j;
        break;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Lattice storage: src/dlbc/lattice.d:96-181, lb/init.d:60-73              */
/* ------------------------------------------------------------------------ */

/* The bit pattern of D's default-initialised double, 0x7FFC000000000000 (a
 * quiet NaN), as the reference writes it for components initialised with
 * FluidInit.None (lb/init-2d fixtures, e.g. population-red-mask-init-none-2d). */
double or_d_double_init(void) {
  const uint64_t bits = 0x7FFC000000000000ull;
  double v;
  memcpy(&v, &bits, sizeof v);
  return v;
}

or_lattice *or_create(int d, int q, int ncomp, const int *n) {
  or_lattice *L;
  int i, vel[OR_MAXQ * 3];
  if (ncomp < 1 || ncomp > OR_MAXC || d < 1 || d > 3) return NULL;
  L = (or_lattice *)calloc(1, sizeof(*L));
  if (!L) return NULL;
  if (or_conn(d, q, vel, L->bounce, L->w, &L->css) != 0) {
    free(L);
    return NULL;
  }
  L->d = d;
  L->q = q;
  L->nc = ncomp;
  /* lattice.d:96-103: halo 2, and 4 for D1Q5, whose +-2 links and the
   * Shan-Chen stencil on them reach four sites out */
  L->h = (d == 1 && q == 5) ? 4 : 2;
  for (i = 0; i < q; i++) {
    int k;
    for (k = 0; k < 3; k++) L->vel[i][k] = k < d ? vel[i * d + k] : 0;
  }
  L->sites = 1;
  L->sites_h = 1;
  for (i = 0; i < 3; i++) {
    if (i < d) {
      if (n[i] < L->h) { /* the self-exchange needs n >= halo */
        free(L);
        return NULL;
      }
      L->n[i] = n[i];
      L->nh[i] = n[i] + 2 * L->h;
    } else {
      L->n[i] = 1;
      L->nh[i] = 1;
    }
    L->sites *= (size_t)L->n[i];
    L->sites_h *= (size_t)L->nh[i];
  }
  for (i = 0; i < ncomp; i++) {
    L->fluids[i] = (double *)malloc(sizeof(double) * L->sites_h * q);
    L->density[i] = (double *)calloc(L->sites_h, sizeof(double));
    L->psi[i] = (double *)calloc(L->sites_h, sizeof(double));
    L->force[i] = (double *)calloc(L->sites_h * d, sizeof(double));
    /* D initialises doubles to its default NaN (field.d: multidimArray
     * default init); a component left uninitialised (FluidInit.None) keeps it */
    {
      size_t s;
      for (s = 0; s < L->sites_h * q; s++) L->fluids[i][s] = or_d_double_init();
    }
  }
  L->advection = (double *)malloc(sizeof(double) * L->sites_h * q);
  {
    size_t s;
    for (s = 0; s < L->sites_h * q; s++) L->advection[s] = or_d_double_init();
  }
  L->mask = (int32_t *)calloc(L->sites_h, sizeof(int32_t));
  L->force_distributed = (double *)calloc(L->sites_h * d, sizeof(double));
  L->wvel = (double *)calloc(L->sites_h * d, sizeof(double));
  for (i = 0; i < ncomp; i++) L->tau[i] = 1.0;
  L->eq_dist_form = OR_SECOND_ORDER;
  L->psi_form = OR_PSI_LINEAR;
  return L;
}

void or_destroy(or_lattice *L) {
  int i;
  if (!L) return;
  for (i = 0; i < L->nc; i++) {
    free(L->fluids[i]);
    free(L->density[i]);
    free(L->psi[i]);
    free(L->force[i]);
  }
  free(L->advection);
  free(L->mask);
  free(L->force_distributed);
  free(L->wvel);
  free(L);
}

int or_set_params(or_lattice *L, const double *tau, const double *gcc, const double *gwc,
                  const double *global_acc, int eq_dist_form, int psi_form,
                  int enable_shan_chen) {
  int i, j;
  for (i = 0; i < L->nc; i++) {
    L->tau[i] = tau ? tau[i] : 1.0;
    L->gwc[i] = gwc ? gwc[i] : 0.0;
    for (j = 0; j < L->nc; j++) L->gcc[i][j] = gcc ? gcc[i * L->nc + j] : 0.0;
  }
  for (i = 0; i < 3; i++) L->global_acc[i] = (global_acc && i < L->d) ? global_acc[i] : 0.0;
  if (eq_dist_form < 0 || eq_dist_form > 2 || psi_form < 0 || psi_form > 1) return -1;
  L->eq_dist_form = eq_dist_form;
  L->psi_form = psi_form;
  L->enable_sc = enable_shan_chen;
  return 0;
}

/* A body force as a plugin would add it: a preCollisionHooks callback
 * (lb/collision.d:29, run at lb/collision.d:68-70 after distributeForce and
 * prepareToCollide) doing L.force[f][p][k] += body[f][k] over the force
 * field.  body[nc*d] row-major; NULL or all zero registers no hook. */
int or_set_body_force(or_lattice *L, const double *body) {
  int f, k;
  L->has_body = 0;
  for (f = 0; f < OR_MAXC; f++)
    for (k = 0; k < 3; k++) {
      L->body[f][k] = (body && f < L->nc && k < L->d) ? body[f * L->d + k] : 0.0;
      if (L->body[f][k] != 0.0) L->has_body = 1;
    }
  return 0;
}

/* preCollisionHooks.execute(L) with the body-force hook above, every stored
 * site (the hook iterates the whole force field) */
void or_pre_collision_hooks(or_lattice *L) {
  int f, k;
  size_t s;
  if (!L->has_body) return;
  for (f = 0; f < L->nc; f++)
    for (s = 0; s < L->sites_h; s++)
      for (k = 0; k < L->d; k++) L->force[f][s * L->d + k] += L->body[f][k];
}

size_t or_sites_h(const or_lattice *L) { return L->sites_h; }
double *or_fluid(or_lattice *L, int f) { return L->fluids[f]; }
int32_t *or_mask(or_lattice *L) { return L->mask; }
double *or_density(or_lattice *L, int f) { return L->density[f]; }
double *or_psi(or_lattice *L, int f) { return L->psi[f]; }
double *or_force(or_lattice *L, int f) { return L->force[f]; }
double *or_weighted_velocity(or_lattice *L) { return L->wvel; }

static inline size_t idx3(const or_lattice *L, long x, long y, long z) {
  return ((size_t)x * L->nh[1] + (size_t)y) * L->nh[2] + (size_t)z;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Scalar primitives                                                        */
/* ------------------------------------------------------------------------ */

/* density(): Phobos std.algorithm.sum over double[q] (lb/density.d:37-39).
 * For a sliceable range of doubles, sum = seed 0.0 + pairwise sum, where the
 * pairwise sum of r is r[0] (n=1), r[0]+r[1] (n=2), else
 * sum(r[0..n/2]) + sum(r[n/2..n]). */
static double pairwise(const double *r, int n) {
  if (n == 1) return r[0];
  if (n == 2) return r[0] + r[1];
  return pairwise(r, n / 2) + pairwise(r + n / 2, n - n / 2);
}

double or_pop_density(const double *pop, int q) { return 0.0 + pairwise(pop, q); }

/* ---- correctly rounded exp ----------------------------------------------
 * D's std.math.exp(double) (called at lb/force.d:327) evaluated in 80-bit
 * extended precision and rounded the result to double; that is correct
 * rounding except in rare double-rounding cases.  The restatement uses the
 * correctly rounded exp: a Ziv fast path in double-double arithmetic
 * (table 2^(j/32), degree-8 polynomial, error < 2^-70 relative) whose
 * rounding test falls back to expq (113-bit) when the result is too close to
 * a rounding boundary.  Subnormal results come from expq directly.        */
static double EXP_T_HI[32], EXP_T_LO[32];
static double LN2_32_HI, LN2_32_MID, LN2_32_LO, INV_LN2_32;
static int exp_ready = 0;
static int exp_fallback_count = 0;

static void exp_init(void) {
  int j;
  __float128 ln2_32 = M_LN2q / 32;
  double hi, mid;
  for (j = 0; j < 32; j++) {
    __float128 t = exp2q((__float128)j / 32);
    EXP_T_HI[j] = (double)t;
    EXP_T_LO[j] = (double)(t - EXP_T_HI[j]);
  }
  /* hi: ln2/32 with the low 16 mantissa bits cleared, so k*hi is exact for
   * |k| < 2^16 (|x| < 745 gives |k| < 34400). */
  {
    double v = (double)ln2_32;
    uint64_t b;
    memcpy(&b, &v, 8);
    b &= ~((uint64_t)0xFFFF);
    memcpy(&hi, &b, 8);
  }
  mid = (double)(ln2_32 - hi);
  LN2_32_HI = hi;
  LN2_32_MID = mid;
  LN2_32_LO = (double)(ln2_32 - hi - mid);
  INV_LN2_32 = (double)(32 / M_LN2q);
  exp_ready = 1;
}

double or_exp_quad(double x) { return (double)expq((__float128)x); }
int or_exp_fallbacks(void) { return exp_fallback_count; }

double or_exp(double x) {
  double kd, t, p, pe, s, bb, se, rh, rl, r2h, r2l, a, c3, u, ue, small, th, tl;
  double p1, p1e, h, e1, l, err, y1, y2;
  int k, j, m;
  if (!exp_ready) exp_init();
  if (isnan(x)) return x + x;
  if (x > 710.0) return INFINITY;
  if (x < -708.39) { /* subnormal or zero result: expq rounds directly */
    exp_fallback_count++;
    return or_exp_quad(x);
  }
  if (fabs(x) < 0x1p-54) return 1.0;
  kd = rint(x * INV_LN2_32);
  k = (int)kd;
  /* r = x - k ln2/32 as a double-double (rh, rl) */
  t = x - kd * LN2_32_HI; /* exact: kd*HI exact, Sterbenz */
  p = kd * LN2_32_MID;
  pe = fma(kd, LN2_32_MID, -p);
  s = t - p; /* TwoSum(t, -p) */
  bb = s - t;
  se = (t - (s - bb)) + (-p - bb);
  rl = se - pe - kd * LN2_32_LO;
  rh = s + rl;
  rl = rl - (rh - s);
  /* e^r = 1 + rh + rh^2/2 + [rl + rh*rl + rh^3/6 (1 + rh/4 + ...)] */
  r2h = rh * rh;
  r2l = fma(rh, rh, -r2h);
  a = 0.5 * r2h; /* exact */
  c3 = r2h * rh *
       (1.0 / 6 + rh * (1.0 / 24 + rh * (1.0 / 120 + rh * (1.0 / 720 + rh * (1.0 / 5040 +
                                                                         rh * (1.0 / 40320))))));
  small = rl + (0.5 * r2l + (rh * rl + c3));
  /* u = rh + a exactly as a double-double */
  u = rh + a;
  bb = u - rh;
  ue = (rh - (u - bb)) + (a - bb);
  j = k & 31;
  m = (k - j) / 32;
  th = EXP_T_HI[j];
  tl = EXP_T_LO[j];
  /* T (1 + u + ue + small) */
  p1 = th * u;
  p1e = fma(th, u, -p1);
  h = th + p1; /* FastTwoSum, |th| > |p1| */
  e1 = p1 - (h - th);
  l = e1 + (p1e + (th * (ue + small) + (tl + tl * u)));
  /* Ziv rounding test with a 2^-68 relative error allowance */
  err = 0x1p-68 * h;
  y1 = h + (l - err);
  y2 = h + (l + err);
  if (y1 != y2) {
    exp_fallback_count++;
    return or_exp_quad(x);
  }
  return ldexp(y1, m);
}

/* psi(), lb/force.d:320-332 */
double or_psi_value(int psi_form, double rho) {
  if (psi_form == OR_PSI_LINEAR) return rho;
  return 1.0 - or_exp(-rho);
}

/* rawVelocity, lb/velocity.d:37-48: vel[j] += n_i c_ij over i, then /= rho */
static void raw_velocity(const double *pop, int q, int d, const int (*vel)[3], double rho,
                         double *out) {
  int i, j;
  for (j = 0; j < d; j++) out[j] = 0.0;
  for (i = 0; i < q; i++)
    for (j = 0; j < d; j++) out[j] += pop[i] * (double)vel[i][j];
  for (j = 0; j < d; j++) out[j] /= rho;
}

/* dotProduct, range.d:57-67: sum0 = 0; sum0 += a[i]*b[i] in order */
static inline double dot_d(const double *a, const double *b, int d) {
  double s = 0;
  int i;
  for (i = 0; i < d; i++) s += a[i] * b[i];
  return s;
}
static inline double dot_c(const double *a, const int *c, int d) {
  double s = 0;
  int i;
  for (i = 0; i < d; i++) s += a[i] * (double)c[i];
  return s;
}

/* eqDist, lb/eqdist.d:61-278 */
static void eq_dist_impl(int d, int q, const int (*vel)[3], const double *w, double css,
                         int form, const double *pop, const double *dv, double rho0,
                         double *dist) {
  int vq;
  if (form == OR_SECOND_ORDER || form == OR_THIRD_ORDER) {
    double pv[3], v[3], vdotv;
    int vd;
    raw_velocity(pop, q, d, vel, rho0, pv); /* eqdist.d:74 */
    for (vd = 0; vd < d; vd++) v[vd] = dv[vd] + pv[vd];
    vdotv = dot_d(v, v, d);
    if (form == OR_SECOND_ORDER) { /* eqdist.d:82-89 */
      const double css2 = 2.0 * css;
      const double css2p2 = css2 * css;
      for (vq = 0; vq < q; vq++) {
        const double vdotcv = dot_c(v, vel[vq], d);
        dist[vq] = rho0 * w[vq] *
                   (1.0 + (vdotcv / css) + (vdotcv * vdotcv / css2p2) - (vdotv / css2));
      }
    } else { /* eqdist.d:90-100 */
      const double css2 = 2.0 * css;
      const double css2p2 = css2 * css;
      const double css6p3 = 3.0 * css2p2 * css;
      for (vq = 0; vq < q; vq++) {
        const double vdotcv = dot_c(v, vel[vq], d);
        dist[vq] = rho0 * w[vq] *
                   (1.0 + (vdotcv / css) + (vdotcv * vdotcv / css2p2) - (vdotv / css2) +
                    (vdotcv * vdotcv * vdotcv / css6p3) - (vdotv * vdotcv / css2p2));
      }
    }
    return;
  }
  /* BDist2 */
  if (d == 3 && q == 19) { /* eqdist.d:102-184 */
    const double aN1 = 1.0 / 36.0, aN0 = 1.0 / 3.0;
    const double haN1 = 0.5 * aN1, aN1_6 = aN1 / 6.0;
    const double T = 1.0 / 3.0, holdit = 1.0 / T;
    const double tux = dv[0], tuy = dv[1], tuz = dv[2];
    const double tuxt = tux * holdit, tuyt = tuy * holdit, tuzt = tuz * holdit;
    const double uke = 0.5 * (tux * tuxt + tuy * tuyt + tuz * tuzt);
    const double cst = 1.0 - uke;
    const double const1 = aN1 * cst;
    double cdotu, even, odd;
#define OR_EO(expr)                                             \
  cdotu = (expr);                                               \
  even = const1 + haN1 * cdotu * cdotu;                         \
  odd = const1 * cdotu + aN1_6 * cdotu * cdotu * cdotu
    OR_EO(tuxt);
    dist[1] = 2.0 * (even + odd);
    dist[2] = 2.0 * (even - odd);
    OR_EO(tuyt);
    dist[3] = 2.0 * (even + odd);
    dist[4] = 2.0 * (even - odd);
    OR_EO(tuzt);
    dist[5] = 2.0 * (even + odd);
    dist[6] = 2.0 * (even - odd);
    OR_EO(tuxt + tuyt);
    dist[7] = even + odd;
    dist[14] = even - odd;
    OR_EO(tuxt - tuyt);
    dist[13] = even + odd;
    dist[8] = even - odd;
    OR_EO(tuxt + tuzt);
    dist[11] = even + odd;
    dist[18] = even - odd;
    OR_EO(tuxt - tuzt);
    dist[12] = even + odd;
    dist[17] = even - odd;
    OR_EO(tuyt + tuzt);
    dist[9] = even + odd;
    dist[16] = even - odd;
    OR_EO(tuyt - tuzt);
    dist[15] = even + odd;
    dist[10] = even - odd;
    dist[0] = aN0 * cst;
    for (vq = 0; vq < q; vq++) dist[vq] *= rho0;
  } else if (d == 2 && q == 9) { /* eqdist.d:185-235 */
    const double aN1 = 1.0 / 24.0, aN0 = 1.0 / 2.0;
    const double haN1 = 0.5 * aN1, aN1_6 = aN1 / 6.0;
    const double T = 1.0 / 3.0, holdit = 1.0 / T;
    const double tux = dv[0], tuy = dv[1];
    const double tuxt = tux * holdit, tuyt = tuy * holdit;
    const double uke = 0.5 * (tux * tuxt + tuy * tuyt);
    const double cst = 1.0 - uke;
    const double const1 = aN1 * cst;
    double cdotu, even, odd;
    OR_EO(tuxt);
    dist[1] = 2.0 * (even + odd);
    dist[2] = 2.0 * (even - odd);
    OR_EO(tuyt);
    dist[3] = 2.0 * (even + odd);
    dist[4] = 2.0 * (even - odd);
    OR_EO(tuxt + tuyt);
    dist[5] = even + odd;
    dist[8] = even - odd;
    OR_EO(tuxt - tuyt);
    dist[6] = even + odd;
    dist[7] = even - odd;
    dist[0] = aN0 * cst;
    for (vq = 0; vq < q; vq++) dist[vq] *= rho0;
#undef OR_EO
  } else { /* eqdist.d:269-271: other connectivities return the population */
    for (vq = 0; vq < q; vq++) dist[vq] = pop[vq];
  }
}

void or_eq_dist(int d, int q, int form, const double *pop, const double *dv, double rho0,
                double *out) {
  int vel[OR_MAXQ * 3], bounce[OR_MAXQ], vv[OR_MAXQ][3], i, k;
  double w[OR_MAXQ], css;
  if (or_conn(d, q, vel, bounce, w, &css) != 0) return;
  for (i = 0; i < q; i++)
    for (k = 0; k < 3; k++) vv[i][k] = k < d ? vel[i * d + k] : 0;
  eq_dist_impl(d, q, (const int(*)[3])vv, w, css, form, pop, dv, rho0, out);
}

/* eqDistUnity, eqdist.d:310-315 */
static void eq_dist_unity(const or_lattice *L, int form, double *out) {
  double pop0[OR_MAXQ] = {0}, dv[3] = {0, 0, 0};
  pop0[0] = 1.0;
  eq_dist_impl(L->d, L->q, (const int(*)[3])L->vel, L->w, L->css, form, pop0, dv,
               or_pop_density(pop0, L->q), out);
}

/* ------------------------------------------------------------------------ */
/* Halo exchange: src/dlbc/fields/parallel.d:40-153, one rank, periodic.    */
/* Send-to-positive copies planes [nH-2h, nH-h) into [0, h); send-to-       */
/* negative copies [h, 2h) into [nH-h, nH); x, then y, then z, each over the */
/* full extent of the other axes (haloOffset = 0), through a copy buffer.    */
/* ------------------------------------------------------------------------ */
static void exchange_generic(const or_lattice *L, void *data, size_t elem) {
  int dim, dir;
  char *base = (char *)data;
  for (dim = 0; dim < L->d; dim++) {
    for (dir = 0; dir < 2; dir++) {
      const int h = L->h, nh = L->nh[dim];
      const int src0 = dir == 0 ? nh - 2 * h : h;
      const int dst0 = dir == 0 ? 0 : nh - h;
      /* the slab [lo, lo+h) of axis dim is (outer) blocks of (h * inner)
       * contiguous elements, inner = product of the faster axes */
      size_t outer = 1, inner = 1, o;
      int a;
      char *buf;
      for (a = 0; a < dim; a++) outer *= (size_t)L->nh[a];
      for (a = dim + 1; a < 3; a++) inner *= (size_t)L->nh[a];
      buf = (char *)malloc(outer * h * inner * elem); /* sbuffer = slab.dup */
      for (o = 0; o < outer; o++)
        memcpy(buf + o * h * inner * elem, base + ((o * nh + src0) * inner) * elem,
               h * inner * elem);
      for (o = 0; o < outer; o++)
        memcpy(base + ((o * nh + dst0) * inner) * elem, buf + o * h * inner * elem,
               h * inner * elem);
      free(buf);
    }
  }
}

/* Lattice.exchangeHalo, lattice.d:200-214: all @Exchange fields
 * (fluids[0..C), mask; the elec fields are off). */
void or_exchange_halo(or_lattice *L) {
  int f;
  double t0 = now_s();
  for (f = 0; f < L->nc; f++) exchange_generic(L, L->fluids[f], sizeof(double) * L->q);
  exchange_generic(L, L->mask, sizeof(int32_t));
  L->timers[0] += now_s() - t0;
}

/* ------------------------------------------------------------------------ */
/* Advection: lb/advection.d:61-99                                          */
/* ------------------------------------------------------------------------ */
void or_advect_fields(or_lattice *L) {
  int f;
  const int q = L->q;
  double t0 = now_s();
  for (f = 0; f < L->nc; f++) {
    const double *src = L->fluids[f];
    double *dst = L->advection;
    long x;
#pragma omp parallel for schedule(static)
    for (x = 0; x < L->nh[0]; x++) {
      long y, z;
      for (y = 0; y < L->nh[1]; y++)
        for (z = 0; z < L->nh[2]; z++) {
          const long p[3] = {x, y, z};
          const size_t ip = idx3(L, x, y, z);
          int i, edge = 0, vq;
          for (i = 0; i < L->d; i++) /* isOnEdge, advection.d:91-99 */
            if (p[i] == 0 || p[i] == L->nh[i] - 1) edge = 1;
          if (edge) continue;
          /* D1Q5 on its halo of 4: the second layer's +-2 sources lie outside
           * the field (the reference indexes past it there); no physical
           * site ever reads that layer, so it keeps its populations */
          for (vq = 0; vq < q; vq++)
            for (i = 0; i < L->d; i++) {
              const long sp = p[i] - L->vel[vq][i];
              if (sp < 0 || sp >= L->nh[i]) edge = 1;
            }
          if (edge) {
            memcpy(dst + ip * q, src + ip * q, sizeof(double) * q);
            continue;
          }
          if (L->mask[ip] == 0) { /* isAdvectable */
            for (vq = 0; vq < q; vq++) {
              const size_t inb = idx3(L, x - L->vel[vq][0], y - L->vel[vq][1],
                                      z - L->vel[vq][2]);
              if (L->mask[inb] != 0) /* isBounceBack */
                dst[ip * q + vq] = src[ip * q + L->bounce[vq]];
              else
                dst[ip * q + vq] = src[inb * q + vq];
            }
          } else {
            memcpy(dst + ip * q, src + ip * q, sizeof(double) * q);
          }
        }
    }
    /* swap(L.fluids[f], L.advection), advection.d:87 */
    L->advection = L->fluids[f];
    L->fluids[f] = dst;
  }
  L->density_fresh = 0; /* postAdvectionHooks: markDensitiesAsStale */
  L->timers[1] += now_s() - t0;
}

/* precalculateDensities / densityField: lb/density.d:89-98,344-353.
 * Every stored site, halo included. */
void or_precalculate_densities(or_lattice *L) {
  int f;
  if (L->density_fresh) return;
  L->density_fresh = 1;
  for (f = 0; f < L->nc; f++) {
    const double *pop = L->fluids[f];
    double *rho = L->density[f];
    long s;
#pragma omp parallel for schedule(static)
    for (s = 0; s < (long)L->sites_h; s++) rho[s] = or_pop_density(pop + (size_t)s * L->q, L->q);
  }
}

/* resetForce, lb/force.d:165-170 */
void or_reset_force(or_lattice *L) {
  int f;
  double t0 = now_s();
  for (f = 0; f < L->nc; f++) memset(L->force[f], 0, sizeof(double) * L->sites_h * L->d);
  memset(L->force_distributed, 0, sizeof(double) * L->sites_h * L->d);
  L->timers[6] += now_s() - t0;
}

/* precalculatePsi + psiField, lb/force.d:137-147,360-367 (every stored site) */
static void precalculate_psi(or_lattice *L) {
  int f;
  or_precalculate_densities(L);
  for (f = 0; f < L->nc; f++) {
    long s;
#pragma omp parallel for schedule(static)
    for (s = 0; s < (long)L->sites_h; s++)
      L->psi[f][s] = or_psi_value(L->psi_form, L->density[f][s]);
  }
}

/* physical-site iteration helper bounds: field.d:152-195 */
#define PHYS_BEGIN(L, d_) ((d_) < (L)->d ? (L)->h : 0)
#define PHYS_END(L, d_) ((d_) < (L)->d ? (L)->nh[d_] - (L)->h : 1)

/* addShanChenForcePsi, lb/force.d:238-309 */
void or_add_shan_chen_force(or_lattice *L) {
  int f1, f2;
  const int d = L->d, q = L->q;
  double t0;
  if (!L->enable_sc) return;
  t0 = now_s();
  precalculate_psi(L);
  for (f1 = 0; f1 < L->nc; f1++) {
    for (f2 = 0; f2 < L->nc; f2++) {
      const double cc = L->gcc[f1][f2];
      long x;
      if (cc == 0.0) continue;
#pragma omp parallel for schedule(static)
      for (x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++) {
        long y, z;
        for (y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)
          for (z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++) {
            const size_t ip = idx3(L, x, y, z);
            double *force = L->force[f1] + ip * d;
            double psiden1;
            int vq, vd;
            if (L->mask[ip] != 0) continue; /* isCollidable */
            psiden1 = L->psi[f1][ip];
            for (vq = 1; vq < q; vq++) { /* Iota!(1, q-1): q = 1 .. Q-1 */
              const size_t inb =
                  idx3(L, x + L->vel[vq][0], y + L->vel[vq][1], z + L->vel[vq][2]);
              const double psiden2 = (L->mask[inb] != 0) ? L->psi[f2][ip] : L->psi[f2][inb];
              const double prefactor = L->w[vq] * psiden1 * psiden2 * cc;
              for (vd = 0; vd < d; vd++) force[vd] -= prefactor * (double)L->vel[vq][vd];
            }
          }
      }
    }
    {
      const double wc = L->gwc[f1]; /* wall interactions, force.d:284-307 */
      long x;
      if (wc == 0.0) continue;
#pragma omp parallel for schedule(static)
      for (x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++) {
        long y, z;
        for (y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)
          for (z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++) {
            const size_t ip = idx3(L, x, y, z);
            double *force = L->force[f1] + ip * d;
            double psiden1;
            int vq, vd;
            if (L->mask[ip] != 0) continue;
            psiden1 = L->psi[f1][ip];
            for (vq = 0; vq < q; vq++) {
              const size_t inb =
                  idx3(L, x + L->vel[vq][0], y + L->vel[vq][1], z + L->vel[vq][2]);
              if (L->mask[inb] != 0) {
                const double prefactor = L->w[vq] * psiden1 * L->density[f1][inb] * wc;
                for (vd = 0; vd < d; vd++) force[vd] -= prefactor * (double)L->vel[vq][vd];
              }
            }
          }
      }
    }
  }
  L->timers[2] += now_s() - t0;
}

/* distributeForce, lb/force.d:179-197 (physical sites) */
void or_distribute_force(or_lattice *L) {
  const int d = L->d;
  long x;
  double t0 = now_s();
  or_precalculate_densities(L);
#pragma omp parallel for schedule(static)
  for (x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++) {
    long y, z;
    for (y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)
      for (z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++) {
        const size_t ip = idx3(L, x, y, z);
        double total = 0.0;
        int f, vd;
        for (f = 0; f < L->nc; f++) total += L->density[f][ip];
        for (f = 0; f < L->nc; f++)
          for (vd = 0; vd < d; vd++)
            L->force[f][ip * d + vd] +=
                L->force_distributed[ip * d + vd] * (L->density[f][ip] / total);
      }
  }
  L->timers[3] += now_s() - t0;
}

/* prepareToCollide + precalculateWeightedVelocity, lb/collision.d:41-51,156-196 */
void or_prepare_to_collide(or_lattice *L) {
  const int d = L->d, q = L->q;
  long x;
  double t0 = now_s();
  if (L->eq_dist_form != OR_BDIST2) return;
#pragma omp parallel for schedule(static)
  for (x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++) {
    long y, z;
    for (y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)
      for (z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++) {
        const size_t ip = idx3(L, x, y, z);
        double *wv = L->wvel + ip * d;
        int f, vq, vd;
        for (vd = 0; vd < d; vd++) wv[vd] = 0.0;
        for (f = 0; f < L->nc; f++) {
          double per[3] = {0.0, 0.0, 0.0};
          const double *pop = L->fluids[f] + ip * q;
          for (vq = 0; vq < q; vq++)
            for (vd = 0; vd < d; vd++) per[vd] += pop[vq] * (double)L->vel[vq][vd];
          for (vd = 0; vd < d; vd++) wv[vd] += per[vd] / L->tau[f];
        }
      }
  }
  or_precalculate_densities(L);
#pragma omp parallel for schedule(static)
  for (x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++) {
    long y, z;
    for (y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)
      for (z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++) {
        const size_t ip = idx3(L, x, y, z);
        double total = 0.0;
        int f, vd;
        for (f = 0; f < L->nc; f++) total += L->density[f][ip];
        for (vd = 0; vd < d; vd++) L->wvel[ip * d + vd] /= total;
      }
  }
  L->timers[4] += now_s() - t0;
}

/* collideFieldsKernel, lb/collision.d:90-119 */
void or_collide_fields(or_lattice *L) {
  const int d = L->d, q = L->q;
  int f;
  double t0 = now_s();
  or_precalculate_densities(L);
  for (f = 0; f < L->nc; f++) {
    const double omega = 1.0 / L->tau[f];
    const double tau = L->tau[f];
    long x;
#pragma omp parallel for schedule(static)
    for (x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++) {
      long y, z;
      for (y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)
        for (z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++) {
          const size_t ip = idx3(L, x, y, z);
          double *pop = L->fluids[f] + ip * q;
          double dv[3], eq[OR_MAXQ], den;
          int vd, vq;
          if (L->mask[ip] != 0) continue;
          den = L->density[f][ip];
          for (vd = 0; vd < d; vd++) {
            dv[vd] = tau * (L->global_acc[vd] + L->force[f][ip * d + vd] / den);
            if (L->eq_dist_form == OR_BDIST2) dv[vd] += L->wvel[ip * d + vd];
          }
          eq_dist_impl(d, q, (const int(*)[3])L->vel, L->w, L->css, L->eq_dist_form, pop, dv,
                       den, eq);
          for (vq = 0; vq < q; vq++) pop[vq] = pop[vq] * (1.0 - omega) + omega * eq[vq];
        }
    }
  }
  L->timers[5] += now_s() - t0;
}

/* runTimeloop body, src/main.d:174-200 (elec and output are off) */
void or_run_timesteps(or_lattice *L, int nsteps) {
  int t;
  for (t = 0; t < nsteps; t++) {
    or_exchange_halo(L);
    or_advect_fields(L);
    or_reset_force(L);
    or_add_shan_chen_force(L);
    or_distribute_force(L);
    or_prepare_to_collide(L);
    or_pre_collision_hooks(L); /* collideFields runs them first, collision.d:68-70 */
    or_collide_fields(L);
  }
}

void or_timers(const or_lattice *L, double *out7) { memcpy(out7, L->timers, sizeof(L->timers)); }

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                              */
/* ------------------------------------------------------------------------ */
#define FOR_PHYS(L)                                            \
  for (long x = PHYS_BEGIN(L, 0); x < PHYS_END(L, 0); x++)     \
    for (long y = PHYS_BEGIN(L, 1); y < PHYS_END(L, 1); y++)   \
      for (long z = PHYS_BEGIN(L, 2); z < PHYS_END(L, 2); z++)

/* localTotalMass, lb/density.d:136-147 */
double or_local_total_mass(const or_lattice *L, int f) {
  double mass = 0.0;
  FOR_PHYS(L) {
    const size_t ip = idx3(L, x, y, z);
    if (L->mask[ip] == 0) mass += or_pop_density(L->fluids[f] + ip * L->q, L->q);
  }
  return mass;
}

/* localTotalMomentum, lb/momentum.d:131-147; momentum() = sum n_i c_i
 * (momentum.d:37-52: vel[j] += n_i * c_ij in i order) */
void or_local_total_momentum(const or_lattice *L, int f, double *out) {
  int j;
  for (j = 0; j < L->d; j++) out[j] = 0.0;
  FOR_PHYS(L) {
    const size_t ip = idx3(L, x, y, z);
    if (L->mask[ip] == 0) {
      double pm[3] = {0, 0, 0};
      const double *pop = L->fluids[f] + ip * L->q;
      int i;
      for (i = 0; i < L->q; i++)
        for (j = 0; j < L->d; j++) pm[j] += pop[i] * (double)L->vel[i][j];
      for (j = 0; j < L->d; j++) out[j] += pm[j];
    }
  }
}

/* rawVelocityField, lb/velocity.d:106-117 (zero on solid sites) */
void or_velocity_field(const or_lattice *L, int f, double *out) {
  size_t o = 0;
  FOR_PHYS(L) {
    const size_t ip = idx3(L, x, y, z);
    const double *pop = L->fluids[f] + ip * L->q;
    int j;
    if (L->mask[ip] == 0) {
      raw_velocity(pop, L->q, L->d, (const int(*)[3])L->vel, or_pop_density(pop, L->q),
                   out + o * L->d);
    } else {
      for (j = 0; j < L->d; j++) out[o * L->d + j] = 0.0;
    }
    o++;
  }
}

void or_density_field_phys(const or_lattice *L, int f, double *out) {
  size_t o = 0;
  FOR_PHYS(L) {
    const size_t ip = idx3(L, x, y, z);
    out[o++] = or_pop_density(L->fluids[f] + ip * L->q, L->q);
  }
}

/* colourField, lb/density.d:284-295 */
void or_colour_field(const or_lattice *L, int f1, int f2, double *out) {
  size_t o = 0;
  FOR_PHYS(L) {
    const size_t ip = idx3(L, x, y, z);
    if (L->mask[ip] == 0)
      out[o] = or_pop_density(L->fluids[f1] + ip * L->q, L->q) -
               or_pop_density(L->fluids[f2] + ip * L->q, L->q);
    else
      out[o] = 0.0;
    o++;
  }
}

void or_get_population(const or_lattice *L, int f, double *out) {
  size_t o = 0;
  FOR_PHYS(L) {
    memcpy(out + o * L->q, L->fluids[f] + idx3(L, x, y, z) * L->q, sizeof(double) * L->q);
    o++;
  }
}
void or_set_population(or_lattice *L, int f, const double *in) {
  size_t o = 0;
  L->density_fresh = 0;
  FOR_PHYS(L) {
    memcpy(L->fluids[f] + idx3(L, x, y, z) * L->q, in + o * L->q, sizeof(double) * L->q);
    o++;
  }
}
void or_get_mask(const or_lattice *L, int32_t *out) {
  size_t o = 0;
  FOR_PHYS(L) out[o++] = L->mask[idx3(L, x, y, z)];
}
void or_set_mask(or_lattice *L, const int32_t *in) {
  size_t o = 0;
  FOR_PHYS(L) L->mask[idx3(L, x, y, z)] = in[o++];
}

/* ------------------------------------------------------------------------ */
/* Random numbers: std.random.Mt19937 (random.d:24) and uniform!"[)"        */
/* (used at fields/init.d:43,64,72,80).                                     */
/* ------------------------------------------------------------------------ */
static void mt_generate(or_mt19937 *g) {
  static const uint32_t mag[2] = {0u, 0x9908b0dfu};
  int k;
  uint32_t y;
  for (k = 0; k < 624 - 397; k++) {
    y = (g->mt[k] & 0x80000000u) | (g->mt[k + 1] & 0x7fffffffu);
    g->mt[k] = g->mt[k + 397] ^ (y >> 1) ^ mag[y & 1u];
  }
  for (; k < 623; k++) {
    y = (g->mt[k] & 0x80000000u) | (g->mt[k + 1] & 0x7fffffffu);
    g->mt[k] = g->mt[k + (397 - 624)] ^ (y >> 1) ^ mag[y & 1u];
  }
  y = (g->mt[623] & 0x80000000u) | (g->mt[0] & 0x7fffffffu);
  g->mt[623] = g->mt[396] ^ (y >> 1) ^ mag[y & 1u];
  g->mti = 0;
}

static uint32_t mt_genrand(or_mt19937 *g) {
  uint32_t y;
  if (g->mti >= 624) mt_generate(g);
  y = g->mt[g->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* seed(): init_genrand, then the first output becomes front */
void or_mt_seed(or_mt19937 *g, uint32_t seed) {
  int i;
  g->mt[0] = seed;
  for (i = 1; i < 624; i++)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->mti = 624;
  g->front = mt_genrand(g);
}

uint32_t or_mt_next(or_mt19937 *g) {
  const uint32_t v = g->front;
  g->front = mt_genrand(g);
  return v;
}

/* uniform!"[)"(a, b): _a + (_b - _a) * (front - min) / (max - min) with
 * _b = nextafter(b, -inf) */
double or_uniform(or_mt19937 *g, double a, double b) {
  const double bb = nextafter(b, -INFINITY);
  const double x = (double)or_mt_next(g);
  return a + (bb - a) * x / 4294967295.0;
}

static or_mt19937 global_rng;
static int global_rng_seeded = 0;
void or_seed_global_rng(uint32_t seed) {
  or_mt_seed(&global_rng, seed);
  global_rng_seeded = 1;
}
or_mt19937 *or_global_rng(void) {
  if (!global_rng_seeded) {
    or_mt_seed(&global_rng, 0u);
    global_rng_seeded = 1;
  }
  return &global_rng;
}

/* ------------------------------------------------------------------------ */
/* Initial conditions: src/dlbc/fields/init.d                               */
/* ------------------------------------------------------------------------ */

/* symmetricLinearTransition, fields/init.d:346-366 */
static double slt(double pos, double width, double negval, double posval) {
  if (width == 0.0) return pos < 0 ? negval : posval;
  if (pos < -0.5 * width) return negval;
  if (pos > 0.5 * width) return posval;
  return negval * (0.5 - pos / width) + posval * (0.5 + pos / width);
}

/* global coordinate of a stored index on one rank: p + 0*n - haloSize */
static inline double gcoord(const or_lattice *L, long p) { return (double)p - (double)L->h; }

/* sphere/two-sphere/cylinder relative position helpers, init.d:94-102,142-153,194-206 */
static double sphere_relpos(const or_lattice *L, const long *p, double radius,
                            const double *offset) {
  double off[3];
  int i;
  for (i = 0; i < L->d; i++) {
    const double gn = gcoord(L, p[i]);
    off[i] = gn - ((double)((size_t)L->n[i]) * 0.5 + offset[i]) + 0.5;
  }
  return sqrt(dot_d(off, off, L->d)) - radius;
}

int or_init_fluid(or_lattice *L, int f, int kind, const double *dens, int ndens,
                  double perturb, int reseed, uint32_t seed, double radius,
                  const double *offset, const double *separation, int axis, double thickness,
                  const double *widths) {
  const int q = L->q, d = L->d;
  double unity[OR_MAXQ];
  or_mt19937 local, *g;
  double *fld = L->fluids[f];
  size_t s;
  int vq;
  L->density_fresh = 0;
  if (reseed) { /* lb/init.d:113-121, shiftSeedByRank=false on one rank */
    or_mt_seed(&local, seed);
    g = &local;
  } else {
    g = or_global_rng();
  }
  eq_dist_unity(L, L->eq_dist_form, unity);
  switch (kind) {
  case OR_INIT_NONE:
  case OR_INIT_FILE:
    return 0;
  case OR_INIT_CONST: /* initConst, init.d:26-37 */
    for (s = 0; s < L->sites_h * q; s++) fld[s] = dens[0];
    return 0;
  case OR_INIT_CONST_RANDOM: /* init.d:39-50: one draw per population value */
    for (s = 0; s < L->sites_h * q; s++) fld[s] = dens[0] * or_uniform(g, 0.0, 2.0);
    return 0;
  case OR_INIT_EQDIST: { /* init.d:52-58 */
    double eqpop[OR_MAXQ];
    for (vq = 0; vq < q; vq++) eqpop[vq] = dens[0] * unity[vq];
    for (s = 0; s < L->sites_h; s++) memcpy(fld + s * q, eqpop, sizeof(double) * q);
    return 0;
  }
  case OR_INIT_EQDIST_RANDOM: /* init.d:60-66 */
    for (s = 0; s < L->sites_h; s++) {
      const double a = dens[0] * or_uniform(g, 0.0, 2.0);
      for (vq = 0; vq < q; vq++) fld[s * q + vq] = a * unity[vq];
    }
    return 0;
  case OR_INIT_EQDIST_PERTURB: /* init.d:68-74 */
    for (s = 0; s < L->sites_h; s++) {
      const double a = dens[0] + or_uniform(g, -perturb, perturb);
      for (vq = 0; vq < q; vq++) fld[s * q + vq] = a * unity[vq];
    }
    return 0;
  case OR_INIT_EQDIST_PERTURB_FRAC: /* init.d:76-82 */
    for (s = 0; s < L->sites_h; s++) {
      const double a = dens[0] * (1.0 + or_uniform(g, -perturb, perturb));
      for (vq = 0; vq < q; vq++) fld[s * q + vq] = a * unity[vq];
    }
    return 0;
  default:
    break;
  }
  /* shaped initialisations loop over every stored site (field.arr) */
  {
    double rad = radius, off[3] = {0, 0, 0}, sep[3] = {0, 0, 0}, wid[16];
    double interfaces[17] = {0};
    int i, nint = 0;
    if (kind == OR_INIT_EQDIST_SPHERE_FRAC || kind == OR_INIT_EQDIST_TWO_SPHERES_FRAC) {
      size_t small = (size_t)L->n[0]; /* init.d:112-123 */
      for (i = 0; i < d; i++)
        if ((size_t)L->n[i] < small) small = (size_t)L->n[i];
      rad = radius * (double)small;
      for (i = 0; i < d; i++) {
        off[i] = offset[i] * (double)L->n[i];
        if (separation) sep[i] = separation[i] * (double)L->n[i];
      }
    } else if (kind == OR_INIT_EQDIST_CYLINDER_FRAC) {
      size_t small = 0; /* init.d:217-229 */
      for (i = 0; i < d; i++)
        if (i != axis && ((size_t)L->n[i] < small || small == 0)) small = (size_t)L->n[i];
      rad = radius * (double)small;
      for (i = 0; i < d; i++) off[i] = offset[i] * (double)L->n[i];
    } else {
      for (i = 0; i < d; i++) {
        if (offset) off[i] = offset[i];
        if (separation) sep[i] = separation[i];
      }
    }
    if (kind == OR_INIT_EQDIST_LAMELLAE || kind == OR_INIT_EQDIST_LAMELLAE_FRAC) {
      if (ndens > 16 || axis >= d) return -1;
      for (i = 0; i < ndens; i++)
        wid[i] = kind == OR_INIT_EQDIST_LAMELLAE_FRAC ? widths[i] * (double)L->n[axis]
                                                      : widths[i];
      /* calculateInterfaces, init.d:335-344 */
      interfaces[0] = 0.0;
      for (i = 0; i < ndens; i++) interfaces[i + 1] = wid[i];
      nint = ndens + 1;
      for (i = 1; i < nint; i++) interfaces[i] += interfaces[i - 1];
      for (i = 0; i < nint; i++) interfaces[i] -= 0.5;
    }
    for (long x = 0; x < L->nh[0]; x++)
      for (long y = 0; y < L->nh[1]; y++)
        for (long z = 0; z < L->nh[2]; z++) {
          const long p[3] = {x, y, z};
          double val = 0.0;
          double *e = fld + idx3(L, x, y, z) * q;
          switch (kind) {
          case OR_INIT_EQDIST_SPHERE:
          case OR_INIT_EQDIST_SPHERE_FRAC: /* init.d:84-103 */
            val = slt(sphere_relpos(L, p, rad, off), thickness, dens[0], dens[1]);
            break;
          case OR_INIT_EQDIST_TWO_SPHERES:
          case OR_INIT_EQDIST_TWO_SPHERES_FRAC: { /* init.d:129-154 */
            double o1[3], o2[3], r1, r2;
            for (i = 0; i < d; i++) {
              const double gn = gcoord(L, p[i]);
              o1[i] = gn - ((double)((size_t)L->n[i]) * 0.5 + off[i] - 0.5 * sep[i]) + 0.5;
              o2[i] = gn - ((double)((size_t)L->n[i]) * 0.5 + off[i] + 0.5 * sep[i]) + 0.5;
            }
            r1 = sqrt(dot_d(o1, o1, d)) - rad;
            r2 = sqrt(dot_d(o2, o2, d)) - rad;
            val = slt(r2 < r1 ? r2 : r1, thickness, dens[0], dens[1]);
            break;
          }
          case OR_INIT_EQDIST_CYLINDER:
          case OR_INIT_EQDIST_CYLINDER_FRAC: { /* init.d:184-208 */
            double o[3];
            for (i = 0; i < d; i++) {
              if (i == axis) {
                o[i] = 0.0;
              } else {
                const double gn = gcoord(L, p[i]);
                o[i] = gn - ((double)((size_t)L->n[i]) * 0.5 + off[i]) + 0.5;
              }
            }
            val = slt(sqrt(dot_d(o, o, d)) - rad, thickness, dens[0], dens[1]);
            break;
          }
          case OR_INIT_EQDIST_LAMELLAE:
          case OR_INIT_EQDIST_LAMELLAE_FRAC: { /* init.d:236-259, 315-333 */
            const double gp = gcoord(L, p[axis]);
            int closest = 0;
            double mind = fabs(interfaces[0] - gp), md;
            for (i = 1; i < nint; i++) {
              const double df = fabs(interfaces[i] - gp);
              if (df < mind) {
                mind = df;
                closest = i;
              }
            }
            md = gp - interfaces[closest];
            if (closest == 0)
              val = slt(md, thickness, dens[ndens - 1], dens[closest]);
            else if (closest == ndens)
              val = slt(md, thickness, dens[closest - 1], dens[0]);
            else
              val = slt(md, thickness, dens[closest - 1], dens[closest]);
            break;
          }
          default:
            return -1;
          }
          for (vq = 0; vq < q; vq++) e[vq] = val * unity[vq];
        }
  }
  return 0;
}

/* initEqDistWall, fields/init.d:469-477 (every stored site) */
void or_init_eq_dist_wall(or_lattice *L, int f) {
  size_t s;
  int vq;
  for (s = 0; s < L->sites_h; s++)
    if (L->mask[s] == 1) {
      for (vq = 0; vq < L->q; vq++) L->fluids[f][s * L->q + vq] = 0.0;
      L->fluids[f][s * L->q] = 1.0;
    }
}

/* initMask, lb/mask.d:134-172; shapes fields/init.d:377-467 */
int or_init_mask(or_lattice *L, int kind, int axis, int wall_offset, const int *dim_has_wall,
                 int wall_thickness) {
  for (long x = 0; x < L->nh[0]; x++)
    for (long y = 0; y < L->nh[1]; y++)
      for (long z = 0; z < L->nh[2]; z++) {
        const long p[3] = {x, y, z};
        int32_t e = 0;
        int vd;
        long gn;
        switch (kind) {
        case 0: /* None */
          break;
        case 2: /* Plates, init.d:425-436 */
          gn = p[axis] - L->h;
          if (gn == wall_offset || gn == (long)L->n[axis] - 1 - wall_offset) e = 1;
          break;
        case 3: /* Tube, init.d:400-412 */
          for (vd = 0; vd < L->d; vd++) {
            gn = p[vd] - L->h;
            if (vd != axis && (gn == 0 || gn == (long)L->n[vd] - 1)) e = 1;
          }
          break;
        case 4: /* Box, init.d:377-388 */
          for (vd = 0; vd < L->d; vd++) {
            gn = p[vd] - L->h;
            if (gn == 0 || gn == (long)L->n[vd] - 1) e = 1;
          }
          break;
        case 5: /* Walls, init.d:449-467 */
          for (vd = 0; vd < L->d; vd++) {
            gn = p[vd] - L->h;
            if (dim_has_wall[2 * vd] && gn < wall_thickness) e = 1;
            /* ptrdiff_t >= size_t compares as unsigned in D: halo sites with
             * gn < 0 satisfy the upper-wall test too (init.d:461) */
            if (dim_has_wall[2 * vd + 1] &&
                (unsigned long)gn >= (unsigned long)((long)L->n[vd] - wall_thickness))
              e = 1;
          }
          break;
        default:
          return -1;
        }
        L->mask[idx3(L, x, y, z)] = e;
      }
  return 0;
}

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
