// kernels_twopass.hpp -- the general two-pass schedule of one time step.
//
// Device layout (all kernels): structure of arrays without halo,
//   pop[(f * Q + i) * sp + s],   s = (x * ny + y) * nz + z,
// sp = the site count padded to a multiple of 64.  The periodic halo exchange
// of the reference (src/dlbc/fields/parallel.d:40-153 on one rank) becomes a
// modulo wrap of the neighbour index; this is exact because, on one rank,
// every halo value the reference reads is the periodic image of a physical
// site (for lattice lengths >= 2).
//
// Side buffer S[f * sp + s]: psi_f at fluid sites (written every step by the
// psi pass); rho_f at solid sites (written once, when populations or mask are
// uploaded; solid populations never change).  The Shan-Chen stencil reads
// psi only at fluid neighbours and the wall term reads rho only at solid
// neighbours (src/dlbc/lb/force.d:273,299), so one buffer serves both.
#pragma once

#include "lattice.hpp"

namespace dlbc {

struct Geo {
  int nx, ny, nz;
  int sites;
  int64_t sp;
  // Block order of the split kernels (kernels_split.hpp split_block): chunk
  // K > 0 hands the XCDs K consecutive blocks each in turn over the first
  // order_main blocks; 0 keeps site order.
  int order_chunk = 0, order_main = 0;
  // s / (n_y n_z) and s / n_z as a multiplication (coords): q = (s m) >> (32 + l)
  // with 2^l >= d and m = ceil(2^(32 + l) / d), which is exact for every
  // s < 2^32 (m d - 2^(32 + l) < d, so s times that excess stays below
  // 2^(32 + l)); a 32-bit division costs some 30 instructions per lane.
  uint64_t div_yz_m = 0, div_z_m = 0;
  int div_yz_l = 0, div_z_l = 0;
};

// The multiplier and shift of division by d (1 <= d <= 2^30) for Geo.
__host__ inline void geo_divisor(uint32_t d, uint64_t &m, int &l) {
  l = 0;
  while ((1ull << l) < d) ++l;
  m = ((1ull << (32 + l)) + d - 1) / d;
}

// Geo's divisors from its extents (after n_x, n_y, n_z are set).
__host__ inline void geo_set_divisors(Geo &g) {
  geo_divisor((uint32_t)g.ny * (uint32_t)g.nz, g.div_yz_m, g.div_yz_l);
  geo_divisor((uint32_t)g.nz, g.div_z_m, g.div_z_l);
}

// floor(s / d) for 0 <= s < 2^31 by Geo's multiplier and shift for d
// (s m < 2^64 since m < 2^33).
__host__ __device__ inline int geo_div(int s, uint64_t m, int l) {
  return (int)(((uint64_t)(uint32_t)s * m) >> (32 + l));
}

struct StepParams {
  double tau[4], omega[4], one_minus_omega[4];
  double gcc[4][4];
  double gwc[4];
  double gacc[3];
  double body[4][3];
  int eq_form, psi_form, sc, has_solid, has_body;
};

__device__ inline void coords(int s, const Geo &g, int &x, int &y, int &z) {
  const int yz = g.ny * g.nz;
  x = geo_div(s, g.div_yz_m, g.div_yz_l);
  const int r = s - x * yz;
  y = geo_div(r, g.div_z_m, g.div_z_l);
  z = r - y * g.nz;
}

__device__ inline int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ inline int site_of(int x, int y, int z, const Geo &g) { return (x * g.ny + y) * g.nz + z; }

// Pull streaming of one site (src/dlbc/lb/advection.d:61-89): for a fluid
// site, n'_i(x) = solid(x - c_i) ? n_bounce(i)(x) : n_i(x - c_i).
template <class L>
__device__ inline void pull_sources(int x, int y, int z, const Geo &g, const uint8_t *mask,
                                    int has_solid, int s, int *src, int *comp) {
#pragma unroll
  for (int i = 0; i < L::q; ++i) {
    const int nb = site_of(wrap(x - L::c[i][0], g.nx), wrap(y - L::c[i][1], g.ny),
                           wrap(z - L::c[i][2], g.nz), g);
    if (has_solid && mask[nb]) {
      src[i] = s;
      comp[i] = L::bounce[i];
    } else {
      src[i] = nb;
      comp[i] = i;
    }
  }
}

// ---- per-site physics, one component at a time ---------------------------
// Both schedules call these, so they produce identical bits.

// sum_i n_i c_i in i order, from 0 (momentum, lb/momentum.d:37-52; the
// per-component sum of precalculateWeightedVelocity, lb/collision.d:271-276).
template <class L>
__device__ inline void first_moment(const double *n, double (&m)[L::d]) {
#pragma unroll
  for (int d = 0; d < L::d; ++d) m[d] = 0.0;
#pragma unroll
  for (int i = 0; i < L::q; ++i)
#pragma unroll
    for (int d = 0; d < L::d; ++d) m[d] += n[i] * static_cast<double>(L::c[i][d]);
}

// addShanChenForcePsi (src/dlbc/lb/force.d:238-309) for component f1 at one
// collidable site, starting from F = 0 (resetForce).  psi(f, i): psi_f at the
// neighbour x + c_i (i = -1: at the site); solid(i): x + c_i is solid;
// rho_solid(f, i): rho_f at the solid neighbour x + c_i.
template <class L, int C, class Psi, class Solid, class RhoSolid>
__device__ inline void sc_force(int f1, double (&F)[L::d], const StepParams &p, Psi psi,
                                Solid solid, RhoSolid rho_solid) {
  constexpr int Q = L::q, D = L::d;
#pragma unroll
  for (int d = 0; d < D; ++d) F[d] = 0.0;
  if (!p.sc) return;
  const double psi1 = psi(f1, -1);
#pragma unroll
  for (int f2 = 0; f2 < C; ++f2) {
    const double cc = p.gcc[f1][f2];
    if (cc == 0.0) continue;
#pragma unroll
    for (int i = 1; i < Q; ++i) {  // Iota!(1, q-1): i = 1 .. Q-1
      // read before the test (every neighbour index is in bounds), so that
      // the reads are not each behind a branch
      const double psi_n = psi(f2, i);
      const double psi2 = solid(i) ? psi(f2, -1) : psi_n;
      const double pf = L::w[i] * psi1 * psi2 * cc;
#pragma unroll
      for (int d = 0; d < D; ++d) F[d] -= pf * static_cast<double>(L::c[i][d]);
    }
  }
  const double wc = p.gwc[f1];
  if (wc != 0.0 && p.has_solid) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      if (solid(i)) {
        const double pf = L::w[i] * psi1 * rho_solid(f1, i) * wc;
#pragma unroll
        for (int d = 0; d < D; ++d) F[d] -= pf * static_cast<double>(L::c[i][d]);
      }
    }
  }
}

// totalDensity of distributeForce / precalculateWeightedVelocity.
template <int C, class Rho>
__device__ inline double total_density(Rho rho) {
  double total = 0.0;
#pragma unroll
  for (int f = 0; f < C; ++f) total += rho(f);
  return total;
}

// BDist2 weighted velocity (lb/collision.d:268-295) from the components'
// first moments mom(f, d) and the total density.
template <class L, int C, class Mom>
__device__ inline void weighted_velocity(Mom mom, double total, const StepParams &p,
                                         double (&wv)[L::d]) {
#pragma unroll
  for (int d = 0; d < L::d; ++d) wv[d] = 0.0;
#pragma unroll
  for (int f = 0; f < C; ++f)
#pragma unroll
    for (int d = 0; d < L::d; ++d) wv[d] += mom(f, d) / p.tau[f];
#pragma unroll
  for (int d = 0; d < L::d; ++d) wv[d] /= total;
}

// distributeForce (forceDistributed == 0, lb/force.d:179-197), the
// per-component body force (a pre-collision hook) and the BGK collision
// (lb/collision.d:90-119) of component f; out(i, value) stores the result.
template <class L, class Out>
__device__ inline void collide_component(int f, const double *n, double rho, double total,
                                         double (&F)[L::d], const double *wv,
                                         const StepParams &p, Out out) {
  constexpr int Q = L::q, D = L::d;
#pragma unroll
  for (int d = 0; d < D; ++d) F[d] += 0.0 * (rho / total);
  if (p.has_body) {
#pragma unroll
    for (int d = 0; d < D; ++d) F[d] += p.body[f][d];
  }
  double dv[D], eq[Q];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    dv[d] = p.tau[f] * (p.gacc[d] + F[d] / rho);
    if (p.eq_form == kBDist2) dv[d] += wv[d];
  }
  eq_dist<L>(p.eq_form, n, dv, rho, eq);
#pragma unroll
  for (int i = 0; i < Q; ++i) out(i, n[i] * p.one_minus_omega[f] + p.omega[f] * eq[i]);
}

// advectFields alone (src/dlbc/lb/advection.d:35-89): A -> B for fluid sites;
// solid sites hold the same populations in both buffers.
template <class L, int C>
__global__ void __launch_bounds__(256) k_advect(const double *__restrict__ A,
                                                double *__restrict__ B,
                                                const uint8_t *__restrict__ mask, Geo g,
                                                int has_solid) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites) return;
  if (has_solid && mask[s]) return;
  int x, y, z;
  coords(s, g, x, y, z);
  int src[L::q], comp[L::q];
  pull_sources<L>(x, y, z, g, mask, has_solid, s, src, comp);
#pragma unroll
  for (int f = 0; f < C; ++f)
#pragma unroll
    for (int i = 0; i < L::q; ++i)
      B[(int64_t)(f * L::q + i) * g.sp + s] = A[(int64_t)(f * L::q + comp[i]) * g.sp + src[i]];
}

// Pass 1: streamed density and psi at every fluid site
// (src/dlbc/lb/density.d:344-353, src/dlbc/lb/force.d:137-147,360-367).
template <class L, int C>
__global__ void __launch_bounds__(256) k_psi(const double *__restrict__ A,
                                             const uint8_t *__restrict__ mask,
                                             double *__restrict__ S, Geo g, int psi_form,
                                             int has_solid) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites) return;
  if (has_solid && mask[s]) return;
  int x, y, z;
  coords(s, g, x, y, z);
  int src[L::q], comp[L::q];
  pull_sources<L>(x, y, z, g, mask, has_solid, s, src, comp);
#pragma unroll
  for (int f = 0; f < C; ++f) {
    double n[L::q];
#pragma unroll
    for (int i = 0; i < L::q; ++i) n[i] = A[(int64_t)(f * L::q + comp[i]) * g.sp + src[i]];
    S[f * g.sp + s] = psi_of(psi_form, density<L>(n));
  }
}

// Pass 2: streaming, Shan-Chen force (src/dlbc/lb/force.d:238-309),
// distributeForce (force.d:179-197), BDist2 weighted velocity
// (src/dlbc/lb/collision.d:156-196) and BGK collision (collision.d:90-119).
template <class L, int C>
__global__ void __launch_bounds__(256) k_collide(const double *__restrict__ A,
                                                 double *__restrict__ B,
                                                 const uint8_t *__restrict__ mask,
                                                 const double *__restrict__ S, Geo g,
                                                 StepParams p) {
  constexpr int Q = L::q, D = L::d;
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites) return;
  if (p.has_solid && mask[s]) return;  // solid populations stay as they are
  int x, y, z;
  coords(s, g, x, y, z);
  double n[C][Q];
  {
    int src[Q], comp[Q];
    pull_sources<L>(x, y, z, g, mask, p.has_solid, s, src, comp);
#pragma unroll
    for (int f = 0; f < C; ++f)
#pragma unroll
      for (int i = 0; i < Q; ++i) n[f][i] = A[(int64_t)(f * Q + comp[i]) * g.sp + src[i]];
  }
  double rho[C];
#pragma unroll
  for (int f = 0; f < C; ++f) rho[f] = density<L>(n[f]);
  int nbr[Q];
  bool solid_nb[Q];
  if (p.sc) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      nbr[i] = site_of(wrap(x + L::c[i][0], g.nx), wrap(y + L::c[i][1], g.ny),
                       wrap(z + L::c[i][2], g.nz), g);
      solid_nb[i] = p.has_solid && mask[nbr[i]];
    }
  }
  double F[C][D];
#pragma unroll
  for (int f = 0; f < C; ++f)
    sc_force<L, C>(
        f, F[f], p, [&](int ff, int i) { return S[ff * g.sp + (i < 0 ? s : nbr[i])]; },
        [&](int i) { return solid_nb[i]; }, [&](int ff, int i) { return S[ff * g.sp + nbr[i]]; });
  const double total = total_density<C>([&](int f) { return rho[f]; });
  double wv[D];
  if (p.eq_form == kBDist2) {
    double mom[C][D];
#pragma unroll
    for (int f = 0; f < C; ++f) first_moment<L>(n[f], mom[f]);
    weighted_velocity<L, C>([&](int f, int d) { return mom[f][d]; }, total, p, wv);
  }
#pragma unroll
  for (int f = 0; f < C; ++f)
    collide_component<L>(f, n[f], rho[f], total, F[f], wv, p, [&](int i, double v) {
      B[(int64_t)(f * Q + i) * g.sp + s] = v;
    });
}

}  // namespace dlbc
