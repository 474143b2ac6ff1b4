// kernels_split.hpp -- two-pass D3Q19 two-component step with one lane per
// (site, component).
//
// The general two-pass schedule (kernels_twopass.hpp) gives a thread a whole
// site: both components' populations (76 doubles) live in its registers and
// every neighbour's solid flag is a separate mask load.  Here a wave holds one
// component of 64 consecutive sites (waves 2k and 2k+1 of a block: the two
// components of the same sites), which halves the registers per thread and
// makes every array base wave-uniform, so loads and stores address with a
// scalar base and a 32-bit offset.  The partner component's density and
// momentum pass through LDS.  The solid neighbourhood is the site's one
// 32-bit solid word; the new populations are written with streaming stores.
//
//   pass 1: pull the 19 streamed populations, density, psi -> S
//   pass 2: pull again, Shan-Chen force from psi in S, velocity shift,
//           BGK collision
//
// When a block's sites lie in one z-row (n_z a multiple of the block's
// sites), pass 2 stages the nine psi rows around the block in LDS with
// coalesced loads issued after the pulls, instead of gathering 19 scattered
// psi values per lane: 2.32 against 2.46 ms per FP64 collide at 256^3, with
// the same bits.
//
// Per-site arithmetic is the shared code of the other schedules (density,
// sc_force, total_density, weighted_velocity, collide_component), so all
// three give identical bits.
#pragma once

#include <type_traits>

#include "kernels_fused.hpp"
#include "kernels_shifted.hpp"

namespace dlbc {

constexpr int kSplitThreads = 256;  // FP64: 4 waves = 128 sites x 2 components
// Shifted storage: one lane per site holding both components (their
// populations sit side by side), 4 waves = 256 sites.
constexpr int kShiftThreads = 256;
constexpr int kShiftSites = 256;

// Grid of a split launch over `sites`: threads / 2 sites per block.  (A grid
// that gave each XCD one contiguous run of sites, an eighth of the lattice,
// measured slower: 3.76 against 3.43 ms per FP64 step at 256^3.)
inline int split_grid(int64_t sites, int threads = kSplitThreads) {
  return (int)((sites + threads / 2 - 1) / (threads / 2));
}

// The block of sites this workgroup takes.  Workgroup b runs on XCD b % 8.
// With g.order_chunk = K the XCDs take turns at K consecutive blocks: the
// j-th workgroup of XCD k takes block ((j / K) 8 + k) K + j % K.  Rows
// neighbouring in y then share an XCD and its L2, and in x too when a plane
// is a whole number of turns (8 K blocks).  Blocks from g.order_main on (the
// part of the grid short of a whole turn) keep site order.
__device__ inline int split_block(const Geo &g) {
  const int b = blockIdx.x;
  const int K = g.order_chunk;
  if (K <= 0 || b >= g.order_main) return b;
  const int j = b >> 3;
  return ((j / K) * 8 + (b & 7)) * K + j % K;
}

// Whether the library picks an XCD block order by itself (an eighth of an
// x-plane per turn) or keeps site order.  Site order: on the 256^3 channel
// (FP64) chunks of 8 to 256 blocks cut the collide pass's L2 misses by a
// quarter (FETCH_SIZE 3.49 -> 2.60 GiB at K = 64) yet ran 2.53-2.74 ms
// against 2.46 ms in site order.
constexpr bool kSplitAutoChunk = false;

// order_main for chunk K over a grid of nb blocks: whole turns of 8 K blocks.
inline int split_order_main(int nb, int K) { return K > 0 ? nb / (8 * K) * (8 * K) : 0; }

// Whether every site of this lane's cache line is solid, for a wave of 64
// consecutive sites whose lines hold LINE sites each (lines aligned to the
// wave: the site stride is a multiple of 64).  Such lines are neither read nor
// written by a collide pass: a solid site keeps its populations.
template <int LINE>
__device__ inline bool line_all_solid(bool solid_site) {
  constexpr uint64_t kLineMask = LINE == 64 ? ~0ull : ((1ull << LINE) - 1ull);
  const uint64_t m = __ballot(solid_site);
  const int first = (threadIdx.x & 63) / LINE * LINE;
  return ((m >> first) & kLineMask) == kLineMask;
}

// Wave-uniform copy of an int: the compiler then keeps it in a scalar
// register, and pointers formed from it stay scalar (the wave index
// threadIdx.x >> 6 is uniform, but the compiler cannot tell).
__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Site and component of this lane (the component is uniform per wave); false
// past the end of the lattice.
__device__ inline bool split_lane(const Geo &g, int &s, int &f) {
  const int w = uniform(threadIdx.x >> 6);
  f = w & 1;
  s = uniform(split_block(g) * (kSplitThreads / 2) + (w >> 1) * 64) + (threadIdx.x & 63);
  return s < g.sites;
}

// Element k of the array at `base` (wave-uniform).  OFF32 (every index of
// the kernel below 2^29): the byte offset is formed in 32 bits, so the access
// takes the scalar base and a 32-bit lane offset instead of a 64-bit lane
// address (one address register and no 64-bit arithmetic per access).
template <bool OFF32, class T>
__device__ inline T *elem(T *base, int64_t k) {
  if constexpr (OFF32) {
    using Byte = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<Byte *>(base) +
                                 (uint32_t)((uint32_t)k * (uint32_t)sizeof(T)));
  } else {
    return base + k;
  }
}
template <bool OFF32, class T>
__device__ inline const T &at(const T *base, int64_t k) {
  return *elem<OFF32>(base, k);
}

// Index offsets from a site to its neighbours along each axis with the
// periodic wrap folded in (m: -1, p: +1), set up once per lane: a neighbour
// index is then the site plus up to three of them, instead of three wraps
// (compare and select each) and two 32-bit multiplies per neighbour.
struct NbrOffsets {
  int xm, xp, ym, yp, zm, zp;
};

__device__ inline NbrOffsets nbr_offsets(const Geo &g, int x, int y, int z) {
  const int plane = g.ny * g.nz;
  NbrOffsets o;
  o.xm = x == 0 ? (g.nx - 1) * plane : -plane;
  o.xp = x == g.nx - 1 ? -(g.nx - 1) * plane : plane;
  o.ym = y == 0 ? (g.ny - 1) * g.nz : -g.nz;
  o.yp = y == g.ny - 1 ? -(g.ny - 1) * g.nz : g.nz;
  o.zm = z == 0 ? g.nz - 1 : -1;
  o.zp = z == g.nz - 1 ? -(g.nz - 1) : 1;
  return o;
}

// Site s + (cx, cy, cz), each component in {-1, 0, 1} (the same index as
// site_of over the wrapped coordinates).
__device__ inline int nbr_site(int s, const NbrOffsets &o, int cx, int cy, int cz) {
  return s + (cx < 0 ? o.xm : (cx > 0 ? o.xp : 0)) + (cy < 0 ? o.ym : (cy > 0 ? o.yp : 0)) +
         (cz < 0 ? o.zm : (cz > 0 ? o.zp : 0));
}

// True when every population index of a lattice (19 sp) is below 2^29.
inline bool split_off32(const Geo &g) { return 19 * g.sp < ((int64_t)1 << 29); }

// The 19 streamed populations of component f (array Af) at site s = (x, y, z)
// whose solid word is sw (pull with bounce-back, src/dlbc/lb/advection.d:61-89).
// Each load takes a wave-uniform base, array min(i, bounce(i)) of Af, and a
// lane offset below 8 sp (bounce partners are at most 7 arrays apart).
template <bool SOLID, bool OFF32, class T>
__device__ inline T pull_one(const T *__restrict__ Af, const Geo &g, int i, int s,
                             const NbrOffsets &o, uint32_t sw) {
  using L = D3Q19;
  const int src = nbr_site(s, o, -L::c[i][0], -L::c[i][1], -L::c[i][2]);
  if (!SOLID) return at<OFF32>(Af + (int64_t)i * g.sp, src);
  const int b = L::bounce[i], mc = i < b ? i : b;
  const bool bb = (sw >> b) & 1u;
  return at<OFF32>(Af + (int64_t)mc * g.sp, (int64_t)((bb ? b : i) - mc) * g.sp + (bb ? s : src));
}

template <bool SOLID, bool OFF32>
__device__ inline void split_pull(const double *__restrict__ Af, const Geo &g, int s,
                                  const NbrOffsets &o, uint32_t sw, double *n) {
#pragma unroll
  for (int i = 0; i < 19; ++i) n[i] = pull_one<SOLID, OFF32>(Af, g, i, s, o, sw);
}

// Coordinates of site s of this lane, where the wave's sites are consecutive
// from s - (lane): the wave's first site is decomposed once with
// wave-uniform (scalar) arithmetic and each lane adds its offset with at
// most one carry when n_z >= 64, instead of two 32-bit divisions per lane.
__device__ inline void lane_coords(int s, const Geo &g, int &x, int &y, int &z) {
  if (g.nz < 64) {
    coords(s, g, x, y, z);
    return;
  }
  const int lane = threadIdx.x & 63;
  int x0, y0, z0;
  coords(uniform(s - lane), g, x0, y0, z0);
  x = x0;
  y = y0;
  z = z0 + lane;
  if (z >= g.nz) {
    z -= g.nz;
    if (++y == g.ny) {
      y = 0;
      ++x;
    }
  }
}

// psi rows of a block of SB consecutive sites that lie in one z-row (n_z a
// multiple of SB), staged in LDS: the nine rows (x + dx, y + dy), dx, dy in
// {-1, 0, 1}, from z0 - 1 to z0 + SB (periodic), of both components.  Every
// neighbour psi (and the solid density in the same array) of the block is in
// them.  The rows are read with 16-byte loads (four FP32 or two FP64 values
// per lane, aligned: n_z, z0 and the component stride are multiples of 64),
// in units of one row and one 64-lane pass, dealt round-robin to the waves:
// a wave-uniform row base and consecutive z per lane, no wrap inside the
// block's span.  The 36 end values (z0 - 1 and z0 + SB of each row) take one
// pass of wave 0.  The loads go out first and the LDS stores follow, so that
// the loads overlap the population pulls issued before; the caller's barrier
// publishes the rows.
template <int SB, int THREADS, bool OFF32, class T>
__device__ inline void stage_psi_rows(const T *__restrict__ S, const Geo &g, int s0,
                                      T (*rows)[9][SB + 2]) {
  constexpr int VEC = 16 / (int)sizeof(T);  // values per 16-byte load
  typedef T Vec __attribute__((ext_vector_type(VEC)));
  constexpr int NW = THREADS / 64, NR = 18, PASS = SB / (64 * VEC);
  constexpr int NU = NR * PASS, UPW = (NU + NW - 1) / NW;  // units, per wave
  static_assert(SB % (64 * VEC) == 0, "staged rows take whole lane passes");
  int x0, y0, z0;
  coords(uniform(s0), g, x0, y0, z0);
  const int w = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // element offset of row pr: component pr / 9, row (dx, dy) = pr % 9 (the
  // wrap is arithmetic: a select among three coordinates became a lookup
  // table in scratch memory)
  auto pair_base = [&](int pr) {
    const int row = pr % 9;
    const int xx = wrap(x0 + row / 3 - 1, g.nx), yy = wrap(y0 + row % 3 - 1, g.ny);
    return (int64_t)(pr / 9) * g.sp + (int64_t)(xx * g.ny + yy) * g.nz;
  };
  Vec v[UPW];
#pragma unroll
  for (int q = 0; q < UPW; ++q) {
    const int u = w + q * NW, pr = (u < NU ? u : 0) / PASS, k = (u < NU ? u : 0) % PASS;
    const T *base = S + pair_base(pr) + z0;
    v[q] = u < NU ? *reinterpret_cast<const Vec *>(elem<OFF32>(base, VEC * (lane + 64 * k)))
                  : Vec(T(0));
  }
  // the two ends of every row's span: lane 2 pr + side of wave 0
  const bool edge = w == 0 && lane < 2 * NR;
  const int epr = lane >> 1, side = lane & 1;
  T ve = T(0);
  if (edge) {
    const int ze = side ? (z0 + SB == g.nz ? 0 : z0 + SB) : (z0 == 0 ? g.nz - 1 : z0 - 1);
    ve = at<OFF32>(S, pair_base(epr) + ze);
  }
#pragma unroll
  for (int q = 0; q < UPW; ++q) {
    const int u = w + q * NW;
    if (u < NU) {
      T *row = rows[u / PASS / 9][u / PASS % 9] + 1 + VEC * (lane + 64 * (u % PASS));
#pragma unroll
      for (int j = 0; j < VEC; ++j) row[j] = v[q][j];
    }
  }
  if (edge) rows[epr / 9][epr % 9][side ? SB + 1 : 0] = ve;
}

// The staged value of component ff at x + c_i for site k of the block (i < 0: x).
template <int SB, class T>
__device__ inline T staged_psi(const T (*rows)[9][SB + 2], int ff, int i, int k) {
  using L = D3Q19;
  if (i < 0) return rows[ff][4][k + 1];
  return rows[ff][(L::c[i][0] + 1) * 3 + L::c[i][1] + 1][k + 1 + L::c[i][2]];
}

template <bool SOLID, bool OFF32>
__global__ void __launch_bounds__(kSplitThreads)
    k_split_psi(const double *__restrict__ A, const uint32_t *__restrict__ nbw,
                double *__restrict__ S, Geo g, int psi_form) {
  int s, f;
  if (!split_lane(g, s, f)) return;
  const uint32_t sw = SOLID ? nbw[s] : 0u;
  if (sw >> 31) return;  // solid: S keeps the site's density
  int x, y, z;
  lane_coords(s, g, x, y, z);
  double n[19];
  split_pull<SOLID, OFF32>(A + (int64_t)f * 19 * g.sp, g, s, nbr_offsets(g, x, y, z), sw, n);
  *elem<OFF32>(S + (int64_t)f * g.sp, s) = psi_of(psi_form, density<D3Q19>(n));
}

// addShanChenForcePsi (src/dlbc/lb/force.d:238-309) for two components whose
// self-couplings gcc[f][f] are zero: the reference skips the f2 == f1 term,
// so only the other component contributes and every lane of a wave runs the
// same loop (sc_force with its lane-dependent skip runs both halves masked).
// Same operations in the same order as sc_force, hence the same bits.
template <class Psi, class Solid, class RhoSolid>
__device__ inline void sc_force_offdiag(int f1, double (&F)[3], const StepParams &p, Psi psi,
                                        Solid solid, RhoSolid rho_solid) {
  using L = D3Q19;
#pragma unroll
  for (int d = 0; d < 3; ++d) F[d] = 0.0;
  if (!p.sc) return;
  const double psi1 = psi(f1, -1);
  const int f2 = f1 ^ 1;
  const double cc = p.gcc[f1][f2];
  if (cc != 0.0) {
#pragma unroll
    for (int i = 1; i < L::q; ++i) {
      const double psi_n = psi(f2, i);  // read before the test, as in sc_force
      const double psi2 = solid(i) ? psi(f2, -1) : psi_n;
      const double pf = L::w[i] * psi1 * psi2 * cc;
#pragma unroll
      for (int d = 0; d < 3; ++d) F[d] -= pf * static_cast<double>(L::c[i][d]);
    }
  }
  const double wc = p.gwc[f1];
  if (wc != 0.0 && p.has_solid) {
#pragma unroll
    for (int i = 0; i < L::q; ++i) {
      if (solid(i)) {
        const double pf = L::w[i] * psi1 * rho_solid(f1, i) * wc;
#pragma unroll
        for (int d = 0; d < 3; ++d) F[d] -= pf * static_cast<double>(L::c[i][d]);
      }
    }
  }
}

// collide_component (distributeForce, body force, velocity shift, eqDist and
// BGK relaxation) for D3Q19 with every equilibrium value stored as soon as it
// is formed, instead of through a 19-value array: the same operations in the
// same order (lattice.hpp eq_dist, kernels_twopass.hpp collide_component), so
// the same bits, with fewer registers live at the stores.
template <class Out>
__device__ inline void collide_d3q19_streamed(int f, const double *n, double rho, double total,
                                              double (&F)[3], const double *wv,
                                              const StepParams &p, Out out) {
  using L = D3Q19;
#pragma unroll
  for (int d = 0; d < 3; ++d) F[d] += 0.0 * (rho / total);
  if (p.has_body) {
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] += p.body[f][d];
  }
  double dv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dv[d] = p.tau[f] * (p.gacc[d] + F[d] / rho);
    if (p.eq_form == kBDist2) dv[d] += wv[d];
  }
  const double omw = p.one_minus_omega[f], om = p.omega[f];
  auto relax = [&](int i, double eq) { out(i, n[i] * omw + om * eq); };
  if (p.eq_form != kBDist2) {  // eqdist.d:82-100
    double pv[3], v[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) pv[j] = 0.0;
#pragma unroll
    for (int i = 0; i < 19; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) pv[j] += n[i] * static_cast<double>(L::c[i][j]);
#pragma unroll
    for (int j = 0; j < 3; ++j) pv[j] /= rho;
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = dv[j] + pv[j];
    double vdotv = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) vdotv += v[j] * v[j];
    constexpr double css = kCss;
    constexpr double css2 = 2.0 * css;
    constexpr double css2p2 = css2 * css;
    constexpr double css6p3 = 3.0 * css2p2 * css;
    constexpr double r_css = 1.0 / css, r_css2 = 1.0 / css2, r_css2p2 = 1.0 / css2p2,
                     r_css6p3 = 1.0 / css6p3;
    const double vv_css2 = div_const(vdotv, css2, r_css2);
    const bool third = p.eq_form == kThirdOrder;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      double vc = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) vc += v[j] * static_cast<double>(L::c[i][j]);
      double eq;
      if (!third)
        eq = rho * L::w[i] *
             (1.0 + div_const(vc, css, r_css) + div_const(vc * vc, css2p2, r_css2p2) - vv_css2);
      else
        eq = rho * L::w[i] *
             (1.0 + div_const(vc, css, r_css) + div_const(vc * vc, css2p2, r_css2p2) - vv_css2 +
              div_const(vc * vc * vc, css6p3, r_css6p3) - div_const(vdotv * vc, css2p2, r_css2p2));
      relax(i, eq);
    }
    return;
  }
  // BDist2, eqdist.d:102-184
  const double aN1 = 1.0 / 36.0, aN0 = 1.0 / 3.0;
  const double haN1 = 0.5 * aN1, aN1_6 = aN1 / 6.0;
  const double T = 1.0 / 3.0, holdit = 1.0 / T;
  const double tux = dv[0], tuy = dv[1], tuz = dv[2];
  const double tuxt = tux * holdit, tuyt = tuy * holdit, tuzt = tuz * holdit;
  const double uke = 0.5 * (tux * tuxt + tuy * tuyt + tuz * tuzt);
  const double cst = 1.0 - uke;
  const double const1 = aN1 * cst;
  auto pair = [&](double cdotu, int ip, int im, bool axis) {
    const double even = const1 + haN1 * cdotu * cdotu;
    const double odd = const1 * cdotu + aN1_6 * cdotu * cdotu * cdotu;
    if (axis) {
      relax(ip, 2.0 * (even + odd) * rho);
      relax(im, 2.0 * (even - odd) * rho);
    } else {
      relax(ip, (even + odd) * rho);
      relax(im, (even - odd) * rho);
    }
  };
  relax(0, aN0 * cst * rho);
  pair(tuxt, 1, 2, true);
  pair(tuyt, 3, 4, true);
  pair(tuzt, 5, 6, true);
  pair(tuxt + tuyt, 7, 14, false);
  pair(tuxt - tuyt, 13, 8, false);
  pair(tuxt + tuzt, 11, 18, false);
  pair(tuxt - tuzt, 12, 17, false);
  pair(tuyt + tuzt, 9, 16, false);
  pair(tuyt - tuzt, 15, 10, false);
}

// STAGED (n_z a multiple of the block's sites): the psi rows come from LDS.
// A template argument rather than a runtime test, so that the body holds one
// psi source: with both, every psi read sat behind a branch and a full wait.
template <bool SOLID, bool OFFDIAG, bool OFF32, bool STAGED>
__global__ void __launch_bounds__(kSplitThreads)
    k_split_collide(const double *__restrict__ A, double *__restrict__ B,
                    const uint32_t *__restrict__ nbw, const double *__restrict__ S, Geo g,
                    StepParams p) {
  using L = D3Q19;
  constexpr int C = 2;
  constexpr int SB = kSplitThreads / 2;
  // density and first moment of both components of the block's sites
  __shared__ double xch[2][4][SB];
  // psi rows of the block when it lies in one z-row
  __shared__ double rows[STAGED ? 2 : 1][9][SB + 2];
  constexpr bool staged = STAGED;
  int s, f;
  const bool inside = split_lane(g, s, f);
  const int k = ((threadIdx.x >> 7) << 6) | (threadIdx.x & 63);  // site within the block
  const uint32_t sw = !inside ? (1u << 31) : (SOLID ? nbw[s] : 0u);
  NbrOffsets o{};
  double n[19];
  double rho = 0.0, mom[3] = {0.0, 0.0, 0.0};
  const bool bdist2 = p.eq_form == kBDist2;
  // A solid site's lane copies its own populations (the reference copies them
  // every step, advection.d:84-86): with every neighbour taken as solid, the
  // pull loads population bounce(i) of the site itself into n[i], in the same
  // load instructions as the fluid lanes' pulls, and the store below writes
  // n[bounce(i)] back to population i.  Whole cache lines are then written;
  // skipping the solid lanes left partially written lines, which cost 7 % of
  // the pass on the 256^3 channel.
  const bool solid_site = inside && (sw >> 31);
  const bool active = inside && !(SOLID && line_all_solid<16>(solid_site));
  if (active) {
    int x, y, z;
    lane_coords(s, g, x, y, z);
    o = nbr_offsets(g, x, y, z);
    split_pull<SOLID, OFF32>(A + (int64_t)f * 19 * g.sp, g, s, o, solid_site ? 0x7ffffu : sw,
                             n);
  }
  if (staged) stage_psi_rows<SB, kSplitThreads, OFF32>(S, g, split_block(g) * SB, rows);
  if (active) {
    rho = density<L>(n);
    if (bdist2) first_moment<L>(n, mom);
  }
  xch[f][0][k] = rho;
  if (bdist2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) xch[f][1 + d][k] = mom[d];
  }
  __syncthreads();
  if (!active) return;
  // (a solid lane runs the arithmetic below on its own populations, whose
  // result it does not store: no divergent branch in a mixed wave)
  const double rho_partner = xch[f ^ 1][0][k];
  double mom_p[3];
  if (bdist2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) mom_p[d] = xch[f ^ 1][1 + d][k];
  }
  auto nbr = [&](int i) { return nbr_site(s, o, L::c[i][0], L::c[i][1], L::c[i][2]); };
  double F[3];
  auto psi = [&](int ff, int i) {
    if (staged) return staged_psi<SB>(rows, ff, i, k);
    return at<OFF32>(S + ff * g.sp, i < 0 ? s : nbr(i));
  };
  auto solid = [&](int i) { return SOLID && ((sw >> i) & 1u); };
  auto rho_solid = [&](int ff, int i) {  // S holds the density at solid sites
    if (staged) return staged_psi<SB>(rows, ff, i, k);
    return at<OFF32>(S + ff * g.sp, nbr(i));
  };
  if constexpr (OFFDIAG)
    sc_force_offdiag(f, F, p, psi, solid, rho_solid);
  else
    sc_force<L, C>(f, F, p, psi, solid, rho_solid);
  const double total = total_density<C>([&](int ff) { return ff == f ? rho : rho_partner; });
  double wv[3];
  if (bdist2)
    weighted_velocity<L, C>([&](int ff, int d) { return ff == f ? mom[d] : mom_p[d]; }, total,
                            p, wv);
  double *__restrict__ Bf = B + (int64_t)f * 19 * g.sp;
  collide_d3q19_streamed(f, n, rho, total, F, wv, p, [&](int i, double v) {
    __builtin_nontemporal_store(solid_site ? n[L::bounce[i]] : v,
                                elem<OFF32>(Bf + (int64_t)i * g.sp, s));
  });
}

// ---- shifted storage (FP32 / FP16 deviations, FP32 arithmetic) -------------
// The same two passes; psi goes through an FP32 array P (the values are FP32
// in every schedule) and the per-site arithmetic is shifted_collide_site, as
// in the tiled kernel, so a lattice gives the same bits under both schedules.
// S still gives the density at solid neighbours (wall force).

// Both components' streamed populations at site s (pull with bounce-back,
// src/dlbc/lb/advection.d:61-89): one 8-byte (FP32) or 4-byte (FP16) load
// per population and lane, from a wave-uniform base, array min(i, bounce(i)),
// at a lane offset below 16 sp elements.
template <class Store, bool SOLID, bool OFF32>
__device__ inline void ssplit_pull(const Store *__restrict__ A, const Geo &g, int s,
                                   const NbrOffsets &o, uint32_t sw, float (&n0)[19],
                                   float (&n1)[19]) {
  using L = D3Q19;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const int src = nbr_site(s, o, -L::c[i][0], -L::c[i][1], -L::c[i][2]);
    const Store *p;
    if (!SOLID) {
      p = elem<OFF32>(A + 2 * (int64_t)i * g.sp, 2 * (int64_t)src);
    } else {
      const int b = L::bounce[i], mc = i < b ? i : b;
      const bool bb = (sw >> b) & 1u;
      p = elem<OFF32>(A + 2 * (int64_t)mc * g.sp,
                      2 * ((int64_t)((bb ? b : i) - mc) * g.sp + (bb ? s : src)));
    }
    pair_load(p, n0[i], n1[i]);
  }
}

// The sites of block blk of a lattice (kShiftSites from blk * kShiftSites),
// one per lane.
__device__ inline int shift_site(int blk) {
  return uniform(blk * kShiftSites + (threadIdx.x & ~63)) + (threadIdx.x & 63);
}

// Pass 1 for block blk of a lattice: psi of both components into P.
template <class Store, bool SOLID, bool OFF32>
__device__ inline void ssplit_psi_block(const Store *__restrict__ A,
                                        const uint32_t *__restrict__ nbw,
                                        float *__restrict__ P, const Geo &g,
                                        const ShiftParams &p, int blk) {
  const int s = shift_site(blk);
  if (s >= g.sites) return;
  const uint32_t sw = SOLID ? nbw[s] : 0u;
  if (sw >> 31) return;  // solid: psi is not read there
  int x, y, z;
  lane_coords(s, g, x, y, z);
  float n0[19], n1[19];
  ssplit_pull<Store, SOLID, OFF32>(A, g, s, nbr_offsets(g, x, y, z), sw, n0, n1);
  *elem<OFF32>(P, s) = psi_f(p.psi_form, p.rho0[0] + shifted_drho(n0));
  *elem<OFF32>(P + g.sp, s) = psi_f(p.psi_form, p.rho0[1] + shifted_drho(n1));
}

// Pass 2 for block blk of a lattice; rows: LDS for the block's psi rows when
// STAGED (n_z a multiple of kShiftSites).  STAGED is a template argument, not
// a runtime test: with both psi sources in one body the compiler merged the
// LDS and global reads into flat loads, each waiting for all memory traffic
// of the wave.  A lane collides component 0 (keeping the result in place of
// its populations) and then component 1, and stores each population of both
// components with one access.
template <class Store, bool SOLID, bool OFF32, bool STAGED = false>
__device__ inline void ssplit_collide_block(const Store *__restrict__ A, Store *__restrict__ B,
                                            const uint32_t *__restrict__ nbw,
                                            const float *__restrict__ P, const Geo &g,
                                            const ShiftParams &p, int blk,
                                            float (*rows)[9][kShiftSites + 2]) {
  using L = D3Q19;
  constexpr int SB = kShiftSites;
  constexpr bool staged = STAGED;
  const int s = shift_site(blk);
  const int k = threadIdx.x;  // site within the block
  const uint32_t sw = s >= g.sites ? (1u << 31) : (SOLID ? nbw[s] : 0u);
  NbrOffsets o{};
  float n0[19], n1[19];
  // a solid site's lane copies its own populations, as in the FP64 pass: the
  // pull with every neighbour taken as solid loads population bounce(i) of the
  // site into n[i], and the store writes n[bounce(i)] back to population i
  const bool inside = s < g.sites, solid_site = inside && (sw >> 31);
  // a 128-byte line holds 16 (FP32) or 32 (FP16) sites' pairs
  const bool active =
      inside && !(SOLID && line_all_solid<128 / (2 * (int)sizeof(Store))>(solid_site));
  if (active) {
    int x, y, z;
    lane_coords(s, g, x, y, z);
    o = nbr_offsets(g, x, y, z);
    ssplit_pull<Store, SOLID, OFF32>(A, g, s, o, solid_site ? 0x7ffffu : sw, n0, n1);
  }
  if (staged) {
    stage_psi_rows<SB, kShiftThreads, OFF32>(P, g, blk * SB, rows);
    __syncthreads();
  }
  if (!active) return;
  const ShiftLane pl0 = shift_lane(p, 0), pl1 = shift_lane(p, 1);
  const float rho0 = pl0.rho0 + shifted_drho(n0), rho1 = pl1.rho0 + shifted_drho(n1);
  float j0[3], j1[3];
  shifted_moment(n0, j0);
  shifted_moment(n1, j1);
  auto nbr = [&](int i) { return nbr_site(s, o, L::c[i][0], L::c[i][1], L::c[i][2]); };
  auto psi = [&](int ff, int i) {
    if (staged) return staged_psi<SB>(rows, ff, i, k);
    return at<OFF32>(P + ff * g.sp, i < 0 ? s : nbr(i));
  };
  // both forces first (P holds the density at solid sites: wall force), so
  // the psi reads are consumed before the populations of a relaxation
  float F0[3], F1[3];
  shifted_force<SOLID>(0, sw, p, pl0, psi, [&](int i, bool) { return psi(0, i); }, F0);
  shifted_force<SOLID>(1, sw, p, pl1, psi, [&](int i, bool) { return psi(1, i); }, F1);
  shifted_relax(n0, rho0, rho1, j0, j1, F0, p, pl0,
                [&](int i, float v) { n0[i] = solid_site ? n0[i] : v; });
  shifted_relax(n1, rho1, rho0, j1, j0, F1, p, pl1, [&](int i, float v) {
    const int b = L::bounce[i];
    pair_store_nt<Store>(elem<OFF32>(B + 2 * (int64_t)i * g.sp, 2 * (int64_t)s),
                         solid_site ? n0[b] : n0[i], solid_site ? n1[b] : v);
  });
}

template <class Store, bool SOLID, bool OFF32>
__global__ void __launch_bounds__(kShiftThreads)
    k_ssplit_psi(const Store *__restrict__ A, const uint32_t *__restrict__ nbw,
                 float *__restrict__ P, Geo g, ShiftParams p) {
  ssplit_psi_block<Store, SOLID, OFF32>(A, nbw, P, g, p, split_block(g));
}

// At most 128 VGPRs: four waves per SIMD.
template <class Store, bool SOLID, bool OFF32, bool STAGED>
__global__ void __launch_bounds__(kShiftThreads) __attribute__((amdgpu_waves_per_eu(4)))
    k_ssplit_collide(const Store *__restrict__ A, Store *__restrict__ B,
                     const uint32_t *__restrict__ nbw, const float *__restrict__ P, Geo g,
                     ShiftParams p) {
  __shared__ float rows[STAGED ? 2 : 1][9][kShiftSites + 2];
  ssplit_collide_block<Store, SOLID, OFF32, STAGED>(A, B, nbw, P, g, p, split_block(g), rows);
}

// Blocks of kShiftSites sites over a lattice.
inline int shift_grid(int64_t sites) { return (int)((sites + kShiftSites - 1) / kShiftSites); }

template <class Store, bool OFF32>
inline void launch_ssplit_impl(const Store *a, Store *b, const uint32_t *nbw, float *P,
                               const Geo &g, const ShiftParams &p, hipStream_t stream,
                               int pass) {
  const int nb = shift_grid(g.sites);
  const dim3 grid(nb), block(kShiftThreads);
  if (pass == 0) {
    if (p.has_solid)
      hipLaunchKernelGGL((k_ssplit_psi<Store, true, OFF32>), grid, block, 0, stream, a, nbw, P,
                         g, p);
    else
      hipLaunchKernelGGL((k_ssplit_psi<Store, false, OFF32>), grid, block, 0, stream, a, nbw, P,
                         g, p);
  } else {
    const bool staged = g.nz % kShiftSites == 0;
    if (p.has_solid && staged)
      hipLaunchKernelGGL((k_ssplit_collide<Store, true, OFF32, true>), grid, block, 0, stream, a,
                         b, nbw, P, g, p);
    else if (p.has_solid)
      hipLaunchKernelGGL((k_ssplit_collide<Store, true, OFF32, false>), grid, block, 0, stream,
                         a, b, nbw, P, g, p);
    else if (staged)
      hipLaunchKernelGGL((k_ssplit_collide<Store, false, OFF32, true>), grid, block, 0, stream,
                         a, b, nbw, P, g, p);
    else
      hipLaunchKernelGGL((k_ssplit_collide<Store, false, OFF32, false>), grid, block, 0, stream,
                         a, b, nbw, P, g, p);
  }
}

template <class Store>
inline void launch_ssplit_d3q19(const void *A, void *B, const uint32_t *nbw, float *P,
                                const Geo &g, const ShiftParams &p, hipStream_t stream,
                                int pass) {
  const Store *a = static_cast<const Store *>(A);
  Store *b = static_cast<Store *>(B);
  if (split_off32(g))
    launch_ssplit_impl<Store, true>(a, b, nbw, P, g, p, stream, pass);
  else
    launch_ssplit_impl<Store, false>(a, b, nbw, P, g, p, stream, pass);
}

template <bool OFF32, bool STAGED>
inline void launch_split_collide(const double *A, double *B, const uint32_t *nbw,
                                 const double *S, const Geo &g, const StepParams &p,
                                 hipStream_t stream, bool solid, bool offdiag) {
  const dim3 grid(split_grid(g.sites)), block(kSplitThreads);
  if (solid && offdiag)
    hipLaunchKernelGGL((k_split_collide<true, true, OFF32, STAGED>), grid, block, 0, stream, A,
                       B, nbw, S, g, p);
  else if (solid)
    hipLaunchKernelGGL((k_split_collide<true, false, OFF32, STAGED>), grid, block, 0, stream, A,
                       B, nbw, S, g, p);
  else if (offdiag)
    hipLaunchKernelGGL((k_split_collide<false, true, OFF32, STAGED>), grid, block, 0, stream, A,
                       B, nbw, S, g, p);
  else
    hipLaunchKernelGGL((k_split_collide<false, false, OFF32, STAGED>), grid, block, 0, stream,
                       A, B, nbw, S, g, p);
}

template <bool OFF32>
inline void launch_split_impl(const double *A, double *B, const uint32_t *nbw, double *S,
                              const Geo &g, const StepParams &p, hipStream_t stream, int pass) {
  const int nb = split_grid(g.sites);
  const dim3 grid(nb), block(kSplitThreads);
  if (pass == 0) {
    if (p.has_solid)
      hipLaunchKernelGGL((k_split_psi<true, OFF32>), grid, block, 0, stream, A, nbw, S, g,
                         p.psi_form);
    else
      hipLaunchKernelGGL((k_split_psi<false, OFF32>), grid, block, 0, stream, A, nbw, S, g,
                         p.psi_form);
    return;
  }
  const bool offdiag = p.gcc[0][0] == 0.0 && p.gcc[1][1] == 0.0;
  if (g.nz % (kSplitThreads / 2) == 0)
    launch_split_collide<OFF32, true>(A, B, nbw, S, g, p, stream, p.has_solid != 0, offdiag);
  else
    launch_split_collide<OFF32, false>(A, B, nbw, S, g, p, stream, p.has_solid != 0, offdiag);
}

inline void launch_split_d3q19(const double *A, double *B, const uint32_t *nbw, double *S,
                               const Geo &g, const StepParams &p, hipStream_t stream,
                               int pass) {
  if (split_off32(g))
    launch_split_impl<true>(A, B, nbw, S, g, p, stream, pass);
  else
    launch_split_impl<false>(A, B, nbw, S, g, p, stream, pass);
}

}  // namespace dlbc
