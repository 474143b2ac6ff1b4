// kernels_fused.hpp -- single-pass D3Q19 two-component Shan-Chen step.
//
// One launch per time step reads every population once and writes it once.
// The collision at x needs psi at x's 18 neighbours, i.e. the streamed
// densities of sites up to two lattice spacings away; the kernel satisfies
// that dependency inside a workgroup:
//
//   * a workgroup owns a TY x 32 tile of the (y, z) plane and marches along x
//     over a chunk [xs, xe) of planes.  Wave w holds tile row y0 + w; lane l
//     holds site z0 + (l & 31) and fluid component l >> 5, so the two
//     components of a site are partner lanes (l, l ^ 32) of one wave and
//     exchange densities and momenta by a cross-lane shuffle;
//   * for the plane ahead (x+1) every lane pulls its site's 19 streamed
//     populations of its component into registers and computes psi;
//   * the psi of the one-site ring around the tile is computed from staged
//     pulls: the ring's streamed populations for plane x+2 are loaded into an
//     LDS stage (double buffered) one iteration before they are reduced;
//   * psi lives in an LDS ring of four planes, so the collision of plane x
//     reads no psi from memory; one barrier per plane separates the
//     producers of psi(x+1) from the collision of x;
//   * solids enter through one 32-bit word per site (bit i: x + c_i is solid,
//     bit 31: x is solid) precomputed when the mask is uploaded: the pull takes
//     the bounce-back population where bit bounce(i) is set, the Shan-Chen
//     stencil substitutes psi(x) where bit i is set;
//   * the collision of plane x uses the registers loaded one iteration
//     earlier and writes the new populations.
// Redundant work: the ring pulls ((TY+2)*34 - 32*TY sites per plane, per tile)
// and the two prologue planes per chunk.
//
// Per-site arithmetic is the same code as the two-pass schedule (sc_force,
// total_density, weighted_velocity, collide_component), so both schedules
// give identical bits.
#pragma once

#include "kernels_twopass.hpp"

namespace dlbc {

// D3Q19 velocity components and bounce index decoded from packed constants
// (for a population index that is not a compile-time constant).
__host__ __device__ constexpr int d3q19_c(int i, int axis) {
  // bit i of P (c = +1) and M (c = -1), connectivity.d:401-418
  constexpr uint32_t PX = (1u << 1) | (1u << 7) | (1u << 11) | (1u << 12) | (1u << 13);
  constexpr uint32_t MX = (1u << 2) | (1u << 8) | (1u << 14) | (1u << 17) | (1u << 18);
  constexpr uint32_t PY = (1u << 3) | (1u << 7) | (1u << 8) | (1u << 9) | (1u << 15);
  constexpr uint32_t MY = (1u << 4) | (1u << 10) | (1u << 13) | (1u << 14) | (1u << 16);
  constexpr uint32_t PZ = (1u << 5) | (1u << 9) | (1u << 10) | (1u << 11) | (1u << 17);
  constexpr uint32_t MZ = (1u << 6) | (1u << 12) | (1u << 15) | (1u << 16) | (1u << 18);
  const uint32_t p = axis == 0 ? PX : (axis == 1 ? PY : PZ);
  const uint32_t m = axis == 0 ? MX : (axis == 1 ? MY : MZ);
  return (int)((p >> i) & 1u) - (int)((m >> i) & 1u);
}

__host__ __device__ constexpr int d3q19_bounce(int i) {
  // 5-bit fields of D3Q19::bounce, packed
  constexpr uint64_t LO = (0ull) | (2ull << 5) | (1ull << 10) | (4ull << 15) | (3ull << 20) |
                          (6ull << 25) | (5ull << 30) | (14ull << 35) | (13ull << 40) |
                          (16ull << 45) | (15ull << 50) | (18ull << 55);
  constexpr uint64_t HI = (17ull) | (8ull << 5) | (7ull << 10) | (10ull << 15) | (9ull << 20) |
                          (12ull << 25) | (11ull << 30);
  return i < 12 ? (int)((LO >> (5 * i)) & 31u) : (int)((HI >> (5 * (i - 12))) & 31u);
}

constexpr bool d3q19_decoders_match() {
  for (int i = 0; i < 19; ++i) {
    if (d3q19_bounce(i) != D3Q19::bounce[i]) return false;
    for (int a = 0; a < 3; ++a)
      if (d3q19_c(i, a) != D3Q19::c[i][a]) return false;
  }
  return true;
}
static_assert(d3q19_decoders_match(), "packed D3Q19 tables disagree with the connectivity");

template <int TY>
struct FusedShape {
  static constexpr int TZ = 32;
  static constexpr int threads = 64 * TY;
  static constexpr int EY = TY + 2, EZ = TZ + 2;  // psi region (ring 1)
  static constexpr int ring = EY * EZ - TY * TZ;
  static constexpr int ring_items = ring * 2 * 19;
};

// Ring index r -> coordinates (ey, ez) in the psi region: the two full rows
// first (contiguous along z), then the two columns.
template <class S>
__device__ inline void ring_coord(int r, int &ey, int &ez) {
  if (r < 2 * S::EZ) {
    ey = r < S::EZ ? 0 : S::EY - 1;
    ez = r < S::EZ ? r : r - S::EZ;
  } else {
    const int k = r - 2 * S::EZ;  // 0 .. 2*TY-1
    ey = 1 + (k >> 1);
    ez = (k & 1) ? S::EZ - 1 : 0;
  }
}

__device__ inline int wrapf(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ inline double shfl_xor32(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __shfl_xor((int)(b & 0xffffffff), 32);
  const int hi = __shfl_xor((int)(b >> 32), 32);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

template <int TY, bool SOLID>
__global__ void __launch_bounds__(64 * TY)
    k_fused_d3q19(const double *__restrict__ A, double *__restrict__ B,
                  const uint32_t *__restrict__ nbw, const double *__restrict__ S, Geo g,
                  StepParams p, int chunk, int ntiles_y, int ntiles_z, int nchunks) {
  using L = D3Q19;
  using Sh = FusedShape<TY>;
  constexpr int Q = 19, C = 2, TZ = Sh::TZ;
  __shared__ double psi_lds[4][C][Sh::EY * Sh::EZ];
  __shared__ double stage[2][Sh::ring_items];

  // XCD-aware block order: hardware places block b on XCD b % 8; consecutive
  // logical tiles (z-neighbours, then y-neighbours) are given to one XCD.
  const int nblocks = gridDim.x;
  const int b = blockIdx.x;
  const int logical = (nblocks % 8 == 0) ? (b % 8) * (nblocks / 8) + b / 8 : b;
  const int tz = logical % ntiles_z;
  const int ty = (logical / ntiles_z) % ntiles_y;
  const int chunk_id = logical / (ntiles_z * ntiles_y);
  if (chunk_id >= nchunks) return;  // uniform per block
  const int xs = chunk_id * chunk;
  const int xe = min(xs + chunk, g.nx);
  const int y0 = ty * TY, z0 = tz * TZ;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int ly = t >> 6;  // wave index = tile row
  const int lz = lane & 31;
  const int f = lane >> 5;  // component
  const int y = y0 + ly, z = z0 + lz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t sp = g.sp;
  const int plane = ny * nz;
  const double *__restrict__ Af = A + (int64_t)f * Q * sp;
  double *__restrict__ Bf = B + (int64_t)f * Q * sp;

  auto gsite = [&](int x, int yy, int zz) { return (x * ny + yy) * nz + zz; };
  auto pslot = [](int x) { return x & 3; };
  // Solid-neighbour word of a site (bit i: x + c_i is solid; bit 31: the
  // site itself), precomputed from the mask when it is uploaded.
  auto solid_word = [&](int x) -> uint32_t {
    if constexpr (SOLID) {
      return nbw[gsite(wrapf(x, nx), y, z)];
    } else {
      return 0u;
    }
  };
  // Streamed populations of the own site and component at plane x (pull with
  // bounce-back, src/dlbc/lb/advection.d:61-89): the upstream site x - c_i is
  // solid iff bit bounce(i) of the solid word is set.
  const int ym = wrapf(y - 1, ny), yp = wrapf(y + 1, ny);
  const int zm = wrapf(z - 1, nz), zpl = wrapf(z + 1, nz);
  auto pull_own = [&](int x, uint32_t sw, double *n) {
    const int xm = wrapf(x - 1, nx), x0 = wrapf(x, nx), xp = wrapf(x + 1, nx);
    const int self = gsite(x0, y, z);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::c[i][0], cy = L::c[i][1], cz = L::c[i][2];
      const int sx = cx == 1 ? xm : (cx == -1 ? xp : x0);
      const int sy = cy == 1 ? ym : (cy == -1 ? yp : y);
      const int sz = cz == 1 ? zm : (cz == -1 ? zpl : z);
      const bool bb = SOLID && ((sw >> L::bounce[i]) & 1u);
      const int comp = bb ? L::bounce[i] : i;
      const int src = bb ? self : gsite(sx, sy, sz);
      n[i] = Af[(int64_t)comp * sp + src];
    }
  };
  // Stage the streamed populations of the ring sites of plane x.
  auto stage_ring = [&](int x, double *stg) {
    for (int w = t; w < Sh::ring_items; w += Sh::threads) {
      const int fq = w / Sh::ring, r = w - fq * Sh::ring;
      const int fr = fq >= Q ? 1 : 0, i = fq - fr * Q;
      int ey, ez;
      ring_coord<Sh>(r, ey, ez);
      const int yy = wrapf(y0 - 1 + ey, ny), zz = wrapf(z0 - 1 + ez, nz);
      const int cx = d3q19_c(i, 0), cy = d3q19_c(i, 1), cz = d3q19_c(i, 2);
      const int self = gsite(wrapf(x, nx), yy, zz);
      const int bi = d3q19_bounce(i);
      const uint32_t rw = SOLID ? nbw[self] : 0u;
      double v = 0.0;
      if (!(rw >> 31)) {
        const bool bb = (rw >> bi) & 1u;
        const int comp = bb ? bi : i;
        const int src = bb ? self : gsite(wrapf(x - cx, nx), wrapf(yy - cy, ny), wrapf(zz - cz, nz));
        v = A[(int64_t)(fr * Q + comp) * sp + src];
      }
      stg[w] = v;
    }
  };
  // psi of the ring sites of plane x from the stage (pairwise order).
  auto reduce_ring = [&](int x, const double *stg) {
    if (t < 2 * Sh::ring) {
      const int fr = t >= Sh::ring ? 1 : 0, r = t - fr * Sh::ring;
      int ey, ez;
      ring_coord<Sh>(r, ey, ez);
      double n[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) n[i] = stg[(fr * Q + i) * Sh::ring + r];
      psi_lds[pslot(x)][fr][ey * Sh::EZ + ez] = psi_of(p.psi_form, density<L>(n));
    }
  };
  const int me_e = (ly + 1) * Sh::EZ + (lz + 1);

  // ---- prologue ---------------------------------------------------------------
  double cur[Q];
  double rho_cur = 0.0;
  uint32_t sw_cur;
  {
    const uint32_t sw = solid_word(xs - 1);
    double n[Q];
    if (!(sw >> 31)) {
      pull_own(xs - 1, sw, n);
      psi_lds[pslot(xs - 1)][f][me_e] = psi_of(p.psi_form, density<L>(n));
    }
    stage_ring(xs - 1, stage[0]);
    __syncthreads();
    reduce_ring(xs - 1, stage[0]);
    sw_cur = solid_word(xs);
    if (!(sw_cur >> 31)) {
      pull_own(xs, sw_cur, cur);
      rho_cur = density<L>(cur);
      psi_lds[pslot(xs)][f][me_e] = psi_of(p.psi_form, rho_cur);
    }
    stage_ring(xs, stage[1]);
    __syncthreads();
    reduce_ring(xs, stage[1]);
    __syncthreads();
    stage_ring(xs + 1, stage[(xs + 1) & 1]);
    __syncthreads();
  }

  uint32_t sw_ahead = solid_word(xs + 1);
  for (int x = xs; x < xe; ++x) {
    // ---- producers of psi(x+1) -------------------------------------------
    double nxt[Q];
    double rho_nxt = 0.0;
    const uint32_t sw_nxt = sw_ahead;
    sw_ahead = solid_word(x + 2);
    if (!(sw_nxt >> 31)) {
      pull_own(x + 1, sw_nxt, nxt);
      rho_nxt = density<L>(nxt);
      psi_lds[pslot(x + 1)][f][me_e] = psi_of(p.psi_form, rho_nxt);
    }
    reduce_ring(x + 1, stage[(x + 1) & 1]);
    stage_ring(x + 2, stage[x & 1]);
    __syncthreads();
    // ---- collision of plane x ----------------------------------------------
    {
      const double rho_partner = shfl_xor32(rho_cur);
      double mom[3], mom_p[3];
      const bool bdist2 = p.eq_form == kBDist2;
      if (bdist2) {
        first_moment<L>(cur, mom);
#pragma unroll
        for (int d = 0; d < 3; ++d) mom_p[d] = shfl_xor32(mom[d]);
      }
      if (!(sw_cur >> 31)) {
        const int xw = x;
        double F[3];
        sc_force<L, C>(
            f, F, p,
            [&](int ff, int i) {
              if (i < 0) return psi_lds[pslot(xw)][ff][me_e];
              return psi_lds[pslot(xw + L::c[i][0])][ff][me_e + L::c[i][1] * Sh::EZ + L::c[i][2]];
            },
            [&](int i) { return SOLID && ((sw_cur >> i) & 1u); },
            [&](int ff, int i) {
              const int nb = gsite(wrapf(xw + L::c[i][0], nx), wrapf(y + L::c[i][1], ny),
                                   wrapf(z + L::c[i][2], nz));
              return S[ff * sp + nb];
            });
        const double total =
            total_density<C>([&](int ff) { return ff == f ? rho_cur : rho_partner; });
        double wv[3];
        if (bdist2)
          weighted_velocity<L, C>(
              [&](int ff, int d) { return ff == f ? mom[d] : mom_p[d]; }, total, p, wv);
        const int s = xw * plane + y * nz + z;
        collide_component<L>(f, cur, rho_cur, total, F, wv, p,
                             [&](int i, double v) {
                               // streaming store: the new populations are not read again
                               // in this step, so they bypass the caches
                               __builtin_nontemporal_store(v, &Bf[(int64_t)i * sp + s]);
                             });
      }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) cur[i] = nxt[i];
    rho_cur = rho_nxt;
    sw_cur = sw_nxt;
  }
}

constexpr int kFusedTY = 8, kFusedTZ = 32;

inline bool fused_d3q19_supported(const Geo &g) {
  return g.nz % kFusedTZ == 0 && g.ny % kFusedTY == 0 && g.nx >= 4 && g.ny >= 4 && g.nz >= 4;
}

// Chunk length: x is cut into as many chunks as it takes for the grid to hold
// at least `target` workgroups (one per CU by default: 256 CUs, and the FP64
// register footprint allows one 512-thread workgroup per CU), each chunk at
// least four planes long.
inline void fused_d3q19_grid(const Geo &g, int target, int &chunk, int &nchunks, int &ty,
                             int &tz) {
  ty = g.ny / kFusedTY;
  tz = g.nz / kFusedTZ;
  const int tiles = ty * tz;
  if (target <= 0) target = 256;
  int want = (target + tiles - 1) / tiles;
  if (want > g.nx / 4) want = g.nx / 4 > 0 ? g.nx / 4 : 1;
  chunk = (g.nx + want - 1) / want;
  nchunks = (g.nx + chunk - 1) / chunk;
}

inline int launch_fused_d3q19(const double *A, double *B, const uint32_t *nbw, double *S,
                              const Geo &g, const StepParams &p, int target_blocks,
                              hipStream_t stream) {
  if (!fused_d3q19_supported(g)) return -1;
  int chunk, nchunks, ty, tz;
  fused_d3q19_grid(g, target_blocks, chunk, nchunks, ty, tz);
  const int nblocks = ty * tz * nchunks;
  constexpr int threads = 64 * kFusedTY;
  if (p.has_solid)
    hipLaunchKernelGGL((k_fused_d3q19<kFusedTY, true>), dim3(nblocks), dim3(threads), 0, stream,
                       A, B, nbw, S, g, p, chunk, ty, tz, nchunks);
  else
    hipLaunchKernelGGL((k_fused_d3q19<kFusedTY, false>), dim3(nblocks), dim3(threads), 0,
                       stream, A, B, nbw, S, g, p, chunk, ty, tz, nchunks);
  return 0;
}

}  // namespace dlbc
