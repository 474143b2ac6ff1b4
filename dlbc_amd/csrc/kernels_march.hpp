// kernels_march.hpp -- single-pass D3Q19 two-component step over whole z-rows.
//
// The collision at site x needs psi at its 18 neighbours, and psi at a site
// needs that site's 19 streamed populations.  The split pair pays for this
// dependency with a second launch that reads every population again (304 B
// per fluid site in FP64).  Here one launch per step reads each population
// once from HBM:
//
//   * a workgroup owns TY whole z-rows (y0 .. y0+TY-1, every z) and marches
//     along x over a chunk [x0, x1) of planes; wave w holds row w / NW,
//     sites z = 32 (w % NW) + (l & 31), component l >> 5, so the two
//     components of a site are partner lanes (l, l ^ 32) and exchange their
//     densities and momenta by a cross-lane shuffle;
//   * for the plane ahead (x+1) the waves pull the streamed populations of
//     the own rows and of the two rows beside them (y0-1, y0+TY: halo rows,
//     pulled for psi only) and write psi into an LDS ring of four planes.
//     A whole row needs no z halo: the ring wraps z itself.  The halo rows
//     are the neighbouring workgroups' own rows of the same plane, so their
//     second read is served by the caches, not HBM (neighbouring rows go to
//     one XCD, and the workgroups march in step);
//   * the own rows' pulled populations stay in registers for one iteration:
//     the collision of plane x uses the populations pulled for psi(x) one
//     iteration earlier, with psi of planes x-1 .. x+1 from the ring.  One
//     barrier per plane separates the producers of psi(x+1) from the
//     collision of plane x; four ring slots make that barrier sufficient.
//
// Solid sites: the pull takes the bounce-back population where bit bounce(i)
// of the site's solid word is set; the ring holds the density at solid sites
// (from S, as the split pair's psi array does), which the Shan-Chen stencil
// reads for the wall force; solid sites are not written.
//
// Per-site arithmetic is the shared code of the other schedules (density,
// psi_of, sc_force / sc_force_offdiag, total_density, weighted_velocity,
// collide_d3q19_streamed), so all of them give identical bits.
#pragma once

#include <type_traits>

#include "kernels_split.hpp"

namespace dlbc {

// Workgroups a march launch aims for.
constexpr int kMarchTargetBlocks = 512;  // two resident workgroups per CU

// Chunking of x: enough workgroups for kMarchTargetBlocks, chunks of equal
// length (the last one shorter).
__host__ inline void march_chunks(const Geo &g, int ty, int target, int &lx, int &nchunks) {
  const int ntiles = g.ny / ty;
  int want = (target + ntiles - 1) / ntiles;
  want = want < 1 ? 1 : (want > g.nx ? g.nx : want);
  lx = (g.nx + want - 1) / want;
  nchunks = (g.nx + lx - 1) / lx;
}

// Element k of the array at a scalar base: with OFF32 the byte offset is
// formed in 32 bits (k < 2^29), so the access takes the scalar base and a
// 32-bit lane offset even where k differs between lanes.
template <bool OFF32, class T>
__device__ inline T *lane_elem(T *base, int64_t k) {
  if constexpr (OFF32) {
    using Byte = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<Byte *>(base) +
                                 (uint32_t)((uint32_t)k * (uint32_t)sizeof(T)));
  } else {
    return base + k;
  }
}

// The 19 streamed populations of component f at site (x, y, z) -- row bases
// rb[dx+1][dy+1] = site index of (x+dx, y+dy, 0), wrapped; z offsets zm/zp of
// this lane -- with bounce-back by the solid word sw (advection.d:61-89).
// Every address is a wave-uniform base (population and row) plus a 32-bit
// lane offset (component f's arrays start at f 19 sp, then the z position);
// without solids the lane offsets are the same three values for every
// population, so the compiler keeps three registers, not 19 addresses.
template <bool SOLID, bool OFF32>
__device__ inline void march_pull(const double *__restrict__ A, const Geo &g, int f,
                                  const int (&rb)[3][3], int z, int zm, int zp, uint32_t sw,
                                  double *n) {
  using L = D3Q19;
  const uint32_t fo = (uint32_t)f * 19u * (uint32_t)g.sp;
  const uint32_t lz[3] = {fo + (uint32_t)zp, fo + (uint32_t)z, fo + (uint32_t)zm};
#pragma unroll
  for (int i = 0; i < L::q; ++i) {
    const int cz = L::c[i][2];
    const int rs = rb[1 - L::c[i][0]][1 - L::c[i][1]];
    if (SOLID) {
      // a base below both candidate elements: population min(i, bounce(i))
      // of row min(rs, own row); the lane picks its offset from it
      const int b = L::bounce[i], mc = i < b ? i : b;
      const int r0 = rb[1][1];
      const int mr = rs < r0 ? rs : r0;
      const double *base = A + (int64_t)mc * g.sp + mr;
      const uint32_t on = (uint32_t)((i - mc) * g.sp + (rs - mr)) + lz[cz + 1];
      const uint32_t ob = (uint32_t)((b - mc) * g.sp + (r0 - mr)) + lz[1];
      n[i] = *lane_elem<OFF32>(base, ((sw >> b) & 1u) ? ob : on);
    } else {
      n[i] = *lane_elem<OFF32>(A + (int64_t)i * g.sp + rs, lz[cz + 1]);
    }
  }
}

// Row bases of the 3 x 3 rows around (x, y), periodic.
__device__ inline void march_rows(const Geo &g, int x, int y, int (&rb)[3][3]) {
  const int xs[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
  const int ys[3] = {y == 0 ? g.ny - 1 : y - 1, y, y == g.ny - 1 ? 0 : y + 1};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) rb[a][b] = (xs[a] * g.ny + ys[b]) * g.nz;
}

// Workgroup -> (row tile, x chunk), XCD-aware: hardware runs workgroup b on
// XCD b % 8; each XCD takes a contiguous run of (chunk, row tile) pairs, so
// neighbouring row tiles of a chunk share an L2 and run together.
__device__ inline void march_tile(const Geo &g, int lx, int ntiles, int &tile, int &x0, int &x1) {
  const int G = gridDim.x, b = blockIdx.x;
  const int lg = (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
  tile = lg % ntiles;
  x0 = (lg / ntiles) * lx;
  x1 = x0 + lx < g.nx ? x0 + lx : g.nx;
}

// psi (or, at a solid site, the density held in S) written into a ring row,
// with the periodic images of z = n_z - 1 and z = 0 at the row's two ends.
template <int NZ>
__device__ inline void ring_put(double *row, int z, double v) {
  row[z + 1] = v;
  if (z == 0) row[NZ + 1] = v;
  if (z == NZ - 1) row[0] = v;
}

// Collision of component f at site s = (x, y, z) of the own row: nc holds its
// streamed populations, sw its solid word, rb the row bases around (x, y);
// psi of planes x-1 .. x+1 comes from the ring (slot sl is plane x, ring row
// rr the own row).  The partner lane (l ^ 32) holds the other component.
template <int NZ, int RY, bool SOLID, bool OFFDIAG, bool OFF32>
__device__ inline void march_collide(double (*ring)[RY][2][NZ + 2], int sl, int rr, int f, int z,
                                     const int (&rb)[3][3], uint32_t sw, const double *nc,
                                     double *__restrict__ B, const Geo &g, const StepParams &p) {
  using L = D3Q19;
  const bool fluid = !(SOLID && (sw >> 31));
  double rho = 0.0, mom[3] = {0.0, 0.0, 0.0};
  if (fluid) {
    rho = density<L>(nc);
    if (p.eq_form == kBDist2) first_moment<L>(nc, mom);
  }
  const double rho_p = shfl_xor32(rho);
  double mom_p[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) mom_p[d] = shfl_xor32(mom[d]);
  if (!fluid) return;  // solid populations stay as they are
  auto psi = [&](int ff, int i) {
    if (i < 0) return ring[sl][rr][ff][z + 1];
    return ring[(sl + L::c[i][0]) & 3][rr + L::c[i][1]][ff][z + 1 + L::c[i][2]];
  };
  auto solid = [&](int i) { return SOLID && ((sw >> i) & 1u); };
  double F[3];
  if constexpr (OFFDIAG)
    sc_force_offdiag(f, F, p, psi, solid, psi);
  else
    sc_force<L, 2>(f, F, p, psi, solid, psi);
  const double total = total_density<2>([&](int ff) { return ff == f ? rho : rho_p; });
  double wv[3];
  if (p.eq_form == kBDist2)
    weighted_velocity<L, 2>([&](int ff, int d) { return ff == f ? mom[d] : mom_p[d]; }, total, p,
                            wv);
  // element f 19 sp + i sp + s of B: the scalar base of population i in this
  // row plus the lane offset f 19 sp + z
  const uint32_t lo = (uint32_t)f * 19u * (uint32_t)g.sp + (uint32_t)z;
  double *__restrict__ Brow = B + rb[1][1];
  collide_d3q19_streamed(f, nc, rho, total, F, wv, p, [&](int i, double v) {
    __builtin_nontemporal_store(v, lane_elem<OFF32>(Brow + (int64_t)i * g.sp, lo));
  });
}

// TY rows per workgroup; every wave pulls its own row and one (TY = 2) or
// both (TY = 1) neighbouring rows for psi, one row's loads in flight at a
// time, and pulls its own row again for the collision (the same lines as
// one iteration earlier).  128 VGPRs: four waves per SIMD.
template <int NW, int TY, bool SOLID, bool OFFDIAG, bool OFF32>
__global__ void __launch_bounds__(64 * NW * TY) __attribute__((amdgpu_waves_per_eu(4)))
    k_march_d3q19(const double *__restrict__ A, double *__restrict__ B,
                  const uint32_t *__restrict__ nbw, const double *__restrict__ S, Geo g,
                  StepParams p, int lx, int ntiles) {
  using L = D3Q19;
  constexpr int NZ = 32 * NW, RY = TY + 2;
  __shared__ double ring[4][RY][2][NZ + 2];
  int tile, x0, x1;
  march_tile(g, lx, ntiles, tile, x0, x1);

  const int w = uniform(threadIdx.x >> 6);
  const int r = w / NW;  // own row within the tile
  const int lane = threadIdx.x & 63;
  const int f = lane >> 5;
  const int z = (w % NW) * 32 + (lane & 31);
  const int zm = z == 0 ? NZ - 1 : z - 1, zp = z == NZ - 1 ? 0 : z + 1;
  const int y = uniform(tile * TY + r);
  const int yh0 = uniform(y == 0 ? g.ny - 1 : y - 1);
  const int yh1 = uniform(y == g.ny - 1 ? 0 : y + 1);

  // Iteration k computes psi of plane x0 - 1 + k (ring slot k & 3) and, from
  // k = 2 on, collides plane x = x0 - 2 + k, whose neighbour planes are in
  // slots (k - 2) & 3 .. k & 3.  One loop, so the psi code (with its exp)
  // is emitted once.
  const int nk = x1 - x0 + 2;
  for (int k = 0; k < nk; ++k) {
    int xa = x0 - 1 + k;
    xa = xa < 0 ? xa + g.nx : (xa >= g.nx ? xa - g.nx : xa);
    // the rows this wave pulls at plane xa: TY = 1 both neighbouring rows
    // and its own, TY = 2 the neighbouring row on its side and its own
#pragma unroll 1
    for (int j = TY == 1 ? 0 : 1; j < 3; ++j) {
      const int yy =
          uniform(j == 2 ? y : (TY == 1 ? (j == 0 ? yh0 : yh1) : (r == 0 ? yh0 : yh1)));
      const int rr = j == 2 ? r + 1 : (TY == 1 ? (j == 0 ? 0 : 2) : (r == 0 ? 0 : RY - 1));
      int rb[3][3];
      march_rows(g, xa, yy, rb);
      const int s = rb[1][1] + z;
      const uint32_t sw = SOLID ? nbw[s] : 0u;
      double v;
      if (SOLID && (sw >> 31)) {
        v = *lane_elem<OFF32>(S, (int64_t)f * g.sp + s);
      } else {
        double t[19];
        march_pull<SOLID, OFF32>(A, g, f, rb, z, zm, zp, sw, t);
        v = psi_of(p.psi_form, density<L>(t));
      }
      ring_put<NZ>(ring[k & 3][rr][f], z, v);
      // one row's loads in flight at a time: hoisting the next row's loads
      // above this reduction would need more than 128 VGPRs
      __builtin_amdgcn_sched_barrier(0);
    }
    if (k < 2) continue;
    const int x = xa == 0 ? g.nx - 1 : xa - 1;
    __syncthreads();
    int rb[3][3];
    march_rows(g, x, y, rb);
    const int s = rb[1][1] + z;
    const uint32_t sw = SOLID ? nbw[s] : 0u;
    double nc[19];
    if (!(SOLID && (sw >> 31))) march_pull<SOLID, OFF32>(A, g, f, rb, z, zm, zp, sw, nc);
    march_collide<NZ, RY, SOLID, OFFDIAG, OFF32>(ring, (k - 1) & 3, r + 1, f, z, rb, sw, nc, B, g,
                                                p);
  }
}

// One row per workgroup: the loads of a plane's three rows (the own row and
// both neighbours) go out together, and the own row's populations stay in
// registers from their psi to their collision, so each population is read
// once from HBM and the neighbouring rows' second reads hit L2 (the row
// tiles of an XCD march together).  Up to 256 VGPRs: two waves per SIMD, one
// workgroup per CU.
template <int NW, bool SOLID, bool OFFDIAG, bool OFF32>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2)))
    k_march_pipe_d3q19(const double *__restrict__ A, double *__restrict__ B,
                       const uint32_t *__restrict__ nbw, const double *__restrict__ S, Geo g,
                       StepParams p, int lx, int ntiles) {
  using L = D3Q19;
  constexpr int NZ = 32 * NW, RY = 3;
  __shared__ double ring[4][RY][2][NZ + 2];
  int tile, x0, x1;
  march_tile(g, lx, ntiles, tile, x0, x1);

  const int lane = threadIdx.x & 63;
  const int f = lane >> 5;
  const int z = (uniform(threadIdx.x >> 6) % NW) * 32 + (lane & 31);
  const int zm = z == 0 ? NZ - 1 : z - 1, zp = z == NZ - 1 ? 0 : z + 1;
  const int y = uniform(tile);
  // rows pulled at every plane, in ring-row order: y - 1, y, y + 1
  const int yr[3] = {uniform(y == 0 ? g.ny - 1 : y - 1), y, uniform(y == g.ny - 1 ? 0 : y + 1)};

  // Per plane xa the three rows' loads go out together (one round trip per
  // plane); the own row's populations (t[1]) are kept for the collision one
  // iteration later.
  double t[3][19];
  uint32_t swt[3];
  double st[3];
  double nc[19];      // own row, plane being collided
  uint32_t swc = 0u;  // its solid word
  const int nk = x1 - x0 + 2;
  for (int k = 0; k < nk; ++k) {
    int xa = x0 - 1 + k;
    xa = xa < 0 ? xa + g.nx : (xa >= g.nx ? xa - g.nx : xa);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      int rb[3][3];
      march_rows(g, xa, yr[j], rb);
      const int s = rb[1][1] + z;
      swt[j] = SOLID ? nbw[s] : 0u;
      if (SOLID && (swt[j] >> 31))
        st[j] = *lane_elem<OFF32>(S, (int64_t)f * g.sp + s);
      else
        march_pull<SOLID, OFF32>(A, g, f, rb, z, zm, zp, swt[j], t[j]);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double v =
          (SOLID && (swt[j] >> 31)) ? st[j] : psi_of(p.psi_form, density<L>(t[j]));
      ring_put<NZ>(ring[k & 3][j][f], z, v);
    }
    if (k >= 2) {
      __syncthreads();
      const int x = xa == 0 ? g.nx - 1 : xa - 1;
      int rb[3][3];
      march_rows(g, x, y, rb);
      march_collide<NZ, RY, SOLID, OFFDIAG, OFF32>(ring, (k - 1) & 3, 1, f, z, rb, swc, nc, B,
                                                   g, p);
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) nc[i] = t[1][i];
    swc = swt[1];
  }
}

// One march launch with NW waves per row: TY rows per workgroup, or (HOLD)
// the pipelined one-row kernel.
template <int NW, int TY, bool HOLD, bool OFF32>
inline void launch_march_nw(const double *A, double *B, const uint32_t *nbw, const double *S,
                            const Geo &g, const StepParams &p, hipStream_t stream, int target,
                            bool solid, bool offdiag) {
  int lx, nchunks;
  march_chunks(g, TY, target, lx, nchunks);
  const int ntiles = g.ny / TY;
  const dim3 grid(ntiles * nchunks), block(64 * NW * TY);
#define DLBC_MARCH_LAUNCH(SOL, OFFD)                                                          \
  do {                                                                                        \
    if constexpr (HOLD)                                                                       \
      hipLaunchKernelGGL((k_march_pipe_d3q19<NW, SOL, OFFD, OFF32>), grid, block, 0, stream,  \
                         A, B, nbw, S, g, p, lx, ntiles);                                     \
    else                                                                                      \
      hipLaunchKernelGGL((k_march_d3q19<NW, TY, SOL, OFFD, OFF32>), grid, block, 0, stream,   \
                         A, B, nbw, S, g, p, lx, ntiles);                                     \
  } while (0)
  if (solid && offdiag)
    DLBC_MARCH_LAUNCH(true, true);
  else if (solid)
    DLBC_MARCH_LAUNCH(true, false);
  else if (offdiag)
    DLBC_MARCH_LAUNCH(false, true);
  else
    DLBC_MARCH_LAUNCH(false, false);
#undef DLBC_MARCH_LAUNCH
}

// Lane offsets of the march kernel, in elements: below 27 sp in the pulls
// with bounce-back (up to 7 sp between a population and its bounce partner,
// one row span, the component's 19 sp), below 20 sp elsewhere.  OFF32 when
// 28 sp fits 2^29 (byte offsets below 2^32).
inline bool march_off32(const Geo &g) { return 28 * g.sp < ((int64_t)1 << 29); }

// Whether the march schedule serves a lattice with `ty` z-rows per workgroup
// (hold: the populations held in registers between the psi pull and the
// collision, one row per workgroup only): n_z = 32 NW with NW in {1, 2, 8}
// (and 16 for one row), 32-bit lane offsets (march_off32) for NW < 16, n_y even for
// two rows per workgroup.
inline bool march_supported(const Geo &g, int ty, bool hold) {
  if (g.nz % 32 != 0) return false;
  const int nw = g.nz / 32;
  if (nw != 1 && nw != 2 && nw != 8 && nw != 16) return false;
  if (nw < 16 && !march_off32(g)) return false;
  if (ty == 2) return !hold && nw <= 8 && g.ny % 2 == 0;
  return ty == 1;
}

// The launch shape the march schedule uses for (rows, hold) requests
// (0 / -1: the library's choice); false when the lattice is not served.
inline bool march_shape(const Geo &g, int rows, int hold, int &ty, bool &h) {
  h = hold > 0;
  ty = rows;
  if (rows == 0) ty = march_supported(g, 2, h) ? 2 : 1;
  return march_supported(g, ty, h);
}

// One step of the march schedule.  Returns -1 when the lattice is not served.
inline int launch_march_d3q19(const double *A, double *B, const uint32_t *nbw, const double *S,
                              const Geo &g, const StepParams &p, hipStream_t stream, int rows,
                              int hold, int target) {
  int ty;
  bool h;
  if (!march_shape(g, rows, hold, ty, h)) return -1;
  if (target <= 0) target = kMarchTargetBlocks;
  const bool solid = p.has_solid != 0;
  const bool offdiag = p.gcc[0][0] == 0.0 && p.gcc[1][1] == 0.0;
  const bool o32 = march_off32(g);
#define DLBC_MARCH(NW, TY, HOLD, O32) \
  launch_march_nw<NW, TY, HOLD, O32>(A, B, nbw, S, g, p, stream, target, solid, offdiag)
  switch (g.nz / 32) {
    case 1:
      if (ty == 2) DLBC_MARCH(1, 2, false, true);
      else if (h) DLBC_MARCH(1, 1, true, true);
      else DLBC_MARCH(1, 1, false, true);
      return 0;
    case 2:
      if (ty == 2) DLBC_MARCH(2, 2, false, true);
      else if (h) DLBC_MARCH(2, 1, true, true);
      else DLBC_MARCH(2, 1, false, true);
      return 0;
    case 8:
      if (ty == 2) DLBC_MARCH(8, 2, false, true);
      else if (h) DLBC_MARCH(8, 1, true, true);
      else DLBC_MARCH(8, 1, false, true);
      return 0;
    case 16:
      if (h && o32) DLBC_MARCH(16, 1, true, true);
      else if (h) DLBC_MARCH(16, 1, true, false);
      else if (o32) DLBC_MARCH(16, 1, false, true);
      else DLBC_MARCH(16, 1, false, false);
      return 0;
    default:
      return -1;
  }
#undef DLBC_MARCH
}

}  // namespace dlbc
