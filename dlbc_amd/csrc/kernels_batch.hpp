// kernels_batch.hpp -- persistent D3Q19 two-component Shan-Chen step over a
// list of tiles of one or many lattices, on shifted FP32 / FP16 storage.
//
// The per-tile schedule is the one of the fused kernels (kernels_fused.hpp):
// a workgroup owns a TY x 32 tile of the (y, z) plane, marches along x over a
// chunk of planes, keeps psi of three planes in an LDS ring and stages the
// one-site ring around the tile two planes ahead.  Two things are added:
//
//   * partial tiles: a tile at the end of y or z covers vy <= TY rows and
//     vz <= 32 columns; the psi ring is the border of the (vy+2) x (vz+2)
//     box, so lattices of any size n >= 4 are handled;
//   * a work list: the launch is persistent (a few workgroups per CU), each
//     workgroup walks a list of items (lattice, tile, chunk), and every item
//     carries its lattice's geometry, buffers and parameters.  A whole
//     ensemble of independent lattices of uneven size and parameters thus
//     advances in one launch per step.
//
// The arithmetic per site is the one of kernels_shifted.hpp (deviation form,
// FP32), evaluated in the same order, so a lattice gives the same bits
// whether it is stepped alone or inside a batch, and whatever the tile
// height, chunk length or grid size.
#pragma once

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <vector>

#include "kernels_shifted.hpp"
#include "kernels_split.hpp"

namespace dlbc {

// One lattice of a batch.  A[0] is read and A[1] written on even steps of
// the batch, the other way round on odd steps.
struct BatchLattice {
  const void *pop[2];
  const uint32_t *nbw;  // solid-neighbour words (used when p.has_solid)
  const double *S;      // density at solid sites (wall force)
  float *P;             // psi of the split schedule
  Geo g;
  ShiftParams p;
};

// The table as the kernels read it: in the constant address space.  A
// pointer loaded from generic memory is generic to the compiler, and every
// access through it becomes a flat instruction, which also waits on the LDS
// counter; a pointer loaded from constant memory is taken for global memory
// (global loads and stores with a scalar base).  The host writes the table
// before the launches that read it, never during one.  (The host pass only
// parses the kernels; there the qualifier is left out.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const BatchLattice ConstLattice;
#else
typedef const BatchLattice ConstLattice;
#endif

__device__ inline ConstLattice &table_entry(const BatchLattice *lats, int l) {
  return ((ConstLattice *)lats)[l];
}

// A work item: lattice index, tile origin (y0, z0), planes [xs, xe).
struct BatchItem {
  int lat, y0, z0, xs, xe, pad0, pad1, pad2;
};

// Ring index r -> (ey, ez) on the border of the (vy+2) x (vz+2) box: the two
// rows first, then the two columns.
__device__ inline void ring_coord_v(int r, int vy, int vz, int &ey, int &ez) {
  const int ez2 = vz + 2;
  if (r < 2 * ez2) {
    ey = r < ez2 ? 0 : vy + 1;
    ez = r < ez2 ? r : r - ez2;
  } else {
    const int k = r - 2 * ez2;
    ey = 1 + (k >> 1);
    ez = (k & 1) ? vz + 1 : 0;
  }
}

// One work item (a tile and a chunk of planes).  FULL: the tile is TY x 32
// (ring size and coordinates known at compile time); SOLID: the lattice has
// solid sites.
template <class Store, int TY, bool FULL, bool SOLID>
__device__ __forceinline__ void batch_tile(const BatchItem &item, ConstLattice &B,
                                           int parity, float (*psi_lds)[2][FusedShape<TY>::EY *
                                                                           FusedShape<TY>::EZ],
                                           float (*stage)[FusedShape<TY>::ring_items]) {
  using L = D3Q19;
  using Sh = FusedShape<TY>;
  constexpr int Q = 19, TZ = Sh::TZ, EZ = Sh::EZ;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int ly = t >> 6;
  const int lz = lane & 31;
  const int f = lane >> 5;
  const int me_e = (ly + 1) * EZ + (lz + 1);
  auto pslot = [](int x) { return x & 3; };
  {
    const Geo g = B.g;
    // Parameters: uniform ones read from the table, per-component ones selected
    // into registers once per item (a lane holds component f).
    const ShiftParams p = B.p;
    const ShiftLane pl = shift_lane(p, f);
    const float rho0_0 = p.rho0[0], rho0_1 = p.rho0[1];
    const int psi_form = p.psi_form;
    const Store *__restrict__ A = static_cast<const Store *>(B.pop[parity]);
    Store *__restrict__ W = static_cast<Store *>(const_cast<void *>(B.pop[parity ^ 1]));
    const uint32_t *__restrict__ nbw = B.nbw;
    constexpr bool solid = SOLID;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t sp = g.sp;
    const int plane = ny * nz;
    const int y0 = item.y0, z0 = item.z0, xs = item.xs, xe = item.xe;
    const int vy = FULL ? TY : min(TY, ny - y0), vz = FULL ? TZ : min(TZ, nz - z0);
    const bool active = FULL || (ly < vy && lz < vz);
    const int ring = FULL ? Sh::ring : 2 * (vz + 2) + 2 * vy;
    const int ring_items = ring * 2 * Q;
    const int y = y0 + ly, z = z0 + lz;
    // component f's populations: elements shift_elem(f, i, s) = 2 (i sp + s) + f
    const Store *__restrict__ Af = A + f;
    Store *__restrict__ Wf = W + f;
    const float rho0 = pl.rho0;

    auto gsite = [&](int x, int yy, int zz) { return (x * ny + yy) * nz + zz; };
    // bit 31 set: nothing to do at this lane and plane (solid site or no site)
    auto solid_word = [&](int x) -> uint32_t {
      if (!active) return 1u << 31;
      return solid ? nbw[gsite(wrapf(x, nx), y, z)] : 0u;
    };
    const int ym = active ? wrapf(y - 1, ny) : 0, yp = active ? wrapf(y + 1, ny) : 0;
    const int zm = active ? wrapf(z - 1, nz) : 0, zpl = active ? wrapf(z + 1, nz) : 0;
    auto pull_own = [&](int x, uint32_t sw, float *n) {
      const int xm = wrapf(x - 1, nx), x0 = wrapf(x, nx), xp = wrapf(x + 1, nx);
      const int self = gsite(x0, y, z);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = L::c[i][0], cy = L::c[i][1], cz = L::c[i][2];
        const int sx = cx == 1 ? xm : (cx == -1 ? xp : x0);
        const int sy = cy == 1 ? ym : (cy == -1 ? yp : y);
        const int sz = cz == 1 ? zm : (cz == -1 ? zpl : z);
        const bool bb = (sw >> L::bounce[i]) & 1u;
        const int comp = bb ? L::bounce[i] : i;
        const int src = bb ? self : gsite(sx, sy, sz);
        n[i] = to_f(Af[2 * ((int64_t)comp * sp + src)]);
      }
    };
    auto drho = [](const float *n) { return shifted_drho(n); };
    auto stage_ring = [&](int x, float *stg) {
      for (int w = t; w < ring_items; w += Sh::threads) {
        const int fq = w / ring, r = w - fq * ring;
        const int fr = fq >= Q ? 1 : 0, i = fq - fr * Q;
        int ey, ez;
        if constexpr (FULL)
          ring_coord<Sh>(r, ey, ez);
        else
          ring_coord_v(r, vy, vz, ey, ez);
        const int yy = wrapf(y0 - 1 + ey, ny), zz = wrapf(z0 - 1 + ez, nz);
        const int cx = d3q19_c(i, 0), cy = d3q19_c(i, 1), cz = d3q19_c(i, 2);
        const int self = gsite(wrapf(x, nx), yy, zz);
        const int bi = d3q19_bounce(i);
        const uint32_t rw = solid ? nbw[self] : 0u;
        float v = 0.0f;
        if (!(rw >> 31)) {
          const bool bb = (rw >> bi) & 1u;
          const int comp = bb ? bi : i;
          const int src =
              bb ? self : gsite(wrapf(x - cx, nx), wrapf(yy - cy, ny), wrapf(zz - cz, nz));
          v = to_f(A[shift_elem(fr, comp, src, sp)]);
        }
        stg[w] = v;
      }
    };
    auto reduce_ring = [&](int x, const float *stg) {
      if (t < 2 * ring) {
        const int fr = t >= ring ? 1 : 0, r = t - fr * ring;
        int ey, ez;
        if constexpr (FULL)
          ring_coord<Sh>(r, ey, ez);
        else
          ring_coord_v(r, vy, vz, ey, ez);
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < Q; ++i) s += stg[(fr * Q + i) * ring + r];
        psi_lds[pslot(x)][fr][ey * EZ + ez] = psi_f(psi_form, (fr ? rho0_1 : rho0_0) + s);
      }
    };

    float cur[Q];
    float rho_cur = 0.0f;
    uint32_t sw_cur;
    {
      const uint32_t sw = solid_word(xs - 1);
      float n[Q];
      if (!(sw >> 31)) {
        pull_own(xs - 1, sw, n);
        psi_lds[pslot(xs - 1)][f][me_e] = psi_f(psi_form, rho0 + drho(n));
      }
      stage_ring(xs - 1, stage[0]);
      __syncthreads();
      reduce_ring(xs - 1, stage[0]);
      sw_cur = solid_word(xs);
      if (!(sw_cur >> 31)) {
        pull_own(xs, sw_cur, cur);
        rho_cur = rho0 + drho(cur);
        psi_lds[pslot(xs)][f][me_e] = psi_f(psi_form, rho_cur);
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
      float nxt[Q];
      float rho_nxt = 0.0f;
      const uint32_t sw_nxt = sw_ahead;
      sw_ahead = solid_word(x + 2);
      if (!(sw_nxt >> 31)) {
        pull_own(x + 1, sw_nxt, nxt);
        rho_nxt = rho0 + drho(nxt);
        psi_lds[pslot(x + 1)][f][me_e] = psi_f(psi_form, rho_nxt);
      }
      reduce_ring(x + 1, stage[(x + 1) & 1]);
      stage_ring(x + 2, stage[x & 1]);
      __syncthreads();
      {
        float j[3];
        shifted_moment(cur, j);
        const float rho_p = shfl_xor32f(rho_cur);
        float j_p[3] = {0.0f, 0.0f, 0.0f};
        if (p.eq_form == kBDist2) {
#pragma unroll
          for (int d = 0; d < 3; ++d) j_p[d] = shfl_xor32f(j[d]);
        }
        if (!(sw_cur >> 31)) {
          const int s = x * plane + y * nz + z;
          shifted_collide_site<SOLID>(
              f, cur, rho_cur, rho_p, j, j_p, sw_cur, p, pl,
              [&](int ff, int i) {
                if (i < 0) return psi_lds[pslot(x)][ff][me_e];
                return psi_lds[pslot(x + L::c[i][0])][ff][me_e + L::c[i][1] * EZ + L::c[i][2]];
              },
              [&](int i, bool b) {
                const int nb = gsite(wrapf(x + L::c[i][0], nx), wrapf(y + L::c[i][1], ny),
                                     wrapf(z + L::c[i][2], nz));
                return B.P[f * sp + (b ? nb : s)];  // the density at solid sites
              },
              [&](int i, float v) { store_nt<Store>(&Wf[2 * ((int64_t)i * sp + s)], v); });
        }
      }
#pragma unroll
      for (int i = 0; i < Q; ++i) cur[i] = nxt[i];
      rho_cur = rho_nxt;
      sw_cur = sw_nxt;
    }
    __syncthreads();  // the next item reuses psi_lds and stage
  }
}

// SOLID: the solid-capable kernel (any lattice: one without solid sites has
// all-zero solid words and takes the same arithmetic path).  It fits in 128
// VGPRs, i.e. two 512-thread workgroups per CU.  !SOLID: lattices without
// solid sites only; the compiler hoists more loads and needs ~166 VGPRs (one
// 512-thread workgroup per CU).  The autotuner picks between them.
template <class Store, int TY, bool SOLID>
__global__ void __launch_bounds__(64 * TY, SOLID ? 4 : 1)
    k_batch_d3q19(const BatchLattice *__restrict__ lats, const BatchItem *__restrict__ items,
                  int nitems, int parity, int per_xcd) {
  using Sh = FusedShape<TY>;
  __shared__ float psi_lds[4][2][Sh::EY * Sh::EZ];
  __shared__ float stage[2][Sh::ring_items];
  // XCD-aware walk: workgroup b runs on XCD b % 8; the item list is cut into
  // eight contiguous segments (neighbouring tiles), one per XCD, and the
  // per_xcd workgroups of an XCD stride through its segment.
  const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int seg = (nitems + 7) >> 3;
  const int it0 = xcd * seg + j, it1 = min(nitems, (xcd + 1) * seg);
  for (int it = it0; it < it1; it += per_xcd) {
    const BatchItem item = items[it];
    ConstLattice &B = table_entry(lats, item.lat);
    const bool full = item.y0 + TY <= B.g.ny && item.z0 + Sh::TZ <= B.g.nz;
    if (full)
      batch_tile<Store, TY, true, SOLID>(item, B, parity, psi_lds, stage);
    else
      batch_tile<Store, TY, false, SOLID>(item, B, parity, psi_lds, stage);
  }
}

// Work list of a set of lattices: tiles of TY x 32 (partial at the ends of y
// and z) and x chunks of at most `chunk` planes (at least 4), lattice-major,
// then chunk, then tile row, then tile column.
inline void batch_items(const Geo *g, int nlat, int ty, int chunk, std::vector<BatchItem> &out) {
  out.clear();
  for (int l = 0; l < nlat; ++l) {
    const int nx = g[l].nx, ny = g[l].ny, nz = g[l].nz;
    const int c = std::max(4, std::min(chunk, nx));
    const int nch = (nx + c - 1) / c;
    const int len = (nx + nch - 1) / nch;  // balanced chunks
    for (int k = 0; k < nch; ++k) {
      const int xs = k * len, xe = std::min(nx, xs + len);
      if (xs >= xe) continue;
      for (int y0 = 0; y0 < ny; y0 += ty)
        for (int z0 = 0; z0 < nz; z0 += 32) out.push_back({l, y0, z0, xs, xe, 0, 0, 0});
    }
  }
}

template <class Store>
inline int launch_batch_d3q19(const BatchLattice *lats, const BatchItem *items, int nitems,
                              int parity, int ty, int grid, bool solid_kernel,
                              hipStream_t stream) {
  if (nitems <= 0) return 0;
  grid = std::max(8, grid / 8 * 8);
  int per_xcd = grid / 8;
  const int seg = (nitems + 7) / 8;
  if (per_xcd > seg) per_xcd = seg;  // no idle workgroups
  grid = per_xcd * 8;
  if (ty == 8 && solid_kernel)
    hipLaunchKernelGGL((k_batch_d3q19<Store, 8, true>), dim3(grid), dim3(512), 0, stream, lats,
                       items, nitems, parity, per_xcd);
  else if (ty == 8)
    hipLaunchKernelGGL((k_batch_d3q19<Store, 8, false>), dim3(grid), dim3(512), 0, stream, lats,
                       items, nitems, parity, per_xcd);
  else if (ty == 4 && solid_kernel)
    hipLaunchKernelGGL((k_batch_d3q19<Store, 4, true>), dim3(grid), dim3(256), 0, stream, lats,
                       items, nitems, parity, per_xcd);
  else if (ty == 4)
    hipLaunchKernelGGL((k_batch_d3q19<Store, 4, false>), dim3(grid), dim3(256), 0, stream, lats,
                       items, nitems, parity, per_xcd);
  else
    return -1;
  return 0;
}

// ---- batches on the split schedule -------------------------------------------
// Two launches per step over all members: pass 1 (psi) then pass 2 (collide).
// Block b of a launch belongs to the lattice l with first[l] <= blk0 + b <
// first[l + 1] (first: the prefix sums of the members' 128-site block counts).

__device__ inline int batch_block_lattice(const int *__restrict__ first, int nlat, int b) {
  int lo = 0, hi = nlat;  // first[lo] <= b < first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <class Store, bool OFF32>
__global__ void __launch_bounds__(kShiftThreads)
    k_batch_split_psi(const BatchLattice *__restrict__ lats, const int *__restrict__ first,
                      int nlat, int blk0, int blk1, int parity) {
  const int b = blk0 + blockIdx.x;
  if (b >= blk1) return;
  const int l = uniform(batch_block_lattice(first, nlat, b));
  ConstLattice &B = table_entry(lats, l);
  const Store *A = static_cast<const Store *>(B.pop[parity]);
  const Geo g = B.g;
  const ShiftParams p = B.p;
  if (!p.sc) return;  // this member's collide pass reads no psi
  if (p.has_solid)
    ssplit_psi_block<Store, true, OFF32>(A, B.nbw, B.P, g, p, b - first[l]);
  else
    ssplit_psi_block<Store, false, OFF32>(A, B.nbw, B.P, g, p, b - first[l]);
}

// At most 128 VGPRs: four waves per SIMD (the compiler's choice, 129, gave three).
template <class Store, bool OFF32>
__global__ void __launch_bounds__(kShiftThreads) __attribute__((amdgpu_waves_per_eu(4)))
    k_batch_split_collide(const BatchLattice *__restrict__ lats, const int *__restrict__ first,
                          int nlat, int blk0, int blk1, int parity) {
  const int b = blk0 + blockIdx.x;
  if (b >= blk1) return;  // whole block: no barrier is left waiting
  const int l = uniform(batch_block_lattice(first, nlat, b));
  ConstLattice &B = table_entry(lats, l);
  const Store *A = static_cast<const Store *>(B.pop[parity]);
  Store *W = static_cast<Store *>(const_cast<void *>(B.pop[parity ^ 1]));
  const Geo g = B.g;
  const ShiftParams p = B.p;
  if (p.has_solid)
    ssplit_collide_block<Store, true, OFF32>(A, W, B.nbw, B.P, g, p, b - first[l], nullptr);
  else
    ssplit_collide_block<Store, false, OFF32>(A, W, B.nbw, B.P, g, p, b - first[l],
                                              nullptr);
}

// The split pair as one persistent launch per step: the workgroups (all
// resident: a cooperative launch) take the 128-site blocks of every member in
// turn (block b on workgroup b mod G, so on the XCD the two-launch form runs
// it on) for the psi pass, wait at a grid barrier, then take them again for
// the stream/force/collide pass.  Same per-block code, hence the same bits,
// as the two launches.  psi is uniform, and nothing returns from the kernel
// before the barrier, so every workgroup reaches it.  (Contiguous runs of
// blocks per workgroup instead were slower: 8.35 against 8.07 ms per step on
// the 256-lattice ensemble.)
template <class Store, bool OFF32>
__global__ void __launch_bounds__(kShiftThreads) __attribute__((amdgpu_waves_per_eu(4)))
    k_batch_split_persistent(const BatchLattice *__restrict__ lats, const int *__restrict__ first,
                             int nlat, int nblk, int parity, int psi) {
  if (psi) {
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
      const int l = uniform(batch_block_lattice(first, nlat, b));
      ConstLattice &B = table_entry(lats, l);
      if (!B.p.sc) continue;  // this member's collide pass reads no psi
      const Store *A = static_cast<const Store *>(B.pop[parity]);
      if (B.p.has_solid)
        ssplit_psi_block<Store, true, OFF32>(A, B.nbw, B.P, B.g, B.p, b - first[l]);
      else
        ssplit_psi_block<Store, false, OFF32>(A, B.nbw, B.P, B.g, B.p, b - first[l]);
    }
    cooperative_groups::this_grid().sync();
  }
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int l = uniform(batch_block_lattice(first, nlat, b));
    ConstLattice &B = table_entry(lats, l);
    const Store *A = static_cast<const Store *>(B.pop[parity]);
    Store *W = static_cast<Store *>(const_cast<void *>(B.pop[parity ^ 1]));
    if (B.p.has_solid)
      ssplit_collide_block<Store, true, OFF32>(A, W, B.nbw, B.P, B.g, B.p, b - first[l],
                                               nullptr);
    else
      ssplit_collide_block<Store, false, OFF32>(A, W, B.nbw, B.P, B.g, B.p, b - first[l],
                                                nullptr);
  }
}

// Workgroups of the persistent split launch: as many as can be resident at
// once on the device (a cooperative launch needs every one resident), at
// most the blocks to do; 0 if the query fails.
template <class Store, bool OFF32>
inline int batch_split_persistent_grid(int device, int nblk) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(&k_batch_split_persistent<Store, OFF32>),
          kShiftThreads, 0) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  const int g = per_cu * cus;
  return g < nblk ? g : nblk;
}

template <class Store, bool OFF32>
inline hipError_t launch_batch_split_persistent(const BatchLattice *lats, const int *first,
                                                int nlat, int nblk, int parity, bool psi,
                                                int grid, hipStream_t stream) {
  int psi_i = psi ? 1 : 0;
  void *args[] = {(void *)&lats, (void *)&first, (void *)&nlat, (void *)&nblk, (void *)&parity,
                  (void *)&psi_i};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_batch_split_persistent<Store, OFF32>),
                                    dim3(grid), dim3(kShiftThreads), args, 0, stream);
}

// One step of the blocks [blk0, blk0 + nblk) of a batch on the split schedule;
// psi: whether any of these members has Shan-Chen on (else pass 1 is skipped).
template <class Store, bool OFF32>
inline void launch_batch_split(const BatchLattice *lats, const int *first, int nlat, int blk0,
                               int nblk, int parity, bool psi, hipStream_t stream) {
  if (nblk <= 0) return;
  const int grid = nblk;
  if (psi)
    hipLaunchKernelGGL((k_batch_split_psi<Store, OFF32>), dim3(grid), dim3(kShiftThreads), 0,
                       stream, lats, first, nlat, blk0, blk0 + nblk, parity);
  hipLaunchKernelGGL((k_batch_split_collide<Store, OFF32>), dim3(grid), dim3(kShiftThreads), 0,
                     stream, lats, first, nlat, blk0, blk0 + nblk, parity);
}

}  // namespace dlbc
