// kernels_split.hpp -- two-pass D3Q19 two-component step with one lane per
// (site, component).
//
// The general two-pass schedule (kernels_twopass.hpp) gives a thread a whole
// site: both components' populations (76 doubles) live in its registers and
// every neighbour's solid flag is a separate mask load.  Here the two
// components of a site sit on partner lanes l and l ^ 32 of one wave (as in
// the fused kernel), which halves the registers per thread, and the solid
// neighbourhood is the site's one 32-bit solid word.  The new populations are
// written with streaming stores.
//
//   pass 1: pull the 19 streamed populations, density, psi -> S
//   pass 2: pull again, Shan-Chen force from psi in S, velocity shift,
//           BGK collision
//
// Per-site arithmetic is the shared code of the other schedules (density,
// sc_force, total_density, weighted_velocity, collide_component), so all
// three give identical bits.
#pragma once

#include "kernels_fused.hpp"

namespace dlbc {

constexpr int kSplitThreads = 256;  // 4 waves = 128 sites x 2 components

// Site and component of this lane; false past the end of the lattice.
__device__ inline bool split_lane(const Geo &g, int &s, int &f) {
  const int lane = threadIdx.x & 63;
  f = lane >> 5;
  s = blockIdx.x * (kSplitThreads / 2) + (threadIdx.x >> 6) * 32 + (lane & 31);
  return s < g.sites;
}

// The 19 streamed populations of component f at site s = (x, y, z) whose
// solid word is sw (pull with bounce-back, src/dlbc/lb/advection.d:61-89).
template <bool SOLID>
__device__ inline void split_pull(const double *__restrict__ Af, const Geo &g, int s, int x,
                                  int y, int z, uint32_t sw, double *n) {
  using L = D3Q19;
#pragma unroll
  for (int i = 0; i < L::q; ++i) {
    const bool bb = SOLID && ((sw >> L::bounce[i]) & 1u);
    const int src = bb ? s
                       : site_of(wrap(x - L::c[i][0], g.nx), wrap(y - L::c[i][1], g.ny),
                                 wrap(z - L::c[i][2], g.nz), g);
    const int comp = bb ? L::bounce[i] : i;
    n[i] = Af[(int64_t)comp * g.sp + src];
  }
}

template <bool SOLID>
__global__ void __launch_bounds__(kSplitThreads)
    k_split_psi(const double *__restrict__ A, const uint32_t *__restrict__ nbw,
                double *__restrict__ S, Geo g, int psi_form) {
  int s, f;
  if (!split_lane(g, s, f)) return;
  const uint32_t sw = SOLID ? nbw[s] : 0u;
  if (sw >> 31) return;  // solid: S keeps the site's density
  int x, y, z;
  coords(s, g, x, y, z);
  double n[19];
  split_pull<SOLID>(A + (int64_t)f * 19 * g.sp, g, s, x, y, z, sw, n);
  S[f * g.sp + s] = psi_of(psi_form, density<D3Q19>(n));
}

template <bool SOLID>
__global__ void __launch_bounds__(kSplitThreads)
    k_split_collide(const double *__restrict__ A, double *__restrict__ B,
                    const uint32_t *__restrict__ nbw, const double *__restrict__ S, Geo g,
                    StepParams p) {
  using L = D3Q19;
  constexpr int C = 2;
  int s, f;
  if (!split_lane(g, s, f)) return;
  const uint32_t sw = SOLID ? nbw[s] : 0u;
  if (sw >> 31) return;  // solid populations stay as they are
  int x, y, z;
  coords(s, g, x, y, z);
  double n[19];
  split_pull<SOLID>(A + (int64_t)f * 19 * g.sp, g, s, x, y, z, sw, n);
  const double rho = density<L>(n);
  const double rho_partner = shfl_xor32(rho);
  double mom[3], mom_p[3];
  const bool bdist2 = p.eq_form == kBDist2;
  if (bdist2) {
    first_moment<L>(n, mom);
#pragma unroll
    for (int d = 0; d < 3; ++d) mom_p[d] = shfl_xor32(mom[d]);
  }
  auto nbr = [&](int i) {
    return site_of(wrap(x + L::c[i][0], g.nx), wrap(y + L::c[i][1], g.ny),
                   wrap(z + L::c[i][2], g.nz), g);
  };
  double F[3];
  sc_force<L, C>(
      f, F, p, [&](int ff, int i) { return S[ff * g.sp + (i < 0 ? s : nbr(i))]; },
      [&](int i) { return SOLID && ((sw >> i) & 1u); },
      [&](int ff, int i) { return S[ff * g.sp + nbr(i)]; });
  const double total = total_density<C>([&](int ff) { return ff == f ? rho : rho_partner; });
  double wv[3];
  if (bdist2)
    weighted_velocity<L, C>([&](int ff, int d) { return ff == f ? mom[d] : mom_p[d]; }, total,
                            p, wv);
  double *__restrict__ Bf = B + (int64_t)f * 19 * g.sp;
  collide_component<L>(f, n, rho, total, F, wv, p, [&](int i, double v) {
    __builtin_nontemporal_store(v, &Bf[(int64_t)i * g.sp + s]);
  });
}

inline void launch_split_d3q19(const double *A, double *B, const uint32_t *nbw, double *S,
                               const Geo &g, const StepParams &p, hipStream_t stream,
                               int pass) {
  const int nb = (g.sites + kSplitThreads / 2 - 1) / (kSplitThreads / 2);
  if (pass == 0) {
    if (p.has_solid)
      hipLaunchKernelGGL(k_split_psi<true>, dim3(nb), dim3(kSplitThreads), 0, stream, A, nbw, S,
                         g, p.psi_form);
    else
      hipLaunchKernelGGL(k_split_psi<false>, dim3(nb), dim3(kSplitThreads), 0, stream, A, nbw,
                         S, g, p.psi_form);
  } else {
    if (p.has_solid)
      hipLaunchKernelGGL(k_split_collide<true>, dim3(nb), dim3(kSplitThreads), 0, stream, A, B,
                         nbw, S, g, p);
    else
      hipLaunchKernelGGL(k_split_collide<false>, dim3(nb), dim3(kSplitThreads), 0, stream, A, B,
                         nbw, S, g, p);
  }
}

}  // namespace dlbc
