// kernels_shifted.hpp -- D3Q19 two-component Shan-Chen step on shifted
// populations: storage of d_i = f_i - w_i rho0 in FP32 or FP16, arithmetic
// in FP32.
//
// The physics is the reference's (pull streaming with bounce-back, Shan-Chen
// force with psi = rho or 1 - exp(-rho), velocity shift, BGK collision with
// the SecondOrder, ThirdOrder or BDist2 equilibrium: src/dlbc/lb/advection.d,
// force.d, collision.d, eqdist.d), rewritten in deviation form so that the
// small quantities never meet the large reference weight in a rounding step:
//   rho  = rho0 + sum_i d_i                (sum_i w_i = 1)
//   j    = sum_i d_i c_i                    (sum_i w_i c_i = 0)
//   d_i^eq = w_i (rho - rho0) + rho (E_i(u) - w_i)
//   d_i'  = d_i + omega (d_i^eq - d_i)
// where E_i(u) is the equilibrium of unit density.  The schedule is the one
// of the FP64 fused kernel (kernels_fused.hpp): x-marching 8 x 32 tiles, two
// lanes per site (one per component), psi in an LDS ring of four planes, the
// one-site psi ring staged two planes ahead, one barrier per plane.
#pragma once

#include <hip/hip_fp16.h>

#include "kernels_fused.hpp"

namespace dlbc {

template <class T>
__device__ inline float to_f(T v);
template <>
__device__ inline float to_f<float>(float v) {
  return v;
}
template <>
__device__ inline float to_f<__half>(__half v) {
  return __half2float(v);
}
template <class T>
__device__ inline T from_f(float v);
template <>
__device__ inline float from_f<float>(float v) {
  return v;
}
template <>
__device__ inline __half from_f<__half>(float v) {
  return __float2half_rn(v);
}

struct ShiftParams {
  float tau[2], omega[2], rho0[2];
  float gcc[2][2], gwc[2];
  float gacc[3], body[2][3];
  int eq_form, psi_form, sc, has_solid, has_body;
};

__device__ inline float psi_f(int form, float rho) { return form == 0 ? rho : 1.0f - __expf(-rho); }

// Deviation of the unit-density equilibrium, E_i(u) - w_i, for D3Q19.
__device__ inline void eq_dev_d3q19(int form, const float (&u)[3], float (&e)[19]) {
  using L = D3Q19;
  if (form != kBDist2) {
    const float uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      const float cu = (float)L::c[i][0] * u[0] + (float)L::c[i][1] * u[1] + (float)L::c[i][2] * u[2];
      float t = 3.0f * cu + 4.5f * cu * cu - 1.5f * uu;
      if (form == kThirdOrder) t += 4.5f * cu * cu * cu - 4.5f * uu * cu;
      e[i] = (float)L::w[i] * t;
    }
    return;
  }
  // BDist2 (eqdist.d:102-184) with the constant part w_i removed
  const float aN1 = 1.0f / 36.0f, haN1 = 0.5f * aN1, aN1_6 = aN1 / 6.0f;
  const float tx = 3.0f * u[0], ty = 3.0f * u[1], tz = 3.0f * u[2];
  const float uke = 0.5f * (u[0] * tx + u[1] * ty + u[2] * tz);
  const float c1 = aN1 * (1.0f - uke);
  auto dev = [&](float cdu, float &ev, float &od) {
    ev = -aN1 * uke + haN1 * cdu * cdu;  // even - aN1
    od = c1 * cdu + aN1_6 * cdu * cdu * cdu;
  };
  float ev, od;
  dev(tx, ev, od);
  e[1] = 2.0f * (ev + od);
  e[2] = 2.0f * (ev - od);
  dev(ty, ev, od);
  e[3] = 2.0f * (ev + od);
  e[4] = 2.0f * (ev - od);
  dev(tz, ev, od);
  e[5] = 2.0f * (ev + od);
  e[6] = 2.0f * (ev - od);
  dev(tx + ty, ev, od);
  e[7] = ev + od;
  e[14] = ev - od;
  dev(tx - ty, ev, od);
  e[13] = ev + od;
  e[8] = ev - od;
  dev(tx + tz, ev, od);
  e[11] = ev + od;
  e[18] = ev - od;
  dev(tx - tz, ev, od);
  e[12] = ev + od;
  e[17] = ev - od;
  dev(ty + tz, ev, od);
  e[9] = ev + od;
  e[16] = ev - od;
  dev(ty - tz, ev, od);
  e[15] = ev + od;
  e[10] = ev - od;
  e[0] = -(1.0f / 3.0f) * uke;
}

__device__ inline float shfl_xor32f(float v) { return __shfl_xor(v, 32); }

template <class Store, int TY, bool SOLID>
__global__ void __launch_bounds__(64 * TY)
    k_shifted_d3q19(const Store *__restrict__ A, Store *__restrict__ B,
                    const uint32_t *__restrict__ nbw, const double *__restrict__ S, Geo g,
                    ShiftParams p, int chunk, int ntiles_y, int ntiles_z, int nchunks) {
  using L = D3Q19;
  using Sh = FusedShape<TY>;
  constexpr int Q = 19, TZ = Sh::TZ;
  __shared__ float psi_lds[4][2][Sh::EY * Sh::EZ];
  __shared__ float stage[2][Sh::ring_items];

  const int nblocks = gridDim.x;
  const int b = blockIdx.x;
  const int logical = (nblocks % 8 == 0) ? (b % 8) * (nblocks / 8) + b / 8 : b;
  const int tz = logical % ntiles_z;
  const int ty = (logical / ntiles_z) % ntiles_y;
  const int chunk_id = logical / (ntiles_z * ntiles_y);
  if (chunk_id >= nchunks) return;
  const int xs = chunk_id * chunk;
  const int xe = min(xs + chunk, g.nx);
  const int y0 = ty * TY, z0 = tz * TZ;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int ly = t >> 6;
  const int lz = lane & 31;
  const int f = lane >> 5;
  const int y = y0 + ly, z = z0 + lz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t sp = g.sp;
  const int plane = ny * nz;
  const Store *__restrict__ Af = A + (int64_t)f * Q * sp;
  Store *__restrict__ Bf = B + (int64_t)f * Q * sp;
  const float rho0 = p.rho0[f];

  auto gsite = [&](int x, int yy, int zz) { return (x * ny + yy) * nz + zz; };
  auto pslot = [](int x) { return x & 3; };
  auto solid_word = [&](int x) -> uint32_t {
    if constexpr (SOLID) {
      return nbw[gsite(wrapf(x, nx), y, z)];
    } else {
      return 0u;
    }
  };
  const int ym = wrapf(y - 1, ny), yp = wrapf(y + 1, ny);
  const int zm = wrapf(z - 1, nz), zpl = wrapf(z + 1, nz);
  auto pull_own = [&](int x, uint32_t sw, float *n) {
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
      n[i] = to_f(Af[(int64_t)comp * sp + src]);
    }
  };
  auto drho = [](const float *n) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < Q; ++i) s += n[i];
    return s;
  };
  auto stage_ring = [&](int x, float *stg) {
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
      float v = 0.0f;
      if (!(rw >> 31)) {
        const bool bb = (rw >> bi) & 1u;
        const int comp = bb ? bi : i;
        const int src = bb ? self : gsite(wrapf(x - cx, nx), wrapf(yy - cy, ny), wrapf(zz - cz, nz));
        v = to_f(A[(int64_t)(fr * Q + comp) * sp + src]);
      }
      stg[w] = v;
    }
  };
  auto reduce_ring = [&](int x, const float *stg) {
    if (t < 2 * Sh::ring) {
      const int fr = t >= Sh::ring ? 1 : 0, r = t - fr * Sh::ring;
      int ey, ez;
      ring_coord<Sh>(r, ey, ez);
      float s = 0.0f;
#pragma unroll
      for (int i = 0; i < Q; ++i) s += stg[(fr * Q + i) * Sh::ring + r];
      psi_lds[pslot(x)][fr][ey * Sh::EZ + ez] = psi_f(p.psi_form, p.rho0[fr] + s);
    }
  };
  const int me_e = (ly + 1) * Sh::EZ + (lz + 1);

  float cur[Q];
  float rho_cur = 0.0f;
  uint32_t sw_cur;
  {
    const uint32_t sw = solid_word(xs - 1);
    float n[Q];
    if (!(sw >> 31)) {
      pull_own(xs - 1, sw, n);
      psi_lds[pslot(xs - 1)][f][me_e] = psi_f(p.psi_form, rho0 + drho(n));
    }
    stage_ring(xs - 1, stage[0]);
    __syncthreads();
    reduce_ring(xs - 1, stage[0]);
    sw_cur = solid_word(xs);
    if (!(sw_cur >> 31)) {
      pull_own(xs, sw_cur, cur);
      rho_cur = rho0 + drho(cur);
      psi_lds[pslot(xs)][f][me_e] = psi_f(p.psi_form, rho_cur);
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
      psi_lds[pslot(x + 1)][f][me_e] = psi_f(p.psi_form, rho_nxt);
    }
    reduce_ring(x + 1, stage[(x + 1) & 1]);
    stage_ring(x + 2, stage[x & 1]);
    __syncthreads();
    {
      float j[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) j[d] = 0.0f;
#pragma unroll
      for (int i = 0; i < Q; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (L::c[i][d] != 0) j[d] += (float)L::c[i][d] * cur[i];
      const float rho_p = shfl_xor32f(rho_cur);
      float j_p[3];
      const bool bdist2 = p.eq_form == kBDist2;
      if (bdist2) {
#pragma unroll
        for (int d = 0; d < 3; ++d) j_p[d] = shfl_xor32f(j[d]);
      }
      if (!(sw_cur >> 31)) {
        // Shan-Chen force on component f (force.d:238-309)
        float F[3] = {0.0f, 0.0f, 0.0f};
        const float psi1 = psi_lds[pslot(x)][f][me_e];
        if (p.sc) {
#pragma unroll
          for (int f2 = 0; f2 < 2; ++f2) {
            const float cc = p.gcc[f][f2];
            if (cc == 0.0f) continue;
            const float psi_self = psi_lds[pslot(x)][f2][me_e];
            float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int i = 1; i < Q; ++i) {
              const bool sn = SOLID && ((sw_cur >> i) & 1u);
              const float psi2 =
                  sn ? psi_self
                     : psi_lds[pslot(x + L::c[i][0])][f2][me_e + L::c[i][1] * Sh::EZ + L::c[i][2]];
              const float wp = (float)L::w[i] * psi2;
#pragma unroll
              for (int d = 0; d < 3; ++d)
                if (L::c[i][d] != 0) acc[d] += (float)L::c[i][d] * wp;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) F[d] -= psi1 * cc * acc[d];
          }
          const float wc = p.gwc[f];
          if (SOLID && wc != 0.0f) {
#pragma unroll
            for (int i = 0; i < Q; ++i) {
              if ((sw_cur >> i) & 1u) {
                const int nb = gsite(wrapf(x + L::c[i][0], nx), wrapf(y + L::c[i][1], ny),
                                     wrapf(z + L::c[i][2], nz));
                const float pf = (float)L::w[i] * psi1 * (float)S[f * sp + nb] * wc;
#pragma unroll
                for (int d = 0; d < 3; ++d) F[d] -= pf * (float)L::c[i][d];
              }
            }
          }
        }
        if (p.has_body) {
#pragma unroll
          for (int d = 0; d < 3; ++d) F[d] += p.body[f][d];
        }
        const float rinv = 1.0f / rho_cur;
        float u[3];
        if (bdist2) {
          const float tot_inv = 1.0f / (rho_cur + rho_p);
          const float tau_p = p.tau[f ^ 1];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const float wv = (j[d] / p.tau[f] + j_p[d] / tau_p) * tot_inv;
            u[d] = p.tau[f] * (p.gacc[d] + F[d] * rinv) + wv;
          }
        } else {
#pragma unroll
          for (int d = 0; d < 3; ++d) u[d] = p.tau[f] * (p.gacc[d] + F[d] * rinv) + j[d] * rinv;
        }
        float e[Q];
        eq_dev_d3q19(p.eq_form, u, e);
        const float drh = rho_cur - rho0;
        const float om = p.omega[f];
        const int s = x * plane + y * nz + z;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const float deq = (float)L::w[i] * drh + rho_cur * e[i];
          Bf[(int64_t)i * sp + s] = from_f<Store>(cur[i] + om * (deq - cur[i]));
        }
      }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) cur[i] = nxt[i];
    rho_cur = rho_nxt;
    sw_cur = sw_nxt;
  }
}

template <class Store>
inline int launch_shifted_d3q19(const void *A, void *B, const uint32_t *nbw, const double *S,
                                const Geo &g, const ShiftParams &p, int target_blocks,
                                hipStream_t stream) {
  if (!fused_d3q19_supported(g)) return -1;
  int chunk, nchunks, ty, tz;
  // FP32 arithmetic needs about half the registers of the FP64 kernel: two
  // workgroups per CU by default
  fused_d3q19_grid(g, target_blocks > 0 ? target_blocks : 512, chunk, nchunks, ty, tz);
  const int nblocks = ty * tz * nchunks;
  constexpr int threads = 64 * kFusedTY;
  const Store *a = static_cast<const Store *>(A);
  Store *b = static_cast<Store *>(B);
  if (p.has_solid)
    hipLaunchKernelGGL((k_shifted_d3q19<Store, kFusedTY, true>), dim3(nblocks), dim3(threads), 0,
                       stream, a, b, nbw, S, g, p, chunk, ty, tz, nchunks);
  else
    hipLaunchKernelGGL((k_shifted_d3q19<Store, kFusedTY, false>), dim3(nblocks), dim3(threads),
                       0, stream, a, b, nbw, S, g, p, chunk, ty, tz, nchunks);
  return 0;
}

}  // namespace dlbc
