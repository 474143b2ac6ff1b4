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
// where E_i(u) is the equilibrium of unit density.
//
// Constants that multiply populations (the weights w_i, BDist2's 1/36 and
// 1/216) are applied as float-float pairs (hi + lo).  A weight rounded to
// FP32 once is off by the same relative amount at every site and step: in
// deviation form that error lands on the conserved moments (sum_i w_i != 1
// creates mass, 36 * fl(1/36) != 1 scales the equilibrium momentum), and it
// builds up with the step count instead of averaging out.  On the 256^3
// channel after 2000 steps this took the velocity error against FP64 from
// 3.0e-5 to 9.0e-6 (relative 2-norm) at the same speed.
//
// The arithmetic uses explicit fused multiply-adds (fmaf) wherever a product
// meets a sum: one rounding instead of two, and one instruction.
//
// This file holds the storage conversions, parameters and per-site pieces;
// the kernels that step shifted lattices are in kernels_split.hpp and
// kernels_batch.hpp.
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
// The FP32 value is materialised before its conversion: otherwise the
// compiler may fuse a preceding fmaf and the conversion into one
// mixed-precision FMA (v_fma_mixlo_f16, rounded once to FP16) in some
// kernels and not in others, and schedules that share the per-site code
// would no longer give the same bits.
__device__ inline float materialise(float v) {
  asm("" : "+v"(v));
  return v;
}
template <>
__device__ inline __half from_f<__half>(float v) {
  return __float2half_rn(materialise(v));
}

// Streaming (nontemporal) store of one population in the storage type.
template <class T>
__device__ inline void store_nt(T *p, float v);
template <>
__device__ inline void store_nt<float>(float *p, float v) {
  __builtin_nontemporal_store(v, p);
}
template <>
__device__ inline void store_nt<__half>(__half *p, float v) {
  __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(materialise(v))),
                              reinterpret_cast<unsigned short *>(p));
}

// Layout of a shifted-storage buffer: the two components of a population sit
// side by side, element (f, i, s) at 2 (i sp + s) + f, so one 8-byte (FP32)
// or 4-byte (FP16) access moves both components of a site.
__host__ __device__ inline int64_t shift_elem(int f, int i, int64_t s, int64_t sp) {
  return 2 * ((int64_t)i * sp + s) + f;
}

// Both components at p (an even element) as FP32.
template <class T>
__device__ inline void pair_load(const T *p, float &a, float &b);
template <>
__device__ inline void pair_load<float>(const float *p, float &a, float &b) {
  const float2 v = *reinterpret_cast<const float2 *>(p);
  a = v.x;
  b = v.y;
}
template <>
__device__ inline void pair_load<__half>(const __half *p, float &a, float &b) {
  const __half2 v = *reinterpret_cast<const __half2 *>(p);
  a = __low2float(v);
  b = __high2float(v);
}

// Streaming (nontemporal) store of both components at p (an even element).
template <class T>
__device__ inline void pair_store_nt(T *p, float a, float b);
template <>
__device__ inline void pair_store_nt<float>(float *p, float a, float b) {
  const uint64_t v = ((uint64_t)__float_as_uint(b) << 32) | __float_as_uint(a);
  __builtin_nontemporal_store(v, reinterpret_cast<uint64_t *>(p));
}
template <>
__device__ inline void pair_store_nt<__half>(__half *p, float a, float b) {
  // both conversions in one packed instruction, rounded to nearest even as
  // __float2half_rn rounds
  typedef float F2 __attribute__((ext_vector_type(2)));
  typedef _Float16 H2 __attribute__((ext_vector_type(2)));
  const F2 in = {materialise(a), materialise(b)};
  const uint32_t v = __builtin_bit_cast(uint32_t, __builtin_convertvector(in, H2));
  __builtin_nontemporal_store(v, reinterpret_cast<uint32_t *>(p));
}

struct ShiftParams {
  float tau[2], itau[2], omega[2], rho0[2];  // itau = 1.0f / tau, rounded once on the host
  float gcc[2][2], gwc[2];
  float gacc[3], body[2][3];
  int eq_form, psi_form, sc, has_solid, has_body;
};

// psi = 1 - exp(-rho) as -expm1(-rho): in the light phase (rho ~ 0.05) the
// difference 1 - exp(-rho) cancels twenty-fold, so a fast exp's error of
// about one FP32 ulp of 1 became a relative error of 2e-6 in psi.
__device__ inline float psi_f(int form, float rho) { return form == 0 ? rho : -expm1f(-rho); }

// A constant as a float-float pair and the product x * (hi + lo), rounded once.
struct FF {
  float hi, lo;
};
constexpr FF ff_of(double v) { return FF{(float)v, (float)(v - (double)(float)v)}; }
constexpr FF kW0 = ff_of(1.0 / 3.0), kW1 = ff_of(1.0 / 18.0), kW2 = ff_of(1.0 / 36.0);
constexpr FF kInv216 = ff_of(1.0 / 216.0);
__device__ inline float ffmul(float x, FF c) { return fmaf(x, c.hi, x * c.lo); }
// D3Q19 weight of population i (connectivity.d:252-256) as a float-float pair.
__device__ inline FF d3q19_wff(int i) { return i == 0 ? kW0 : (i < 7 ? kW1 : kW2); }

// Deviation of the equilibrium from the reference state,
//   d_i^eq = w_i (rho - rho0) + rho (E_i(u) - w_i),
// for D3Q19, with E_i(u) the equilibrium of unit density; drh = rho - rho0.
// The weight classes share their products: w_i drh takes one float-float
// product per class (w_0 = 1/3, w_1..6 = 1/18, w_7..18 = 1/36), and BDist2's
// 216 (E_i - w_i) (exact constants) meets rho / 216, formed once, in one FMA.
__device__ inline void eq_dev_d3q19(int form, const float (&u)[3], float rho, float drh,
                                    float (&e)[19]) {
  using L = D3Q19;
  if (form != kBDist2) {
    const float uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      const float cu = (float)L::c[i][0] * u[0] + (float)L::c[i][1] * u[1] + (float)L::c[i][2] * u[2];
      float t = fmaf(4.5f * cu, cu, fmaf(3.0f, cu, -1.5f * uu));
      if (form == kThirdOrder) t += (4.5f * cu) * fmaf(cu, cu, -uu);
      e[i] = ffmul(fmaf(rho, t, drh), d3q19_wff(i));
    }
    return;
  }
  // BDist2 (eqdist.d:102-184) with the constant part w_i removed, scaled by
  // 216 so that every constant is exact (aN1 = 1/36 -> 6, aN1 / 2 -> 3,
  // aN1 / 6 -> 1, 1/3 -> 72)
  const float tx = 3.0f * u[0], ty = 3.0f * u[1], tz = 3.0f * u[2];
  const float uke = 0.5f * (u[0] * tx + u[1] * ty + u[2] * tz);
  const float c1 = 6.0f * (1.0f - uke);
  auto dev = [&](float cdu, float &ev, float &od) {
    ev = fmaf(3.0f * cdu, cdu, -6.0f * uke);  // 216 (even - aN1)
    od = fmaf(cdu * cdu, cdu, c1 * cdu);      // 216 odd
  };
  float E[19];
  float ev, od;
  dev(tx, ev, od);
  E[1] = 2.0f * (ev + od);
  E[2] = 2.0f * (ev - od);
  dev(ty, ev, od);
  E[3] = 2.0f * (ev + od);
  E[4] = 2.0f * (ev - od);
  dev(tz, ev, od);
  E[5] = 2.0f * (ev + od);
  E[6] = 2.0f * (ev - od);
  dev(tx + ty, ev, od);
  E[7] = ev + od;
  E[14] = ev - od;
  dev(tx - ty, ev, od);
  E[13] = ev + od;
  E[8] = ev - od;
  dev(tx + tz, ev, od);
  E[11] = ev + od;
  E[18] = ev - od;
  dev(tx - tz, ev, od);
  E[12] = ev + od;
  E[17] = ev - od;
  dev(ty + tz, ev, od);
  E[9] = ev + od;
  E[16] = ev - od;
  dev(ty - tz, ev, od);
  E[15] = ev + od;
  E[10] = ev - od;
  E[0] = -72.0f * uke;
  const float r216 = ffmul(rho, kInv216);
  const float dw0 = ffmul(drh, kW0), dw1 = ffmul(drh, kW1), dw2 = ffmul(drh, kW2);
#pragma unroll
  for (int i = 0; i < 19; ++i) e[i] = fmaf(r216, E[i], i == 0 ? dw0 : (i < 7 ? dw1 : dw2));
}

// All ones when bit i of w is set, else zero.  The empty asm hides that the
// mask is one of two values: the compiler would otherwise turn the blends
// below back into selects and then move the load of the unused operand into a
// branch of its own (one branch and one wait per neighbour).
__device__ inline uint32_t bit_mask(uint32_t w, int i) {
  uint32_t m = 0u - ((w >> i) & 1u);
  asm("" : "+v"(m));
  return m;
}

// a where the mask is set, else b, bit for bit.
__device__ inline float blend(uint32_t m, float a, float b) {
  return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// sum_i w_i c_i v_i (i = 1..18) for D3Q19, as a sum over the nine pairs of
// opposite links c_i = -c_j of w_i c_i (v_i - v_j): near an interface the
// v_i are O(1) and the sum a gradient ten times smaller, and a difference of
// neighbouring values rounds less than a running sum of the values (on the
// channel the velocity error against FP64 stayed the same; the pairs take
// fewer instructions).  The unweighted sums of the two weight classes then
// take one float-float product by w_1 (w_2 = w_1 / 2 exactly).
template <class V>
__device__ inline void weighted_c_sum(V v, float (&acc)[3]) {
  using L = D3Q19;
  float a1[3] = {0.0f, 0.0f, 0.0f}, a2[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 1; i < 19; ++i) {
    if (L::bounce[i] < i) continue;  // each pair once, from its lower index
    const float dv = v(i) - v(L::bounce[i]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (L::c[i][d] == 0) continue;
      float &a = i < 7 ? a1[d] : a2[d];
      a = L::c[i][d] > 0 ? a + dv : a - dv;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) acc[d] = ffmul(fmaf(0.5f, a2[d], a1[d]), kW1);
}

__device__ inline float shfl_xor32f(float v) { return __shfl_xor(v, 32); }

// The parameters of the component a lane holds, selected once.
struct ShiftLane {
  float rho0, tau_f, itau_f, itau_o, om_f, g_self, g_other, wc_f, body_f[3];
};

__device__ inline ShiftLane shift_lane(const ShiftParams &p, int f) {
  auto sel = [f](float a0, float a1) { return f ? a1 : a0; };
  ShiftLane l;
  l.rho0 = sel(p.rho0[0], p.rho0[1]);
  l.tau_f = sel(p.tau[0], p.tau[1]);
  l.itau_f = sel(p.itau[0], p.itau[1]);
  l.itau_o = sel(p.itau[1], p.itau[0]);
  l.om_f = sel(p.omega[0], p.omega[1]);
  l.g_self = sel(p.gcc[0][0], p.gcc[1][1]);
  l.g_other = sel(p.gcc[0][1], p.gcc[1][0]);
  l.wc_f = sel(p.gwc[0], p.gwc[1]);
  for (int d = 0; d < 3; ++d) l.body_f[d] = sel(p.body[0][d], p.body[1][d]);
  return l;
}

// Sum of the 19 stored deviations in i order (rho - rho0).
__device__ inline float shifted_drho(const float *n) {
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 19; ++i) s += n[i];
  return s;
}

// First moment sum_i d_i c_i over the c_i != 0 terms in i order.
__device__ inline void shifted_moment(const float *n, float (&j)[3]) {
  using L = D3Q19;
#pragma unroll
  for (int d = 0; d < 3; ++d) j[d] = 0.0f;
#pragma unroll
  for (int i = 0; i < 19; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (L::c[i][d] != 0) j[d] += (float)L::c[i][d] * n[i];
}

// Shan-Chen and wall force (force.d:238-309) and body force on component f
// at one fluid site.  sw: the site's solid word; psi(ff, i): psi_ff at
// x + c_i (i < 0: at x); wall_rho(i, b): rho_f at the solid neighbour x + c_i
// when b, else any value read in bounds (the site's own).
template <bool SOLID, class Psi, class WallRho>
__device__ inline void shifted_force(int f, uint32_t sw, const ShiftParams &p,
                                     const ShiftLane &pl, Psi psi, WallRho wall_rho,
                                     float (&F)[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) F[d] = 0.0f;
  if (p.sc) {
    const float psi1 = psi(f, -1);
    // the other component first, then the own one: for the usual coupling
    // matrix (zero diagonal) both lanes of a site take the same branch
#pragma unroll
    for (int which = 0; which < 2; ++which) {
      const int f2 = which ? f : (f ^ 1);
      const float cc = which ? pl.g_self : pl.g_other;
      if (cc == 0.0f) continue;
      const float psi_self = psi(f2, -1);
      float acc[3];
      weighted_c_sum(
          [&](int i) {
            const float psi_n = psi(f2, i);
            return SOLID ? blend(bit_mask(sw, i), psi_self, psi_n) : psi_n;
          },
          acc);
      const float pc = psi1 * cc;
#pragma unroll
      for (int d = 0; d < 3; ++d) F[d] = fmaf(-pc, acc[d], F[d]);
    }
    const float wc = pl.wc_f;
    // wall force: sum over the solid neighbours; skipped by waves none of
    // whose sites has one
    if (SOLID && wc != 0.0f && __any((sw & 0x7fffeu) != 0u)) {
      float acc[3];
      weighted_c_sum(
          [&](int i) {
            const uint32_t m = bit_mask(sw, i);
            return blend(m, wall_rho(i, m != 0u), 0.0f);  // every lane loads
          },
          acc);
      const float pw = psi1 * wc;
#pragma unroll
      for (int d = 0; d < 3; ++d) F[d] = fmaf(-pw, acc[d], F[d]);
    }
  }
  if (p.has_body) {
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] += pl.body_f[d];
  }
}

// Velocity shift, equilibrium and BGK relaxation of component f at one fluid
// site, in deviation form (force.d, collision.d, eqdist.d).  cur: streamed
// deviations; rho_p, j_p: the partner component's density and first moment;
// F: the force on f; out(i, d): the new deviation of population i.
template <class Out>
__device__ inline void shifted_relax(const float *cur, float rho_cur, float rho_p,
                                     const float (&j)[3], const float (&j_p)[3],
                                     const float (&F)[3], const ShiftParams &p,
                                     const ShiftLane &pl, Out out) {
  constexpr int Q = 19;
  const float rinv = 1.0f / rho_cur;
  float u[3];
  if (p.eq_form == kBDist2) {
    const float tot_inv = 1.0f / (rho_cur + rho_p);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float wv = fmaf(j[d], pl.itau_f, j_p[d] * pl.itau_o) * tot_inv;
      u[d] = fmaf(pl.tau_f, fmaf(F[d], rinv, p.gacc[d]), wv);
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = fmaf(pl.tau_f, fmaf(F[d], rinv, p.gacc[d]), j[d] * rinv);
  }
  // rho - rho0 summed from the deviations (not rho_cur - rho0, which would
  // bring back the rounding of rho_cur)
  const float drh = shifted_drho(cur);
  float deq[Q];
  eq_dev_d3q19(p.eq_form, u, rho_cur, drh, deq);
#pragma unroll
  for (int i = 0; i < Q; ++i) out(i, fmaf(pl.om_f, deq[i] - cur[i], cur[i]));
}

// Force and relaxation of component f at one fluid site (the two above in
// turn).  Every schedule of shifted storage goes through shifted_force and
// shifted_relax, so they all give the same bits.
template <bool SOLID, class Psi, class WallRho, class Out>
__device__ inline void shifted_collide_site(int f, const float *cur, float rho_cur, float rho_p,
                                            const float (&j)[3], const float (&j_p)[3],
                                            uint32_t sw, const ShiftParams &p,
                                            const ShiftLane &pl, Psi psi, WallRho wall_rho,
                                            Out out) {
  float F[3];
  shifted_force<SOLID>(f, sw, p, pl, psi, wall_rho, F);
  shifted_relax(cur, rho_cur, rho_p, j, j_p, F, p, pl, out);
}

}  // namespace dlbc
