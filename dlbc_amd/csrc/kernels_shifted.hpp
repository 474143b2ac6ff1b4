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
// where E_i(u) is the equilibrium of unit density.  This file holds the
// storage conversions, parameters and per-site pieces; the kernel that steps
// shifted lattices (alone or in batches) is in kernels_batch.hpp.
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

// Streaming (nontemporal) store of one population in the storage type.
template <class T>
__device__ inline void store_nt(T *p, float v);
template <>
__device__ inline void store_nt<float>(float *p, float v) {
  __builtin_nontemporal_store(v, p);
}
template <>
__device__ inline void store_nt<__half>(__half *p, float v) {
  __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v)),
                              reinterpret_cast<unsigned short *>(p));
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

}  // namespace dlbc
