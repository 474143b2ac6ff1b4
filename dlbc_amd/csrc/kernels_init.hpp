// kernels_init.hpp -- the reference's initial conditions written straight into
// the device populations and mask (src/dlbc/lb/init.d:263-339,
// src/dlbc/fields/init.d:68-477, src/dlbc/lb/mask.d:134-172).
//
// A site's value comes from the same functions as the host initialisation
// (init_shapes.hpp), and each population is stored with the same conversion
// as an upload of the host array, so the device state equals host init +
// upload bit for bit.  Random kinds draw on the host (the Mersenne Twister is
// sequential and the reference's draw order is the halo-padded row-major
// order) and hand over one value per site, or per population for
// ConstRandom: 8 B per site instead of the 8 q B of a population upload.
#pragma once

#include <type_traits>

#include "init_shapes.hpp"
#include "kernels_shifted.hpp"
#include "kernels_twopass.hpp"

namespace dlbc {

// A population in the storage type: FP64 as is; shifted storage keeps
// f_i - w_i rho0 rounded to FP32 and then to the storage type (D3Q19 only).
template <class St>
__device__ inline St to_storage(double v, int i, double rho0) {
  if constexpr (std::is_same<St, double>::value) {
    return v;
  } else {
    return from_f<St>((float)(v - D3Q19::w[i] * rho0));
  }
}

// Element of population i of component f at site s: FP64 storage is
// [C][q][sp]; shifted storage keeps the two components side by side
// (shift_elem).
template <class St>
__device__ inline int64_t pop_elem(int f, int i, int s, int q, int64_t sp) {
  if constexpr (std::is_same<St, double>::value)
    return (int64_t)(f * q + i) * sp + s;
  else
    return shift_elem(f, i, s, sp);
}

// Where a site's value comes from.
enum InitMode : int {
  kInitConst = 0,  // value * u_i (Const with u = 1, EqDist)
  kInitSite = 1,   // vals[t] * u_i (EqDistRandom, EqDistPerturb, EqDistPerturbFrac)
  kInitPop = 2,    // vals[t q + i] (ConstRandom)
  kInitShape = 3   // shape_value(p) * u_i (spheres, cylinders, lamellae)
};

struct FluidInitArgs {
  int mode, q, f;
  int64_t halo;
  double value, rho0;
  double u[19];
};

// Sites s0 .. s0+cnt-1 of component f, in both population buffers.
template <class St>
__global__ void __launch_bounds__(256)
    k_init_fluid(St *__restrict__ A, St *__restrict__ B, Geo g, int s0, int cnt,
                 const double *__restrict__ vals, FluidInitArgs a, ShapeSpec sh) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= cnt) return;
  const int s = s0 + t;
  double val = a.value;
  if (a.mode == kInitSite) {
    val = vals[t];
  } else if (a.mode == kInitShape) {
    int x, y, z;
    coords(s, g, x, y, z);
    const int64_t p[3] = {x + a.halo, y + a.halo, z + a.halo};
    val = shape_value(sh, p);
  }
  for (int i = 0; i < a.q; ++i) {
    const double v = a.mode == kInitPop ? vals[(int64_t)t * a.q + i] : val * a.u[i];
    const int64_t o = pop_elem<St>(a.f, i, s, a.q, g.sp);
    const St sv = to_storage<St>(v, i, a.rho0);
    A[o] = sv;
    B[o] = sv;
  }
}

// initEqDistWall (src/dlbc/fields/init.d:469-477): solid sites get [1, 0, ..., 0].
template <class St>
__global__ void __launch_bounds__(256)
    k_eq_dist_wall(St *__restrict__ A, St *__restrict__ B, const uint8_t *__restrict__ mask,
                   Geo g, int q, int f, double rho0) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites || !mask[s]) return;
  for (int i = 0; i < q; ++i) {
    const int64_t o = pop_elem<St>(f, i, s, q, g.sp);
    const St sv = to_storage<St>(i == 0 ? 1.0 : 0.0, i, rho0);
    A[o] = sv;
    B[o] = sv;
  }
}

// initMask shapes (Plates, Tube, Box, Walls) of every physical site.
__global__ void __launch_bounds__(256) k_init_mask(uint8_t *__restrict__ mask, Geo g, MaskSpec m) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites) return;
  int x, y, z;
  coords(s, g, x, y, z);
  const int64_t p[3] = {x + m.h, y + m.h, z + m.h};
  mask[s] = (uint8_t)mask_value(m, p);
}

}  // namespace dlbc
