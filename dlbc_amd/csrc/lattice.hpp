// lattice.hpp -- connectivities and per-site physics shared by all kernels.
//
// Every function restates one reference routine (cited) with the same
// operation order, so that FP64 results are bit-identical to the reference's
// D code compiled without fused multiply-add (build with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "crexp.hpp"

namespace dlbc {

enum EqForm : int { kSecondOrder = 0, kThirdOrder = 1, kBDist2 = 2 };

// Connectivities (src/dlbc/lb/connectivity.d:198-418).  A d < 3 lattice is
// stored as a 3-D one with unit extent in the unused axes; its velocities have
// zero components there.
struct D3Q19 {
  static constexpr int d = 3, q = 19;
  static constexpr double css = 1.0 / 3.0;  // c_s^2, connectivity.d:214-248
  // connectivity.d:401-418
  static constexpr int c[19][3] = {
      {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
      {1, 1, 0},  {-1, 1, 0},  {0, 1, 1},  {0, -1, 1},  {1, 0, 1},  {1, 0, -1},  {1, -1, 0},
      {-1, -1, 0}, {0, 1, -1}, {0, -1, -1}, {-1, 0, 1}, {-1, 0, -1}};
  // generateBounce (connectivity.d:198-212): index of -c_i
  static constexpr int bounce[19] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 16, 15, 18, 17, 8, 7, 10, 9, 12, 11};
  // connectivity.d:252-256
  static constexpr double w[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                   1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

struct D2Q9 {
  static constexpr int d = 2, q = 9;
  static constexpr double css = 1.0 / 3.0;
  // connectivity.d:344-374
  static constexpr int c[9][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                  {1, 1, 0},  {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}};
  static constexpr int bounce[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};
  // connectivity.d:267-271
  static constexpr double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

struct D1Q3 {
  static constexpr int d = 1, q = 3;
  static constexpr double css = 1.0 / 3.0;
  // connectivity.d:318-324
  static constexpr int c[3][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}};
  static constexpr int bounce[3] = {0, 2, 1};
  // connectivity.d:285-288
  static constexpr double w[3] = {4.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
};

// connectivity.d:236-238 (c_s^2 = 1), 279-283, 347-353: the +-2 links reach
// two sites out, so the reference gives this lattice a halo of 4
// (lattice.d:96-103); here the periodic wrap serves any length >= 4.
struct D1Q5 {
  static constexpr int d = 1, q = 5;
  static constexpr double css = 1.0;
  static constexpr int c[5][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {2, 0, 0}, {-2, 0, 0}};
  static constexpr int bounce[5] = {0, 2, 1, 4, 3};
  static constexpr double w[5] = {6.0 / 12.0, 2.0 / 12.0, 2.0 / 12.0, 1.0 / 12.0, 1.0 / 12.0};
};

constexpr double kCss = D3Q19::css;

// density(): std.algorithm.sum = 0.0 + pairwise sum by recursive halving
// (src/dlbc/lb/density.d:37-39).
template <int B, int N, class T>
__host__ __device__ inline T pairwise(const T *r) {
  if constexpr (N == 1) {
    return r[B];
  } else if constexpr (N == 2) {
    return r[B] + r[B + 1];
  } else {
    return pairwise<B, N / 2>(r) + pairwise<B + N / 2, N - N / 2>(r);
  }
}

template <class L, class T>
__host__ __device__ inline T density(const T *pop) {
  return T(0) + pairwise<0, L::q>(pop);
}

// RN(a / c) for a constant divisor c with rc = RN(1/c), by Markstein's
// correction: q0 = RN(a rc) is within 1.5 ulp of a/c, one correction
// q1 = RN(q0 + RN(a - q0 c) rc) makes it faithful, and for a faithful q1 the
// remainder a - q1 c is exact and RN(q1 + (a - q1 c) rc) is the correctly
// rounded quotient (Markstein's theorem).  Five fused operations instead of
// the general IEEE division sequence.  Valid while a, a/c and the
// intermediates are normal numbers; the equilibria below only divide
// velocity moments (|a| << 2^1000) and a result below 2^-900 only ever enters
// a sum with 1.0, where any value that small gives the same rounded sum.
__host__ __device__ inline double div_const(double a, double c, double rc) {
  const double q0 = a * rc;
  const double q1 = fma(fma(-q0, c, a), rc, q0);
  return fma(fma(-q1, c, a), rc, q1);
}

// eqDist (src/dlbc/lb/eqdist.d:61-278).  pop: the population, dv: the
// velocity shift, rho0: its density.
template <class L>
__device__ inline void eq_dist(int form, const double *pop, const double *dv, double rho0,
                               double *dist) {
  constexpr int Q = L::q, D = L::d;
  if (form != kBDist2) {
    double pv[D], v[D];
    // rawVelocity (lb/velocity.d:37-48): vel[j] += n_i c_ij over i, / rho
#pragma unroll
    for (int j = 0; j < D; ++j) pv[j] = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) pv[j] += pop[i] * static_cast<double>(L::c[i][j]);
#pragma unroll
    for (int j = 0; j < D; ++j) pv[j] /= rho0;
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = dv[j] + pv[j];
    // dotProduct (range.d:57-67): sequential from 0
    double vdotv = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) vdotv += v[j] * v[j];
    // c_s^2 products exactly as the reference forms them; every quotient
    // by them is the correctly rounded one (div_const)
    constexpr double css = L::css;
    constexpr double css2 = 2.0 * css;
    constexpr double css2p2 = css2 * css;
    constexpr double css6p3 = 3.0 * css2p2 * css;
    constexpr double r_css = 1.0 / css, r_css2 = 1.0 / css2, r_css2p2 = 1.0 / css2p2,
                     r_css6p3 = 1.0 / css6p3;
    const double vv_css2 = div_const(vdotv, css2, r_css2);
    if (form == kSecondOrder) {  // eqdist.d:82-89
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        double vc = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) vc += v[j] * static_cast<double>(L::c[i][j]);
        dist[i] = rho0 * L::w[i] *
                  (1.0 + div_const(vc, css, r_css) + div_const(vc * vc, css2p2, r_css2p2) -
                   vv_css2);
      }
    } else {  // ThirdOrder, eqdist.d:90-100
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        double vc = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) vc += v[j] * static_cast<double>(L::c[i][j]);
        dist[i] = rho0 * L::w[i] *
                  (1.0 + div_const(vc, css, r_css) + div_const(vc * vc, css2p2, r_css2p2) -
                   vv_css2 + div_const(vc * vc * vc, css6p3, r_css6p3) -
                   div_const(vdotv * vc, css2p2, r_css2p2));
      }
    }
    return;
  }
  // BDist2
  if constexpr (D == 3 && Q == 19) {  // eqdist.d:102-184
    const double aN1 = 1.0 / 36.0, aN0 = 1.0 / 3.0;
    const double haN1 = 0.5 * aN1, aN1_6 = aN1 / 6.0;
    const double T = 1.0 / 3.0, holdit = 1.0 / T;
    const double tux = dv[0], tuy = dv[1], tuz = dv[2];
    const double tuxt = tux * holdit, tuyt = tuy * holdit, tuzt = tuz * holdit;
    const double uke = 0.5 * (tux * tuxt + tuy * tuyt + tuz * tuzt);
    const double cst = 1.0 - uke;
    const double const1 = aN1 * cst;
    auto eo = [&](double cdotu, double &even, double &odd) {
      even = const1 + haN1 * cdotu * cdotu;
      odd = const1 * cdotu + aN1_6 * cdotu * cdotu * cdotu;
    };
    double e, o;
    eo(tuxt, e, o);
    dist[1] = 2.0 * (e + o);
    dist[2] = 2.0 * (e - o);
    eo(tuyt, e, o);
    dist[3] = 2.0 * (e + o);
    dist[4] = 2.0 * (e - o);
    eo(tuzt, e, o);
    dist[5] = 2.0 * (e + o);
    dist[6] = 2.0 * (e - o);
    eo(tuxt + tuyt, e, o);
    dist[7] = e + o;
    dist[14] = e - o;
    eo(tuxt - tuyt, e, o);
    dist[13] = e + o;
    dist[8] = e - o;
    eo(tuxt + tuzt, e, o);
    dist[11] = e + o;
    dist[18] = e - o;
    eo(tuxt - tuzt, e, o);
    dist[12] = e + o;
    dist[17] = e - o;
    eo(tuyt + tuzt, e, o);
    dist[9] = e + o;
    dist[16] = e - o;
    eo(tuyt - tuzt, e, o);
    dist[15] = e + o;
    dist[10] = e - o;
    dist[0] = aN0 * cst;
#pragma unroll
    for (int i = 0; i < Q; ++i) dist[i] *= rho0;
  } else if constexpr (D == 2 && Q == 9) {  // eqdist.d:185-235
    const double aN1 = 1.0 / 24.0, aN0 = 1.0 / 2.0;
    const double haN1 = 0.5 * aN1, aN1_6 = aN1 / 6.0;
    const double T = 1.0 / 3.0, holdit = 1.0 / T;
    const double tux = dv[0], tuy = dv[1];
    const double tuxt = tux * holdit, tuyt = tuy * holdit;
    const double uke = 0.5 * (tux * tuxt + tuy * tuyt);
    const double cst = 1.0 - uke;
    const double const1 = aN1 * cst;
    auto eo = [&](double cdotu, double &even, double &odd) {
      even = const1 + haN1 * cdotu * cdotu;
      odd = const1 * cdotu + aN1_6 * cdotu * cdotu * cdotu;
    };
    double e, o;
    eo(tuxt, e, o);
    dist[1] = 2.0 * (e + o);
    dist[2] = 2.0 * (e - o);
    eo(tuyt, e, o);
    dist[3] = 2.0 * (e + o);
    dist[4] = 2.0 * (e - o);
    eo(tuxt + tuyt, e, o);
    dist[5] = e + o;
    dist[8] = e - o;
    eo(tuxt - tuyt, e, o);
    dist[6] = e + o;
    dist[7] = e - o;
    dist[0] = aN0 * cst;
#pragma unroll
    for (int i = 0; i < Q; ++i) dist[i] *= rho0;
  } else {  // eqdist.d:269-271: other connectivities return the population
#pragma unroll
    for (int i = 0; i < Q; ++i) dist[i] = pop[i];
  }
}

__device__ inline double psi_of(int psi_form, double rho) {
  return psi_form == 0 ? rho : psi_exponential(rho);  // lb/force.d:320-332
}

}  // namespace dlbc
