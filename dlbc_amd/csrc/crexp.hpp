// crexp.hpp -- correctly rounded exp(double) for host and device.
//
// The Shan-Chen pseudopotential psi = 1 - exp(-rho) (src/dlbc/lb/force.d:327)
// is evaluated with D's std.math.exp, which computes in x87 extended
// precision and rounds to double.  This file provides the correctly rounded
// exp, so that the CPU restatement and every GPU kernel produce the same bits
// from the same density.
//
// Method (Ziv): x = k ln2/64 + r, |r| <= ln2/128, exp(x) = 2^(k/64) e^r with
// 2^(j/64) from a 64-entry double-double table.
//   fast path:     e^r with a degree-7 polynomial and double-double leading
//                  terms, total error < 2^-72; the result is returned when
//                  both ends of a 2^-68 error interval round to the same double
//                  (all but ~2^-14 of arguments);
//   accurate path: degree-11 Taylor series in double-double arithmetic
//                  (error ~2^-100), rounded once; subnormal results are
//                  rounded directly on the 2^-1074 grid.
// Every product error term uses fma(), which is exact on both the CPU and
// CDNA hardware; compile with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace dlbc {

#include "exp_table.inc"

struct DD {
  double hi, lo;
};

__host__ __device__ inline DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

__host__ __device__ inline DD fast_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}

__host__ __device__ inline DD two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}

__host__ __device__ inline DD dd_mul(DD a, DD b) {
  DD p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ inline DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  DD t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

// Round (h + l) * 2^m onto the subnormal grid 2^-1074 (m + 1074 in [-1, 53]).
__host__ __device__ inline double round_subnormal(double h, double l, int m) {
  const int e = m + 1074;
  const double u = ldexp(h, e);
  const double ul = ldexp(l, e);
  double k = rint(u);
  const double frac = (u - k) + ul;
  if (frac > 0.5) {
    k += 1.0;
  } else if (frac < -0.5) {
    k -= 1.0;
  } else if (frac == 0.5 || frac == -0.5) {  // exact tie: to even
    const double kk = k + (frac > 0 ? 1.0 : -1.0);
    if (fmod(kk, 2.0) == 0.0) k = kk;
  }
  return ldexp(k, -1074);
}

__host__ __device__ inline double exp_cr(double x) {
  if (x != x) return x + x;
  if (x > 710.0) return INFINITY;
  if (x < -745.2) return 0.0;
  if (fabs(x) < 0x1p-54) return 1.0;

  // Range reduction: r = x - k ln2/64 = rh + rl.
  const double kd = rint(x * kInvLn2_64);
  const int k = static_cast<int>(kd);
  const double t = x - kd * kLn2_64Hi;  // exact (kd*Hi exact; Sterbenz)
  const DD p = two_prod(kd, kLn2_64Mid);
  DD r = two_sum(t, -p.hi);
  r.lo = (r.lo - p.lo) - kd * kLn2_64Lo;
  r = fast_two_sum(r.hi, r.lo);
  const double rh = r.hi, rl = r.lo;

  const int j = k & 63;
  const int m = (k - j) / 64;
  const double th = kExpTab[2 * j], tl = kExpTab[2 * j + 1];

  if (x >= -708.39) {
    // Fast path: e^r = 1 + rh + rh^2/2 + [rl + rh rl + rh^3/6 (...)]
    const DD r2 = two_prod(rh, rh);
    const double a = 0.5 * r2.hi;
    const double c3 =
        r2.hi * rh *
        (1.0 / 6 + rh * (1.0 / 24 + rh * (1.0 / 120 + rh * (1.0 / 720 + rh * (1.0 / 5040)))));
    const double small = rl + (0.5 * r2.lo + (rh * rl + c3));
    const DD u = two_sum(rh, a);
    const DD p1 = two_prod(th, u.hi);
    const DD h = fast_two_sum(th, p1.hi);
    const double l = h.lo + (p1.lo + (th * (u.lo + small) + (tl + tl * u.hi)));
    const double err = 0x1p-68 * h.hi;
    const double y1 = h.hi + (l - err);
    const double y2 = h.hi + (l + err);
    if (y1 == y2) return ldexp(y1, m);
  }

  // Accurate path: e^r - 1 by Horner in double-double, then T (1 + that).
  DD acc = {kInvFact[18], kInvFact[19]};  // 1/11!
  for (int i = 8; i >= 0; --i) {          // 1/10! .. 1/2!
    acc = dd_mul(acc, r);
    acc = dd_add(acc, DD{kInvFact[2 * i], kInvFact[2 * i + 1]});
  }
  acc = dd_mul(acc, r);
  acc = dd_add(acc, DD{1.0, 0.0});
  acc = dd_mul(acc, r);                    // e^r - 1
  const DD e = dd_add(DD{1.0, 0.0}, acc);  // e^r
  const DD res = dd_mul(DD{th, tl}, e);
  if (x >= -708.39) return ldexp(res.hi + res.lo, m);
  return round_subnormal(res.hi, res.lo, m);
}

// Shan-Chen psi(rho) (src/dlbc/lb/force.d:320-332).
__host__ __device__ inline double psi_exponential(double rho) { return 1.0 - exp_cr(-rho); }

}  // namespace dlbc
