// host_init.cpp -- the reference's initial conditions on the host
// (src/dlbc/fields/init.d, src/dlbc/lb/init.d:263-339, src/dlbc/lb/mask.d).
//
// Fields are written into halo-padded arrays [n0+2h]..[q] exactly as the
// reference fills its own: every stored site including the halo, row-major,
// with one random draw per site (per value for ConstRandom) in that order.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/dlbc_hip.h"
#include "lattice.hpp"

namespace {

// ---- std.random.Mt19937 (src/dlbc/random.d:24) ------------------------------
struct MT {
  uint32_t mt[624];
  uint32_t mti;
  uint32_t front;
  uint32_t pad;
};
static_assert(sizeof(MT) <= sizeof(dlbc_rng), "rng state size");

uint32_t mt_raw(MT &g) {
  if (g.mti >= 624) {
    for (int k = 0; k < 624; ++k) {
      const uint32_t y = (g.mt[k] & 0x80000000u) | (g.mt[(k + 1) % 624] & 0x7fffffffu);
      g.mt[k] = g.mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    g.mti = 0;
  }
  uint32_t y = g.mt[g.mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

MT &as_mt(dlbc_rng *r) { return *reinterpret_cast<MT *>(r->state); }

// ---- connectivity weights and eqDistUnity (src/dlbc/lb/eqdist.d:310-315) ----
// For SecondOrder/ThirdOrder the unit equilibrium is the weight vector; for
// BDist2 it is LB3D's constant set (D3Q19: 1/3, 2/36, 1/36; D2Q9: 1/2, 1/12,
// 1/24); for D1Q3 BDist2 returns the population [1, 0, 0] itself.
bool unity(int d, int q, int form, double *u) {
  if (d == 3 && q == 19) {
    if (form == 2) {
      const double aN1 = 1.0 / 36.0, aN0 = 1.0 / 3.0;
      u[0] = aN0 * 1.0;
      for (int i = 1; i < 7; ++i) u[i] = 2.0 * (aN1 * 1.0 + 0.0);
      for (int i = 7; i < 19; ++i) u[i] = aN1 * 1.0 + 0.0;
      for (int i = 0; i < 19; ++i) u[i] *= 1.0;
    } else {
      u[0] = 1.0 / 3.0;
      for (int i = 1; i < 7; ++i) u[i] = 1.0 / 18.0;
      for (int i = 7; i < 19; ++i) u[i] = 1.0 / 36.0;
    }
    return true;
  }
  if (d == 2 && q == 9) {
    if (form == 2) {
      const double aN1 = 1.0 / 24.0;
      u[0] = 1.0 / 2.0;
      for (int i = 1; i < 5; ++i) u[i] = 2.0 * aN1;
      for (int i = 5; i < 9; ++i) u[i] = aN1;
    } else {
      u[0] = 4.0 / 9.0;
      for (int i = 1; i < 5; ++i) u[i] = 1.0 / 9.0;
      for (int i = 5; i < 9; ++i) u[i] = 1.0 / 36.0;
    }
    return true;
  }
  if (d == 1 && q == 3) {
    if (form == 2) {
      u[0] = 1.0;
      u[1] = u[2] = 0.0;
    } else {
      u[0] = 4.0 / 6.0;
      u[1] = u[2] = 1.0 / 6.0;
    }
    return true;
  }
  return false;
}

// symmetricLinearTransition (src/dlbc/fields/init.d:346-366)
double transition(double pos, double width, double negval, double posval) {
  if (width == 0.0) return pos < 0 ? negval : posval;
  if (pos < -0.5 * width) return negval;
  if (pos > 0.5 * width) return posval;
  return negval * (0.5 - pos / width) + posval * (0.5 + pos / width);
}

struct Box {
  int d;
  int64_t n[3], nh[3];
  int h;
  int64_t sites_h() const { return nh[0] * nh[1] * nh[2]; }
};

bool make_box(Box &b, int d, const int64_t *n, int halo) {
  if (d < 1 || d > 3 || halo < 0 || !n) return false;
  b.d = d;
  b.h = halo;
  for (int i = 0; i < 3; ++i) {
    b.n[i] = i < d ? n[i] : 1;
    b.nh[i] = i < d ? n[i] + 2 * halo : 1;
    if (b.n[i] < 1) return false;
  }
  return true;
}

// dotProduct (src/dlbc/range.d:57-67)
double dot(const double *a, int d) {
  double s = 0;
  for (int i = 0; i < d; ++i) s += a[i] * a[i];
  return s;
}

}  // namespace

extern "C" {

void dlbc_rng_seed(dlbc_rng *r, uint32_t seed) {
  MT &g = as_mt(r);
  g.mt[0] = seed;  // init_genrand
  for (uint32_t i = 1; i < 624; ++i) g.mt[i] = 1812433253u * (g.mt[i - 1] ^ (g.mt[i - 1] >> 30)) + i;
  g.mti = 624;
  g.front = mt_raw(g);  // seed() primes front with the first output
}

uint32_t dlbc_rng_next(dlbc_rng *r) {
  MT &g = as_mt(r);
  const uint32_t v = g.front;
  g.front = mt_raw(g);
  return v;
}

// uniform!"[)"(a, b) = a + (nextafter(b, -inf) - a) * (front - min) / (max - min)
double dlbc_rng_uniform(dlbc_rng *r, double a, double b) {
  const double bb = std::nextafter(b, -INFINITY);
  const double x = static_cast<double>(dlbc_rng_next(r));
  return a + (bb - a) * x / 4294967295.0;
}

int dlbc_init_fluid(double *host, int d, int q, const int64_t *n, int halo, int form, int kind,
                    const double *dens, int ndens, double perturb, dlbc_rng *rng, double radius,
                    const double *offset, const double *separation, int axis, double thickness,
                    const double *widths) {
  Box b;
  double u[19];
  if (!host || !make_box(b, d, n, halo) || !unity(d, q, form, u)) return DLBC_ERR_ARG;
  const int64_t S = b.sites_h();
  auto needs = [&](int k) { return dens && ndens >= k; };
  switch (kind) {
    case DLBC_INIT_NONE:
    case DLBC_INIT_FILE:
      return DLBC_OK;
    case DLBC_INIT_CONST:  // initConst, init.d:26-37
      if (!needs(1)) return DLBC_ERR_ARG;
      for (int64_t s = 0; s < S * q; ++s) host[s] = dens[0];
      return DLBC_OK;
    case DLBC_INIT_CONST_RANDOM:  // init.d:39-50
      if (!needs(1) || !rng) return DLBC_ERR_ARG;
      for (int64_t s = 0; s < S * q; ++s) host[s] = dens[0] * dlbc_rng_uniform(rng, 0.0, 2.0);
      return DLBC_OK;
    case DLBC_INIT_EQDIST: {  // init.d:52-58
      if (!needs(1)) return DLBC_ERR_ARG;
      double e[19];
      for (int i = 0; i < q; ++i) e[i] = dens[0] * u[i];
      for (int64_t s = 0; s < S; ++s) std::memcpy(host + s * q, e, sizeof(double) * q);
      return DLBC_OK;
    }
    case DLBC_INIT_EQDIST_RANDOM:  // init.d:60-66
    case DLBC_INIT_EQDIST_PERTURB:  // init.d:68-74
    case DLBC_INIT_EQDIST_PERTURB_FRAC: {  // init.d:76-82
      if (!needs(1) || !rng) return DLBC_ERR_ARG;
      for (int64_t s = 0; s < S; ++s) {
        double a;
        if (kind == DLBC_INIT_EQDIST_RANDOM)
          a = dens[0] * dlbc_rng_uniform(rng, 0.0, 2.0);
        else if (kind == DLBC_INIT_EQDIST_PERTURB)
          a = dens[0] + dlbc_rng_uniform(rng, -perturb, perturb);
        else
          a = dens[0] * (1.0 + dlbc_rng_uniform(rng, -perturb, perturb));
        for (int i = 0; i < q; ++i) host[s * q + i] = a * u[i];
      }
      return DLBC_OK;
    }
    default:
      break;
  }
  // Shapes (init.d:84-274).  Fractional variants scale by the global lengths.
  const bool frac = kind == DLBC_INIT_EQDIST_SPHERE_FRAC || kind == DLBC_INIT_EQDIST_TWO_SPHERES_FRAC ||
                    kind == DLBC_INIT_EQDIST_CYLINDER_FRAC || kind == DLBC_INIT_EQDIST_LAMELLAE_FRAC;
  double off[3] = {0, 0, 0}, sep[3] = {0, 0, 0}, rad = radius;
  for (int i = 0; i < d; ++i) {
    off[i] = offset ? offset[i] : 0.0;
    sep[i] = separation ? separation[i] : 0.0;
  }
  if (frac && (kind == DLBC_INIT_EQDIST_SPHERE_FRAC || kind == DLBC_INIT_EQDIST_TWO_SPHERES_FRAC)) {
    int64_t small = b.n[0];  // init.d:112-118
    for (int i = 0; i < d; ++i) small = b.n[i] < small ? b.n[i] : small;
    rad = radius * (double)small;
    for (int i = 0; i < d; ++i) {
      off[i] = off[i] * (double)b.n[i];
      sep[i] = sep[i] * (double)b.n[i];
    }
  } else if (kind == DLBC_INIT_EQDIST_CYLINDER_FRAC) {
    int64_t small = 0;  // init.d:217-224
    for (int i = 0; i < d; ++i)
      if (i != axis && (b.n[i] < small || small == 0)) small = b.n[i];
    rad = radius * (double)small;
    for (int i = 0; i < d; ++i) off[i] = off[i] * (double)b.n[i];
  }
  std::vector<double> inter;
  if (kind == DLBC_INIT_EQDIST_LAMELLAE || kind == DLBC_INIT_EQDIST_LAMELLAE_FRAC) {
    if (axis < 0 || axis >= d || !widths || !needs(1)) return DLBC_ERR_ARG;
    inter.push_back(0.0);  // calculateInterfaces, init.d:335-344
    for (int i = 0; i < ndens; ++i)
      inter.push_back(kind == DLBC_INIT_EQDIST_LAMELLAE_FRAC ? widths[i] * (double)b.n[axis]
                                                              : widths[i]);
    for (size_t i = 1; i < inter.size(); ++i) inter[i] += inter[i - 1];
    for (double &v : inter) v -= 0.5;
  } else if (!needs(2)) {
    return DLBC_ERR_ARG;
  }
  const double hd = (double)b.h;
  int64_t s = 0;
  for (int64_t x = 0; x < b.nh[0]; ++x)
    for (int64_t y = 0; y < b.nh[1]; ++y)
      for (int64_t z = 0; z < b.nh[2]; ++z, ++s) {
        const int64_t p[3] = {x, y, z};
        double val;
        if (kind == DLBC_INIT_EQDIST_SPHERE || kind == DLBC_INIT_EQDIST_SPHERE_FRAC) {
          double o[3];  // init.d:94-102
          for (int i = 0; i < d; ++i)
            o[i] = ((double)p[i] - hd) - ((double)b.n[i] * 0.5 + off[i]) + 0.5;
          val = transition(std::sqrt(dot(o, d)) - rad, thickness, dens[0], dens[1]);
        } else if (kind == DLBC_INIT_EQDIST_TWO_SPHERES ||
                   kind == DLBC_INIT_EQDIST_TWO_SPHERES_FRAC) {
          double o1[3], o2[3];  // init.d:142-153
          for (int i = 0; i < d; ++i) {
            const double g = (double)p[i] - hd;
            o1[i] = g - ((double)b.n[i] * 0.5 + off[i] - 0.5 * sep[i]) + 0.5;
            o2[i] = g - ((double)b.n[i] * 0.5 + off[i] + 0.5 * sep[i]) + 0.5;
          }
          const double r1 = std::sqrt(dot(o1, d)) - rad, r2 = std::sqrt(dot(o2, d)) - rad;
          val = transition(r2 < r1 ? r2 : r1, thickness, dens[0], dens[1]);
        } else if (kind == DLBC_INIT_EQDIST_CYLINDER || kind == DLBC_INIT_EQDIST_CYLINDER_FRAC) {
          double o[3];  // init.d:194-206
          for (int i = 0; i < d; ++i)
            o[i] = i == axis ? 0.0 : ((double)p[i] - hd) - ((double)b.n[i] * 0.5 + off[i]) + 0.5;
          val = transition(std::sqrt(dot(o, d)) - rad, thickness, dens[0], dens[1]);
        } else if (kind == DLBC_INIT_EQDIST_LAMELLAE || kind == DLBC_INIT_EQDIST_LAMELLAE_FRAC) {
          const double gp = (double)p[axis] - hd;  // init.d:245-258, 315-333
          size_t best = 0;
          double md = std::fabs(inter[0] - gp);
          for (size_t i = 1; i < inter.size(); ++i) {
            const double df = std::fabs(inter[i] - gp);
            if (df < md) {
              md = df;
              best = i;
            }
          }
          const double diff = gp - inter[best];
          if (best == 0)
            val = transition(diff, thickness, dens[ndens - 1], dens[0]);
          else if ((int)best == ndens)
            val = transition(diff, thickness, dens[best - 1], dens[0]);
          else
            val = transition(diff, thickness, dens[best - 1], dens[best]);
        } else {
          return DLBC_ERR_ARG;
        }
        for (int i = 0; i < q; ++i) host[s * q + i] = val * u[i];
      }
  return DLBC_OK;
}

int dlbc_init_mask(int32_t *host, int d, const int64_t *n, int halo, int kind, int axis,
                   int wall_offset, const int *dim_has_wall, int wall_thickness) {
  Box b;
  if (!host || !make_box(b, d, n, halo)) return DLBC_ERR_ARG;
  if ((kind == 2 || kind == 3) && (axis < 0 || axis >= d)) return DLBC_ERR_ARG;
  if (kind == 5 && !dim_has_wall) return DLBC_ERR_ARG;
  int64_t s = 0;
  for (int64_t x = 0; x < b.nh[0]; ++x)
    for (int64_t y = 0; y < b.nh[1]; ++y)
      for (int64_t z = 0; z < b.nh[2]; ++z, ++s) {
        const int64_t p[3] = {x, y, z};
        int32_t e = 0;
        if (kind == 0) {
        } else if (kind == 2) {  // Plates, init.d:425-436
          const int64_t g = p[axis] - b.h;
          if (g == wall_offset || g == b.n[axis] - 1 - wall_offset) e = 1;
        } else if (kind == 3 || kind == 4) {  // Tube, Box: init.d:377-412
          for (int i = 0; i < d; ++i) {
            const int64_t g = p[i] - b.h;
            if ((kind == 4 || i != axis) && (g == 0 || g == b.n[i] - 1)) e = 1;
          }
        } else if (kind == 5) {  // Walls, init.d:449-467
          for (int i = 0; i < d; ++i) {
            const int64_t g = p[i] - b.h;
            if (dim_has_wall[2 * i] && g < wall_thickness) e = 1;
            // the reference compares signed with unsigned here: negative
            // (halo) coordinates pass the upper-wall test too
            if (dim_has_wall[2 * i + 1] && (uint64_t)g >= (uint64_t)(b.n[i] - wall_thickness))
              e = 1;
          }
        } else {
          return DLBC_ERR_ARG;
        }
        host[s] = e;
      }
  return DLBC_OK;
}

int dlbc_init_eq_dist_wall(double *host, const int32_t *mask, int d, int q, const int64_t *n,
                           int halo) {
  Box b;
  if (!host || !mask || !make_box(b, d, n, halo)) return DLBC_ERR_ARG;
  const int64_t S = b.sites_h();
  for (int64_t s = 0; s < S; ++s)
    if (mask[s] == 1) {
      for (int i = 0; i < q; ++i) host[s * q + i] = 0.0;
      host[s * q] = 1.0;
    }
  return DLBC_OK;
}

int64_t dlbc_selftest_division(int64_t n, uint64_t seed) {
  using dlbc::div_const;
  const double css = dlbc::kCss, css2 = 2.0 * css, css2p2 = css2 * css;
  const double css6p3 = 3.0 * css2p2 * css;
  const double cs[4] = {css, css2, css2p2, css6p3};
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {  // splitmix64
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  int64_t bad = 0;
  for (int64_t k = 0; k < n; ++k) {
    const double c = cs[k & 3];
    const uint64_t r = next();
    double a;
    uint64_t bits;
    if ((k >> 2) & 1) {  // random mantissa, exponent in [-200, 200]
      bits = (r & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 - 200 + (int)((r >> 52) % 401)) << 52);
      std::memcpy(&a, &bits, 8);
    } else {  // a few ulp from an exact multiple of the divisor
      bits = (r & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 + (int)((r >> 52) % 64) - 32) << 52);
      double m;
      std::memcpy(&m, &bits, 8);
      a = m * c;
      std::memcpy(&bits, &a, 8);
      bits += (int64_t)(next() % 9) - 4;
      std::memcpy(&a, &bits, 8);
    }
    if (r >> 63) a = -a;
    if (div_const(a, c, 1.0 / c) != a / c) ++bad;
  }
  return bad;
}

}  // extern "C"
