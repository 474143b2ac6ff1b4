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
#include "init_shapes.hpp"
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
  if (!host || !make_box(b, d, n, halo) || !dlbc::eq_unity(d, q, form, u)) return DLBC_ERR_ARG;
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
        const double a = dlbc::random_density(kind, dens[0], perturb, rng);
        for (int i = 0; i < q; ++i) host[s * q + i] = a * u[i];
      }
      return DLBC_OK;
    }
    default:
      break;
  }
  // Shapes (init.d:84-274)
  dlbc::ShapeSpec sh;
  if (!dlbc::resolve_shape(sh, kind, d, n, halo, dens, ndens, radius, offset, separation, axis,
                           thickness, widths))
    return DLBC_ERR_ARG;
  int64_t s = 0;
  for (int64_t x = 0; x < b.nh[0]; ++x)
    for (int64_t y = 0; y < b.nh[1]; ++y)
      for (int64_t z = 0; z < b.nh[2]; ++z, ++s) {
        const int64_t p[3] = {x, y, z};
        const double val = dlbc::shape_value(sh, p);
        for (int i = 0; i < q; ++i) host[s * q + i] = val * u[i];
      }
  return DLBC_OK;
}

int dlbc_init_mask(int32_t *host, int d, const int64_t *n, int halo, int kind, int axis,
                   int wall_offset, const int *dim_has_wall, int wall_thickness) {
  Box b;
  dlbc::MaskSpec m;
  if (!host || !make_box(b, d, n, halo) ||
      !dlbc::resolve_mask(m, d, n, halo, kind, axis, wall_offset, dim_has_wall, wall_thickness))
    return DLBC_ERR_ARG;
  int64_t s = 0;
  for (int64_t x = 0; x < b.nh[0]; ++x)
    for (int64_t y = 0; y < b.nh[1]; ++y)
      for (int64_t z = 0; z < b.nh[2]; ++z, ++s) {
        const int64_t p[3] = {x, y, z};
        host[s] = dlbc::mask_value(m, p);
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
