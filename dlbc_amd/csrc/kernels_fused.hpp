// kernels_fused.hpp -- single-pass D3Q19 two-component kernel (see below).
#pragma once

#include "kernels_twopass.hpp"

namespace dlbc {

inline bool fused_d3q19_supported(const Geo &) { return false; }

inline int launch_fused_d3q19(const double *, double *, const uint8_t *, double *, const Geo &,
                              const StepParams &, hipStream_t) {
  return -1;
}

}  // namespace dlbc
