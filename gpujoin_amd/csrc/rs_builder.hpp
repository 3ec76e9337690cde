// rs_builder.hpp - host RadixSpline builder (see rs_builder.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace gj {

struct RadixSplineHost {
  uint64_t min_key = 0, max_key = 0, n = 0;
  uint32_t radix_bits = 0, shift_bits = 0, max_error = 0;
  std::vector<uint32_t> radix;   // (max_prefix + 2) entries
  std::vector<uint64_t> knot_x;  // knot keys (widened to 64 bits)
  std::vector<double> knot_y;    // knot positions
};

// keys: n ascending keys of key_bytes (4 or 8) bytes each.
RadixSplineHost build_radix_spline(const void* keys, uint64_t n, int key_bytes, uint32_t radix_bits,
                                   uint32_t max_error);

}  // namespace gj
