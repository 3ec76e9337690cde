// rs_builder.cpp - host RadixSpline builder.
//
// The reference builds its RadixSpline with the third-party rs::Builder
// (learnedsystems/RadixSpline, fetched at HEAD by lib/RadixSpline.cmake:9-16;
// call site src/rs.cuh:54-64), which is absent here.  This is a restatement of
// the published method:
//   * spline: GreedySplineCorridor (Neumann & Michel, "Smooth interpolating
//     histograms with error guarantees", BNCOD 2008) over the CDF points
//     (key, first position of key); a knot is emitted when the next CDF point
//     leaves the error corridor [pos - err, pos + err] spanned from the last knot;
//     the first and last keys are always knots;
//   * radix table: 2^radix_bits + 2 entries over the top radix_bits significant
//     bits of (key - min_key); entry p = index of the first knot whose prefix is >= p.
#include "rs_builder.hpp"

#include <cmath>
#include <stdexcept>

namespace gj {
namespace {

template <class K>
class CorridorBuilder {
 public:
  CorridorBuilder(RadixSplineHost& out) : out_(out) {}

  void run(const K* keys, uint64_t n) {
    const double err = static_cast<double>(out_.max_error);
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t key = keys[i];
      const double pos = static_cast<double>(i);
      if (i == 0) {
        emit(key, pos);
        cdf_ = {key, pos};
        last_key_ = key;
        distinct_ = 1;
        continue;
      }
      if (key == last_key_) continue;  // same CDF point
      last_key_ = key;
      ++distinct_;
      const Point up{key, pos + err};
      const Point down{key, pos < err ? 0.0 : pos - err};
      if (distinct_ == 2) {
        upper_ = up;
        lower_ = down;
        cdf_ = {key, pos};
        continue;
      }
      const Point& knot = knots_back();
      const Point d{key, pos};
      if (turn(knot, upper_, d) != kClockwise || turn(knot, lower_, d) != kCounterClockwise) {
        emit(cdf_.x, cdf_.y);  // the corridor cannot reach the new point
        upper_ = up;
        lower_ = down;
      } else {
        if (turn(knot, upper_, up) == kClockwise) upper_ = up;
        if (turn(knot, lower_, down) == kCounterClockwise) lower_ = down;
      }
      cdf_ = {key, pos};
    }
    if (out_.knot_x.back() != last_key_) emit(last_key_, static_cast<double>(n - 1));
    for (uint64_t p = static_cast<uint64_t>(prefix_) + 1; p < out_.radix.size(); ++p)
      out_.radix[p] = static_cast<uint32_t>(out_.knot_x.size());
  }

 private:
  struct Point {
    uint64_t x;
    double y;
  };
  enum Turn { kNone, kClockwise, kCounterClockwise };

  // orientation of (a - o) x (b - o) with a tolerance of one double epsilon
  static Turn turn(const Point& o, const Point& a, const Point& b) {
    const double ax = static_cast<double>(a.x - o.x), ay = a.y - o.y;
    const double bx = static_cast<double>(b.x - o.x), by = b.y - o.y;
    const double cross = std::fma(ay, bx, -std::fma(by, ax, 0.0));
    constexpr double eps = 2.220446049250313e-16;
    if (cross > eps) return kClockwise;
    if (cross < -eps) return kCounterClockwise;
    return kNone;
  }

  const Point& knots_back() {
    back_ = {out_.knot_x.back(), out_.knot_y.back()};
    return back_;
  }

  void emit(uint64_t x, double y) {
    out_.knot_x.push_back(x);
    out_.knot_y.push_back(y);
    const uint64_t prefix = (x - out_.min_key) >> out_.shift_bits;
    if (prefix != prefix_) {
      const uint32_t idx = static_cast<uint32_t>(out_.knot_x.size() - 1);
      for (uint64_t p = prefix_ + 1; p <= prefix; ++p) out_.radix[p] = idx;
      prefix_ = prefix;
    }
  }

  RadixSplineHost& out_;
  Point upper_{}, lower_{}, cdf_{}, back_{};
  uint64_t last_key_ = 0, distinct_ = 0, prefix_ = 0;
};

}  // namespace

RadixSplineHost build_radix_spline(const void* keys, uint64_t n, int key_bytes, uint32_t radix_bits,
                                   uint32_t max_error) {
  if (n == 0) throw std::invalid_argument("empty column");
  RadixSplineHost rs;
  rs.n = n;
  rs.radix_bits = radix_bits;
  rs.max_error = max_error;
  auto key_at = [&](uint64_t i) -> uint64_t {
    return key_bytes == 8 ? static_cast<const uint64_t*>(keys)[i] : static_cast<const uint32_t*>(keys)[i];
  };
  rs.min_key = key_at(0);
  rs.max_key = key_at(n - 1);
  const uint64_t span = rs.max_key - rs.min_key;
  const uint32_t used = span ? 64u - static_cast<uint32_t>(__builtin_clzll(span)) : 0u;
  rs.shift_bits = used < radix_bits ? 0u : used - radix_bits;
  rs.radix.assign((span >> rs.shift_bits) + 2, 0u);
  if (key_bytes == 8) {
    CorridorBuilder<uint64_t>(rs).run(static_cast<const uint64_t*>(keys), n);
  } else {
    CorridorBuilder<uint32_t>(rs).run(static_cast<const uint32_t*>(keys), n);
  }
  return rs;
}

}  // namespace gj
