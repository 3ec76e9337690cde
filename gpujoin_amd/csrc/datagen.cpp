// datagen.cpp - seeded workload generators (host, OpenMP).
//
// Successor of generate_datasets and helpers (src/index_lookup_common.cuh:44-206)
// and of zipf_distribution (src/zipf.hpp:13-102).  Every random number is
// gj_rng_u64(stream_seed, element_index) - a splitmix64 output - so the data
// depends only on the arguments, never on the thread count or schedule.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <parallel/algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "gj_runtime.h"

namespace gj {
void set_last_error(const std::string& msg);
}

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * kGamma); }
inline uint64_t sub_seed(uint64_t seed, uint64_t tag) { return mix64(seed ^ (tag * 0xD1B54A32D192ED03ull)); }
// uniform double in [0, 1) with 53 random bits
inline double unit(uint64_t r) { return static_cast<double>(r >> 11) * 0x1.0p-53; }
// uniform integer in [0, n)
inline uint64_t below(uint64_t r, uint64_t n) {
  return static_cast<uint64_t>((static_cast<unsigned __int128>(r) * n) >> 64);
}

int nthreads(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

template <class F>
int guarded(F&& f) {
  try {
    f();
    return GJ_OK;
  } catch (const std::invalid_argument& e) {
    gj::set_last_error(e.what());
    return GJ_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    gj::set_last_error("host allocation failed");
    return GJ_ERR_NOMEM;
  } catch (const std::exception& e) {
    gj::set_last_error(e.what());
    return GJ_ERR_INTERNAL;
  }
}

void require(bool c, const char* msg) {
  if (!c) throw std::invalid_argument(msg);
}

// Keyed bijection on [0, 2^bits): rounds of odd multiply, add and xorshift,
// each a permutation of the bits-bit integers.  The first n outputs are n
// distinct keys spread uniformly over the range.
struct KeyedPerm {
  uint64_t mask;
  unsigned shift;
  uint64_t mul[4], add[4];
  KeyedPerm(unsigned bits, uint64_t seed) {
    mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1ull);
    shift = std::max(1u, bits / 2);
    for (int r = 0; r < 4; ++r) {
      mul[r] = rng(seed, 2 * r) | 1ull;
      add[r] = rng(seed, 2 * r + 1);
    }
  }
  uint64_t operator()(uint64_t x) const {
    for (int r = 0; r < 4; ++r) {
      x = (x * mul[r] + add[r]) & mask;
      x ^= x >> shift;
    }
    return x;
  }
};

template <class K>
void psort(K* b, K* e, int threads) {
  __gnu_parallel::sort(b, e, std::less<K>(), __gnu_parallel::default_parallel_tag(nthreads(threads)));
}

// Draws keys by `sample(i)` for increasing i until n distinct values exist;
// drops any excess at seeded random positions.  Output ascending.
template <class K, class Sample>
void distinct_sorted(K* out, uint64_t n, uint64_t seed, int threads, Sample&& sample) {
  std::vector<K> v;
  uint64_t next = 0;
  auto draw = [&](uint64_t count) {
    const uint64_t old = v.size();
    v.resize(old + count);
#pragma omp parallel for schedule(static) num_threads(nthreads(threads))
    for (uint64_t j = 0; j < count; ++j) v[old + j] = sample(next + j);
    next += count;
  };
  draw(n);
  for (int round = 0;; ++round) {
    psort(v.data(), v.data() + v.size(), threads);
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (v.size() >= n) break;
    if (round > 64) throw std::invalid_argument("key distribution too narrow for the requested number of distinct keys");
    // a narrow distribution yields fewer new values per draw as it fills up:
    // the draw grows geometrically with the round (bounded by 4n)
    const uint64_t need = n - v.size();
    const uint64_t count = (need + need / 4 + 64) << std::min(round, 8);
    draw(std::min<uint64_t>(count, 4 * n + 64));
  }
  if (v.size() > n) {
    const uint64_t excess = v.size() - n;
    std::vector<uint8_t> drop(v.size(), 0);
    const uint64_t ds = sub_seed(seed, 0x7711);
    uint64_t marked = 0;
    for (uint64_t t = 0; marked < excess; ++t) {
      const uint64_t p = below(rng(ds, t), v.size());
      if (!drop[p]) {
        drop[p] = 1;
        ++marked;
      }
    }
    uint64_t w = 0;
    for (uint64_t i = 0; i < v.size(); ++i)
      if (!drop[i]) out[w++] = v[i];
  } else {
    std::memcpy(out, v.data(), n * sizeof(K));
  }
}

template <class K>
void gen_build_keys(K* out, uint64_t n, gj_key_dist dist, unsigned upper_bits, uint64_t seed, int threads) {
  const unsigned kbits = 8 * sizeof(K);
  const unsigned bits = (upper_bits == 0 || upper_bits > kbits) ? kbits : upper_bits;
  const double upper = std::ldexp(1.0, static_cast<int>(bits));
  const uint64_t maxv = (bits >= 64) ? ~0ull : ((1ull << bits) - 1ull);
  switch (dist) {
    case GJ_KEYS_DENSE: {
      require(n - 1 <= maxv, "dense keys exceed the key range");
#pragma omp parallel for schedule(static) num_threads(nthreads(threads))
      for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<K>(i);
      return;
    }
    case GJ_KEYS_UNIFORM: {
      require(bits >= 64 || n <= (1ull << bits), "more keys requested than the range holds");
      const KeyedPerm perm(bits, sub_seed(seed, 1));
#pragma omp parallel for schedule(static) num_threads(nthreads(threads))
      for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<K>(perm(i));
      psort(out, out + n, threads);
      return;
    }
    case GJ_KEYS_LOGNORMAL: {
      // lognormal(mu = 0, sigma = 2), scaled so that exp(sigma * 6) maps to the top of the range
      const uint64_t s = sub_seed(seed, 2);
      const double sigma = 2.0, scale = upper / std::exp(sigma * 6.0);
      distinct_sorted(out, n, seed, threads, [&](uint64_t i) -> K {
        const double u1 = 1.0 - unit(rng(s, 2 * i));  // (0, 1]
        const double u2 = unit(rng(s, 2 * i + 1));
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        const double v = std::exp(sigma * z) * scale;
        return static_cast<K>(v >= static_cast<double>(maxv) ? maxv : static_cast<uint64_t>(v));
      });
      return;
    }
    case GJ_KEYS_CLUSTERED: {
      // 64 centres uniform in the range, Gaussian spread 2^20 around each.  Centre and
      // offset are added as integers: a double near 2^63 resolves only multiples of 2^10.
      const uint64_t sc = sub_seed(seed, 3), s = sub_seed(seed, 4);
      std::vector<uint64_t> centre(64);
      for (int c = 0; c < 64; ++c) centre[c] = maxv == ~0ull ? rng(sc, c) : below(rng(sc, c), maxv + 1);
      distinct_sorted(out, n, seed, threads, [&](uint64_t i) -> K {
        const uint64_t c = below(rng(s, 3 * i), 64);
        const double u1 = 1.0 - unit(rng(s, 3 * i + 1));
        const double u2 = unit(rng(s, 3 * i + 2));
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        const int64_t off = static_cast<int64_t>(std::floor(z * 1048576.0));  // |off| < 2^24
        const uint64_t ctr = centre[c];
        uint64_t v;
        if (off < 0)
          v = ctr < static_cast<uint64_t>(-off) ? 0 : ctr - static_cast<uint64_t>(-off);
        else
          v = maxv - ctr < static_cast<uint64_t>(off) ? maxv : ctr + static_cast<uint64_t>(off);
        return static_cast<K>(v);
      });
      return;
    }
  }
  throw std::invalid_argument("unknown key distribution");
}

// Zipf(q) on {1..n} by rejection-inversion (Hoermann & Derflinger, ACM TOMACS
// 6.3, 1996), with the numerically stable H / H^-1 forms of zipf_distribution
// (src/zipf.hpp:52-93) so that equal uniforms give equal ranks.
struct Zipf {
  double n, q, h_x1, h_n;
  static constexpr double kEps = 1e-8;
  static double expxm1bx(double x) {
    return (std::fabs(x) > kEps) ? std::expm1(x) / x : (1.0 + x / 2.0 * (1.0 + x / 3.0 * (1.0 + x / 4.0)));
  }
  static double log1pxbx(double x) {
    return (std::fabs(x) > kEps) ? std::log1p(x) / x : 1.0 - x * ((1 / 2.0) - x * ((1 / 3.0) - x * (1 / 4.0)));
  }
  double H(double x) const {
    const double lx = std::log(x);
    return expxm1bx((1.0 - q) * lx) * lx;
  }
  double H_inv(double x) const {
    const double t = std::max(-1.0, x * (1.0 - q));
    return std::exp(log1pxbx(t) * x);
  }
  double h(double x) const { return std::exp(-q * std::log(x)); }
  Zipf(uint64_t n_, double q_) : n(static_cast<double>(n_)), q(q_) {
    h_x1 = H(1.5) - 1.0;
    h_n = H(n + 0.5);
  }
  // one rejection-inversion trial with uniform u01 in [0, 1); returns 0 on rejection
  uint64_t trial(double u01) const {
    const double u = u01 * (h_n - h_x1) + h_x1;
    const double x = H_inv(u);
    double kd = std::round(x);
    if (kd < 1.0) kd = 1.0;
    if (kd > n) kd = n;
    const uint64_t k = static_cast<uint64_t>(kd);
    return (u >= H(kd + 0.5) - h(kd)) ? k : 0;
  }
};

template <class K>
uint64_t find_pos(const K* keys, uint64_t n, K x) {
  return static_cast<uint64_t>(std::lower_bound(keys, keys + n, x) - keys);
}

template <class K>
void gen_probes(const K* keys, uint64_t n, K* out, uint64_t m, gj_probe_dist dist, double theta, double miss,
                unsigned upper_bits, uint64_t seed, uint32_t* expected, int threads) {
  const unsigned kbits = 8 * sizeof(K);
  const unsigned bits = (upper_bits == 0 || upper_bits > kbits) ? kbits : upper_bits;
  const uint64_t maxv = (bits >= 64) ? ~0ull : ((1ull << bits) - 1ull);
  const uint64_t s_rank = sub_seed(seed, 11), s_miss = sub_seed(seed, 12), s_key = sub_seed(seed, 13);
  require(dist != GJ_PROBES_ZIPF || n >= 2, "zipf probes need at least two keys");
  const Zipf zipf(n >= 2 ? n - 1 : 1, theta);
  if (miss > 0.0) {
    // there must be room for absent keys
    require(n - 1 < maxv, "no absent keys exist in the key range");
  }
#pragma omp parallel for schedule(static, 4096) num_threads(nthreads(threads))
  for (uint64_t j = 0; j < m; ++j) {
    if (miss > 0.0 && unit(rng(s_miss, j)) < miss) {
      for (uint64_t t = 0;; ++t) {
        const uint64_t r = rng(s_key ^ (t * kGamma), j);
        const K x = static_cast<K>(bits >= 64 ? r : (r & maxv));
        const uint64_t p = find_pos(keys, n, x);
        if (p == n || keys[p] != x) {
          out[j] = x;
          if (expected) expected[j] = GJ_NOT_FOUND;
          break;
        }
      }
      continue;
    }
    uint64_t rank;
    if (dist == GJ_PROBES_UNIFORM) {
      rank = below(rng(s_rank, j), n);
    } else {
      // the reference draws k in [1, n-1] and probes keys[k] (src/index_lookup_common.cuh:147, 203)
      uint64_t k = 0;
      for (uint64_t t = 0; k == 0; ++t) k = zipf.trial(unit(rng(s_rank ^ (t * kGamma), j)));
      rank = k;
    }
    out[j] = keys[rank];
    if (expected) expected[j] = static_cast<uint32_t>(rank);
  }
}

}  // namespace

extern "C" {

uint64_t gj_rng_u64(uint64_t seed, uint64_t index) { return rng(seed, index); }

uint64_t gj_zipf_trial(uint64_t n, double theta, double u01) {
  if (n == 0) return 0;
  return Zipf(n, theta).trial(u01);
}

int gj_gen_build_keys(void* h_out, uint64_t n, int key_bytes, gj_key_dist dist, uint32_t upper_bits, uint64_t seed,
                      int threads) {
  return guarded([&] {
    require(h_out != nullptr, "h_out is null");
    require(key_bytes == 4 || key_bytes == 8, "key_bytes must be 4 or 8");
    if (n == 0) return;
    if (key_bytes == 4)
      gen_build_keys(static_cast<uint32_t*>(h_out), n, dist, upper_bits, seed, threads);
    else
      gen_build_keys(static_cast<uint64_t*>(h_out), n, dist, upper_bits, seed, threads);
  });
}

int gj_gen_probes(const void* h_sorted_keys, uint64_t n, int key_bytes, void* h_out, uint64_t m, gj_probe_dist dist,
                  double zipf_theta, double miss_fraction, uint32_t upper_bits, uint64_t seed, uint32_t* h_expected,
                  int threads) {
  return guarded([&] {
    require(h_sorted_keys && h_out, "null argument");
    require(n >= 1, "empty build column");
    require(key_bytes == 4 || key_bytes == 8, "key_bytes must be 4 or 8");
    require(miss_fraction >= 0.0 && miss_fraction < 1.0, "miss_fraction must be in [0, 1)");
    require(dist == GJ_PROBES_UNIFORM || dist == GJ_PROBES_ZIPF, "unknown probe distribution");
    if (key_bytes == 4)
      gen_probes(static_cast<const uint32_t*>(h_sorted_keys), n, static_cast<uint32_t*>(h_out), m, dist, zipf_theta,
                 miss_fraction, upper_bits, seed, h_expected, threads);
    else
      gen_probes(static_cast<const uint64_t*>(h_sorted_keys), n, static_cast<uint64_t*>(h_out), m, dist, zipf_theta,
                 miss_fraction, upper_bits, seed, h_expected, threads);
  });
}

int gj_gen_dimension(int32_t* h_key, uint8_t* h_flag, uint64_t dim_rows, uint64_t seed, int threads) {
  return guarded([&] {
    require(h_key && h_flag, "null argument");
    require(dim_rows < 0x7FFFFFFFull, "dimension too large for int32 keys");
    for (uint64_t i = 0; i < dim_rows; ++i) h_key[i] = static_cast<int32_t>(i + 1);
    const uint64_t sp = sub_seed(seed, 21), sf = sub_seed(seed, 22);
    for (uint64_t i = dim_rows; i > 1; --i) {  // Fisher-Yates
      const uint64_t j = below(rng(sp, i - 1), i);
      std::swap(h_key[i - 1], h_key[j]);
    }
#pragma omp parallel for schedule(static) num_threads(nthreads(threads))
    for (uint64_t i = 0; i < dim_rows; ++i) h_flag[i] = unit(rng(sf, i)) < (1.0 / 3.0) ? 1 : 0;
  });
}

int gj_gen_fact(int32_t* h_fkey, int32_t* h_date, int64_t* h_price, int8_t* h_discount, uint64_t rows,
                uint64_t dim_rows, int32_t days, int64_t price_lo, int64_t price_hi, int32_t disc_max, uint64_t seed,
                int threads) {
  return guarded([&] {
    require(h_fkey && h_date && h_price && h_discount, "null argument");
    require(dim_rows >= 1 && days >= 1 && price_hi >= price_lo && disc_max >= 0 && disc_max <= 100, "bad range");
    const uint64_t s1 = sub_seed(seed, 31), s2 = sub_seed(seed, 32), s3 = sub_seed(seed, 33),
                   s4 = sub_seed(seed, 34);
    const uint64_t pspan = static_cast<uint64_t>(price_hi - price_lo) + 1;
#pragma omp parallel for schedule(static) num_threads(nthreads(threads))
    for (uint64_t i = 0; i < rows; ++i) {
      h_fkey[i] = static_cast<int32_t>(1 + below(rng(s1, i), dim_rows));
      h_date[i] = static_cast<int32_t>(below(rng(s2, i), static_cast<uint64_t>(days)));
      h_price[i] = price_lo + static_cast<int64_t>(below(rng(s3, i), pspan));
      h_discount[i] = static_cast<int8_t>(below(rng(s4, i), static_cast<uint64_t>(disc_max) + 1));
    }
  });
}

int gj_sort_dimension(const int32_t* h_key, const uint8_t* h_flag, uint64_t rows, uint32_t* h_sorted_key,
                      uint8_t* h_sorted_flag) {
  return guarded([&] {
    require(h_key && h_flag && h_sorted_key && h_sorted_flag, "null argument");
    std::vector<uint64_t> order(rows);
    for (uint64_t i = 0; i < rows; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
      return static_cast<uint32_t>(h_key[a]) < static_cast<uint32_t>(h_key[b]);
    });
    for (uint64_t i = 0; i < rows; ++i) {
      h_sorted_key[i] = static_cast<uint32_t>(h_key[order[i]]);
      h_sorted_flag[i] = h_flag[order[i]];
    }
  });
}

int gj_sort_keys(void* h_keys, uint64_t n, int key_bytes, int threads) {
  return guarded([&] {
    require(h_keys != nullptr, "h_keys is null");
    if (key_bytes == 4)
      psort(static_cast<uint32_t*>(h_keys), static_cast<uint32_t*>(h_keys) + n, threads);
    else if (key_bytes == 8)
      psort(static_cast<uint64_t*>(h_keys), static_cast<uint64_t*>(h_keys) + n, threads);
    else
      throw std::invalid_argument("key_bytes must be 4 or 8");
  });
}

}  // extern "C"
