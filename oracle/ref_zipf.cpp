// ref_zipf.cpp - extern "C" wrapper that compiles the reference's own Zipf
// sampler (src/zipf.hpp, included from the reference tree at build time) into
// oracle/_ref/libref_zipf.so.  Test infrastructure: it pins the oracle's
// rejection-inversion restatement against the reference bit for bit.
#include <cstdint>
#include <random>

#include "zipf.hpp"

extern "C" void ref_zipf_mt19937(uint64_t n, double q, uint32_t seed, uint64_t count, uint64_t* out) {
  std::mt19937 gen(seed);
  zipf_distribution<uint64_t> dist(n, q);
  for (uint64_t i = 0; i < count; ++i) out[i] = dist(gen);
}
