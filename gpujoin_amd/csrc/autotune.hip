// autotune.hip - picks the lookup driver configuration (approach, block size,
// partition fan-out, passes, window) by timing candidates on a probe sample.
// No reference counterpart: the reference fixes block size 128, fan-out 2^11
// and the window on the command line (src/index_lookup_config.hpp:18-37).
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "lookup_dispatch.hpp"

using namespace gj;

extern "C" int gj_autotune_lookup(const gj_index* index, const void* d_probe_sample, uint64_t m,
                                  uint32_t* d_rowids_scratch, gj_lookup_params* best, double* best_lookups_per_s,
                                  gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(index && d_probe_sample && d_rowids_scratch && best, GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(m > 0, GJ_ERR_INVALID, "empty probe sample");
    const hipStream_t s = S(stream);
    std::vector<gj_lookup_params> cand;
    auto base = [] {
      gj_lookup_params p;
      gj_lookup_params_default(&p);
      return p;
    };
    for (uint32_t block : {128u, 256u}) {
      gj_lookup_params p = base();
      p.approach = GJ_PLAIN;
      p.block_size = block;
      cand.push_back(p);
    }
    {
      gj_lookup_params p = base();
      p.approach = GJ_BWS;
      cand.push_back(p);
    }
    for (uint32_t bits : {8u, 10u, 12u}) {
      for (uint64_t window : {1ull << 20, 1ull << 22, 1ull << 24, 1ull << 26}) {
        if (window > 4 * m && window != (1ull << 20)) continue;
        for (uint32_t passes : {1u, 2u}) {
          if (passes == 2 && bits < 10) continue;
          gj_lookup_params p = base();
          p.approach = GJ_PARTITIONING;
          p.radix_bits = bits;
          p.window = window;
          p.passes = passes;
          cand.push_back(p);
        }
      }
    }
    hipEvent_t e0, e1;
    GJ_HIP_CHECK(hipEventCreate(&e0));
    GJ_HIP_CHECK(hipEventCreate(&e1));
    double best_ms = -1.0;
    try {
      for (const auto& p : cand) {
        try {
          lookup_dispatch(*index, d_probe_sample, m, d_rowids_scratch, p, s);  // warm-up (and scratch growth)
          GJ_HIP_CHECK(hipEventRecord(e0, s));
          const int reps = 3;
          for (int r = 0; r < reps; ++r) lookup_dispatch(*index, d_probe_sample, m, d_rowids_scratch, p, s);
          GJ_HIP_CHECK(hipEventRecord(e1, s));
          GJ_HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0;
          GJ_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          const double per = ms / reps;
          if (best_ms < 0 || per < best_ms) {
            best_ms = per;
            *best = p;
          }
        } catch (const error& e) {
          if (e.status == GJ_ERR_HIP || e.status == GJ_ERR_NOMEM) throw;  // real failures propagate
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    GJ_REQUIRE(best_ms > 0, GJ_ERR_INTERNAL, "no lookup configuration ran");
    if (best_lookups_per_s) *best_lookups_per_s = static_cast<double>(m) / (best_ms * 1e-3);
  });
}
