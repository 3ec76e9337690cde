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
    // fan-outs above 2^12 always take two passes, and pay only with large windows (binary
    // search over 2^26 keys: 14.4 G lookups/s at 2^16 partitions, 12.4 at 2^10)
    for (uint32_t bits : {8u, 10u, 12u, 14u, 16u}) {
      for (uint64_t window : {1ull << 20, 1ull << 22, 1ull << 24, 1ull << 26}) {
        if (window > 4 * m && window != (1ull << 20)) continue;
        if (bits > 12 && window < (1ull << 24)) continue;
        for (uint32_t passes : {1u, 2u}) {
          if (passes == 2 && bits < 10) continue;
          if (passes == 1 && bits > 12) continue;
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

// ---- index autotuner -------------------------------------------------------
// Picks the index's own parameters: threads per lookup (Harmonia's lanes per
// node search), the B-tree's node size and LDS-resident levels, RadixSpline's radix
// bits and error bound, the learned index's layout, line load, leaf count and error
// bound.  Each
// candidate is built over the caller's column and timed with the plain driver
// on the probe sample; the candidates are destroyed again.
namespace {

std::vector<gj_index_params> index_candidates(gj_index_kind kind, uint64_t n, int key_bytes,
                                              const gj_index_params& base) {
  std::vector<gj_index_params> c;
  auto add = [&](auto set) {
    gj_index_params p = base;
    set(p);
    c.push_back(p);
  };
  switch (kind) {
    case GJ_HARMONIA:
      for (uint32_t g : {8u, 16u, 32u})
        if (g != 32 || key_bytes == 4) add([&](gj_index_params& p) { p.harmonia_group_lanes = g; });
      break;
    case GJ_BTREE:
      // node size (64 B, one 128 B line, 256 B) x the LDS budget of the upper levels
      for (uint32_t nb : {64u, 128u, 256u})
        for (uint32_t b : {4096u, 16384u, 32768u})
          add([&](gj_index_params& p) {
            p.btree_node_bytes = nb;
            p.btree_lds_bytes = b;
          });
      break;
    case GJ_RADIX_SPLINE:
      // radix bits whose table LDS holds (12, 13) and the larger global tables (16, 18) x the
      // error bound
      for (uint32_t r : {12u, 13u, 16u, 18u})
        for (uint32_t e : {16u, 32u, 64u})
          add([&](gj_index_params& p) {
            p.rs_radix_bits = r;
            p.rs_max_error = e;
          });
      break;
    case GJ_LEARNED_PLR:
    case GJ_LEARNED_MLP: {
      // the default (compact leaves where the build admits them), then explicit counts
      // of 16-byte leaves around the default
      add([&](gj_index_params& p) { p.learned_leaf_bits = 0; });
      // the layouts: the column itself, and lines filled to 8 or 12 keys (the default 10
      // is the candidate above where the auto layout picks lines)
      add([&](gj_index_params& p) { p.learned_layout = GJ_LAYOUT_COLUMN; });
      for (uint32_t lk : {8u, 12u})
        add([&](gj_index_params& p) {
          p.learned_layout = GJ_LAYOUT_LINES;
          p.learned_line_keys = key_bytes == 4 ? 2 * lk : lk;
        });
      const int a = std::clamp(static_cast<int>(ilog2_ceil(n)) - 10, 16, 24);  // the build's default leaf bits
      for (int d : {-1, 0, 1})
        for (uint32_t e : {64u, 128u, 256u}) {
          if (kind == GJ_LEARNED_MLP && e != 128u) continue;  // each MLP candidate retrains the root
          add([&](gj_index_params& p) {
            p.learned_leaf_bits = static_cast<uint32_t>(std::clamp(a + d, 1, 24));
            p.learned_max_error = e;
          });
        }
      break;
    }
    default:
      c.push_back(base);
  }
  return c;
}

}  // namespace

extern "C" int gj_autotune_index(gj_index_kind kind, const void* h_sorted_keys, const void* d_sorted_keys, uint64_t n,
                                 int key_bytes, const gj_index_params* base, const void* d_probe_sample, uint64_t m,
                                 uint32_t* d_rowids_scratch, gj_index_params* best, double* best_lookups_per_s,
                                 gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_probe_sample && d_rowids_scratch && best, GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(m > 0, GJ_ERR_INVALID, "empty probe sample");
    gj_index_params b;
    gj_index_params_default(&b);
    if (base) b = *base;
    const hipStream_t s = S(stream);
    gj_lookup_params lp;
    gj_lookup_params_default(&lp);
    hipEvent_t e0, e1;
    GJ_HIP_CHECK(hipEventCreate(&e0));
    GJ_HIP_CHECK(hipEventCreate(&e1));
    double best_ms = -1.0;
    gj_index* ix = nullptr;
    try {
      for (const gj_index_params& p : index_candidates(kind, n, key_bytes, b)) {
        ix = nullptr;
        const int st = gj_index_build(kind, h_sorted_keys, d_sorted_keys, n, key_bytes, &p, stream, &ix);
        if (st == GJ_ERR_HIP || st == GJ_ERR_NOMEM) throw error(st, gj_last_error());
        if (st != GJ_OK) continue;  // a candidate this column does not admit
        lookup_dispatch(*ix, d_probe_sample, m, d_rowids_scratch, lp, s);  // warm-up
        GJ_HIP_CHECK(hipEventRecord(e0, s));
        const int reps = 3;
        for (int r = 0; r < reps; ++r) lookup_dispatch(*ix, d_probe_sample, m, d_rowids_scratch, lp, s);
        GJ_HIP_CHECK(hipEventRecord(e1, s));
        GJ_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        GJ_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        gj_index_destroy(ix);
        ix = nullptr;
        const double per = ms / reps;
        if (best_ms < 0 || per < best_ms) {
          best_ms = per;
          *best = p;
        }
      }
    } catch (...) {
      if (ix) gj_index_destroy(ix);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    GJ_REQUIRE(best_ms > 0, GJ_ERR_INVALID, "no candidate index could be built over this column");
    if (best_lookups_per_s) *best_lookups_per_s = static_cast<double>(m) / (best_ms * 1e-3);
  });
}
