// index_impl.hpp - the gj_index object behind the opaque C handle.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "indexes.cuh"

namespace gj {

struct BTreeLayout {
  uint32_t levels = 0;      // inner levels
  uint32_t lds_levels = 0;  // top levels staged in LDS
  uint64_t lds_nodes = 0;
  uint64_t total_nodes = 0;
  uint64_t level_off[kMaxLevels] = {};
  uint64_t level_cnt[kMaxLevels + 1] = {};
};

// Host-side layout of the implicit B-tree over n keys with NK keys per node.
BTreeLayout btree_layout(uint64_t n, int key_bytes, uint32_t lds_budget_bytes);

}  // namespace gj

struct gj_index {
  gj_index_kind kind = GJ_NO_OP;
  int key_bytes = 8;
  uint64_t n = 0;
  const void* d_col = nullptr;
  const void* h_col = nullptr;  // optional host view of the column, valid during build only
  gj_index_params params{};
  uint64_t min_key = 0, max_key = 0;
  int cu = 0;

  // binary search: bucket maxima staged in LDS
  gj::dbuf samples;
  uint32_t sample_bits = 0;
  uint64_t sample_step = 0;

  // B-tree (also the fallback of the learned indexes)
  gj::dbuf btree_nodes;
  gj::BTreeLayout bt;

  // scratch reused by the bws / partitioning / hj drivers
  mutable gj::dbuf scratch;

  gj_index_stats stats{};

  template <class K>
  gj::BSearchDev<K> bsearch_dev() const {
    gj::BSearchDev<K> d{};
    d.col = static_cast<const K*>(d_col);
    d.n = n;
    d.samples = samples.as<const K>();
    d.step = sample_step;
    d.sample_bits = sample_bits;
    return d;
  }

  template <class K>
  gj::BTreeDev<K> btree_dev() const {
    gj::BTreeDev<K> d{};
    d.col = static_cast<const K*>(d_col);
    d.n = n;
    d.nodes = btree_nodes.as<const K>();
    d.levels = bt.levels;
    d.lds_levels = bt.lds_levels;
    d.lds_nodes = bt.lds_nodes;
    for (int i = 0; i < gj::kMaxLevels; ++i) d.level_off[i] = bt.level_off[i];
    for (int i = 0; i <= gj::kMaxLevels; ++i) d.level_cnt[i] = bt.level_cnt[i];
    return d;
  }
};
