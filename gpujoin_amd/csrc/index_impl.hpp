// index_impl.hpp - the gj_index object behind the opaque C handle.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "indexes.cuh"
#include "indexes_ext.cuh"

namespace gj {

struct BTreeLayout {
  uint32_t node_bytes = 128;  // bytes per node and per leaf chunk (64, 128 or 256)
  uint32_t levels = 0;      // inner levels
  uint32_t lds_levels = 0;  // top levels staged in LDS
  uint64_t lds_nodes = 0;
  uint64_t total_nodes = 0;
  uint64_t level_off[kMaxLevels] = {};
  uint64_t level_cnt[kMaxLevels + 1] = {};
};

// Host-side layout of the implicit B-tree over n keys with node_bytes / key_bytes keys per node.
BTreeLayout btree_layout(uint64_t n, int key_bytes, uint32_t lds_budget_bytes, uint32_t node_bytes = 128);

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
  bool col_in_host = false;  // the column is page-locked host memory (read over the host link)

  // binary search: bucket maxima staged in LDS
  gj::dbuf samples;
  uint32_t sample_bits = 0;
  uint64_t sample_step = 0;

  // B-tree (also the fallback of the learned indexes)
  gj::dbuf btree_nodes;
  gj::BTreeLayout bt;

  // Harmonia
  gj::dbuf hm_keys;
  uint64_t hm_inner_nodes = 0, hm_lds_nodes = 0, hm_nodes = 0;
  uint64_t hm_first[gj::kMaxLevels] = {};  // first node of each level
  uint32_t hm_depth = 0, hm_lds_levels = 0;

  // RadixSpline
  gj::dbuf rs_radix, rs_knots;
  uint32_t rs_shift = 0, rs_err = 0;
  bool rs_lds = false;  // the radix table is staged in LDS
  uint64_t rs_points = 0;

  // learned two-stage models (B-tree above is their fallback)
  gj::dbuf lm_leaves, lm_mlp, lm_knots, lm_seps, lm_subs;
  gj::dbuf lm_groups;        // compact leaves: group bases (LeafGroup)
  bool lm_compact = false;   // lm_leaves holds LeafCompact (8 B) instead of LeafModel (16 B)
  uint32_t lm_gshift = 0;    // compact leaves per group = 2^lm_gshift
  uint32_t lm_splits = 0;  // learned leaves split into sub-leaves
  gj::RootModel lm_root{};
  uint32_t lm_num_leaves = 0, lm_fallback_any = 0;
  int32_t lm_err_lo = 0, lm_err_hi = 0;  // learned window bounds over all leaves
  // learned lines layout: the key lines and the two-stage leaves (none: root only)
  uint32_t lm_layout = GJ_LAYOUT_COLUMN;
  gj::dbuf lm_lines, lm_line_leaves;
  uint64_t lm_num_lines = 0, lm_mult = 0;
  uint32_t lm_num_line_leaves = 0;

  // hash table
  gj::dbuf ht;
  uint64_t ht_cap = 0;

  // Device scratch of the calls that need memory beyond their inputs and outputs
  // (the partitioned driver's tuples, the hash-join driver's table, the
  // join-aggregate's partial sums): one set per stream, grown on demand and kept
  // until the index is destroyed, so calls on different streams never share it.
  struct StreamScratch {
    gj::dbuf a, b;
  };
  mutable std::mutex scratch_mu;
  mutable std::map<hipStream_t, std::unique_ptr<StreamScratch>> scratch_by_stream;
  StreamScratch& scratch_for(hipStream_t s) const {
    std::lock_guard<std::mutex> g(scratch_mu);
    auto& p = scratch_by_stream[s];
    if (!p) p = std::make_unique<StreamScratch>();
    return *p;
  }
  // u64 device counter of tuples the partitioned driver's scatter could not place
  gj::dbuf dropped;

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

  template <class K, int G>
  gj::HarmoniaDev<K, G> harmonia_dev() const {
    gj::HarmoniaDev<K, G> d{};
    d.col = static_cast<const K*>(d_col);
    d.n = n;
    d.keys = hm_keys.as<const K>();
    d.inner_nodes = hm_inner_nodes;
    for (int i = 0; i < gj::kMaxLevels; ++i) d.first[i] = hm_first[i];
    d.depth = hm_depth;
    d.lds_levels = hm_lds_levels;
    d.lds_nodes = hm_lds_nodes;
    return d;
  }

  template <class K, bool Lds>
  gj::RadixSplineDev<K, Lds> rs_dev() const {
    gj::RadixSplineDev<K, Lds> d{};
    d.col = static_cast<const K*>(d_col);
    d.n = n;
    d.radix = rs_radix.as<const uint32_t>();
    d.radix_count = static_cast<uint32_t>(rs_radix.bytes / sizeof(uint32_t));
    d.knots = rs_knots.as<const gj::RsKnot>();
    d.min_key = static_cast<K>(min_key);
    d.max_key = static_cast<K>(max_key);
    d.shift_bits = rs_shift;
    d.err = rs_err;
    return d;
  }

  template <class K, int Mode>
  gj::LearnedDev<K, Mode> learned_dev() const {
    gj::LearnedDev<K, Mode> d{};
    d.col = static_cast<const K*>(d_col);
    d.n = n;
    d.min_key = static_cast<K>(min_key);
    d.max_key = static_cast<K>(max_key);
    d.root = lm_root;
    d.root.mlp = lm_mlp.as<const float>();
    d.root.knots = lm_knots.as<const uint64_t>();
    d.leaves = lm_leaves.as<const gj::LeafModel>();
    d.num_leaves = lm_num_leaves;
    d.err_lo = lm_err_lo;
    d.err_hi = lm_err_hi;
    d.num_splits = lm_splits;
    d.seps = lm_seps.as<const uint64_t>();
    d.subs = lm_subs.as<const gj::LeafModel>();
    d.btree = btree_dev<K>();
    return d;
  }

  template <class K>
  gj::LearnedCompactDev<K> learned_compact_dev() const {
    gj::LearnedCompactDev<K> d{};
    d.col = static_cast<const K*>(d_col);
    d.n = n;
    d.min_key = static_cast<K>(min_key);
    d.max_key = static_cast<K>(max_key);
    d.root = lm_root;
    d.root.mlp = lm_mlp.as<const float>();
    d.root.knots = lm_knots.as<const uint64_t>();
    d.leaves = lm_leaves.as<const gj::LeafCompact>();
    d.groups = lm_groups.as<const gj::LeafGroup>();
    d.num_leaves = lm_num_leaves;
    d.gshift = lm_gshift;
    d.num_groups = static_cast<uint32_t>(lm_groups.bytes / sizeof(gj::LeafGroup));
    d.err_lo = lm_err_lo;
    d.err_hi = lm_err_hi;
    return d;
  }

  template <class K>
  gj::LinesDev<K> lines_dev() const {
    gj::LinesDev<K> d{};
    d.lines = lm_lines.as<const K>();
    d.num_lines = lm_num_lines;
    d.min_key = static_cast<K>(min_key);
    d.max_key = static_cast<K>(max_key);
    d.mult = lm_mult;
    d.root = lm_root;
    d.root.mlp = lm_mlp.as<const float>();
    d.root.knots = lm_knots.as<const uint64_t>();
    d.leaves = lm_line_leaves.as<const gj::LineLeaf>();
    d.num_leaves = lm_num_line_leaves;
    return d;
  }

  template <class K>
  gj::HashDev<K> hash_dev() const {
    gj::HashDev<K> d{};
    d.table = ht.as<const gj::HashEntry>();
    d.mask = ht_cap - 1;
    return d;
  }

  // NB: the node size the tree was built with (bt.node_bytes); the learned kinds'
  // fallback tree is always built with 128
  template <class K, int NB = 128>
  gj::BTreeDev<K, NB> btree_dev() const {
    gj::BTreeDev<K, NB> d{};
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
