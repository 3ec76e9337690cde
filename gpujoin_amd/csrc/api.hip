// api.hip - the extern "C" index entry points of gj.h.
#include <cstring>
#include <memory>
#include <string>

#include "common.hpp"
#include "index_impl.hpp"
#include "lookup_dispatch.hpp"

using namespace gj;

namespace {

void fill_defaults(gj_index_params& p) {
  if (!p.rs_radix_bits) p.rs_radix_bits = 18;
  if (!p.rs_max_error) p.rs_max_error = 32;
  if (!p.rs_lds_bits) p.rs_lds_bits = 13;
  if (!p.harmonia_group_lanes) p.harmonia_group_lanes = 8;  // one 128-byte line per load
  if (!p.btree_lds_bytes) p.btree_lds_bytes = 16384;
  if (!p.btree_node_bytes) p.btree_node_bytes = 128;
  // learned_leaf_bits 0: chosen from the column size at build
  if (!p.learned_max_error) p.learned_max_error = 128;
  if (!p.mlp_hidden) p.mlp_hidden = 32;
  if (!p.mlp_steps) p.mlp_steps = 400;
  if (!p.mlp_batch) p.mlp_batch = 65536;
  if (!(p.mlp_lr > 0.0f)) p.mlp_lr = 3e-3f;
  if (!p.seed) p.seed = 0x5eedull;
}

uint64_t memory_of(const gj_index& ix) {
  uint64_t b = ix.samples.bytes + ix.btree_nodes.bytes;
  b += extended_memory(ix);
  return b;
}

}  // namespace

extern "C" {

void gj_index_params_default(gj_index_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  fill_defaults(*p);
}

void gj_lookup_params_default(gj_lookup_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->approach = GJ_PLAIN;
  p->block_size = 256;
  p->radix_bits = 10;
  p->passes = 1;
  p->ignore_bits = 0xFFFFFFFFu;
  p->window = 1ull << 24;
}

int gj_parse_index_type(const char* name, gj_index_kind* out) {
  return guarded([&] {
    GJ_REQUIRE(name && out, GJ_ERR_INVALID, "null argument");
    const std::string s(name);
    if (s == "binary_search") *out = GJ_BINARY_SEARCH;
    else if (s == "btree") *out = GJ_BTREE;
    else if (s == "harmonia") *out = GJ_HARMONIA;
    else if (s == "radix_spline") *out = GJ_RADIX_SPLINE;
    else if (s == "learned_plr") *out = GJ_LEARNED_PLR;
    else if (s == "learned_mlp") *out = GJ_LEARNED_MLP;
    else if (s == "no_op") *out = GJ_NO_OP;
    else if (s == "hash") *out = GJ_HASH;
    else throw error(GJ_ERR_INVALID, "unknown index type");  // src/indexes.cpp:17
  });
}

const char* gj_index_type_name(gj_index_kind kind) {
  switch (kind) {
    case GJ_BINARY_SEARCH: return "binary_search";
    case GJ_BTREE: return "btree";
    case GJ_HARMONIA: return "harmonia";
    case GJ_RADIX_SPLINE: return "radix_spline";
    case GJ_LEARNED_PLR: return "learned_plr";
    case GJ_LEARNED_MLP: return "learned_mlp";
    case GJ_NO_OP: return "no_op";
    case GJ_HASH: return "hash";
  }
  return "unknown";
}

int gj_index_build(gj_index_kind kind, const void* h_sorted_keys, const void* d_sorted_keys, uint64_t n,
                   int key_bytes, const gj_index_params* params, gj_stream_t stream, gj_index** out) {
  return guarded([&] {
    GJ_REQUIRE(out, GJ_ERR_INVALID, "out is null");
    *out = nullptr;
    GJ_REQUIRE(kind >= GJ_BINARY_SEARCH && kind <= GJ_HASH, GJ_ERR_INVALID, "unknown index kind");
    GJ_REQUIRE(key_bytes == 4 || key_bytes == 8, GJ_ERR_INVALID, "key_bytes must be 4 or 8");
    GJ_REQUIRE(n >= 1, GJ_ERR_INVALID, "the column must hold at least one key");
    // row ids are uint32 with 0xFFFFFFFF reserved for not-found
    GJ_REQUIRE(n < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "'value_t' does not suffice: n must be < 2^32 - 1");
    GJ_REQUIRE(kind == GJ_NO_OP || d_sorted_keys, GJ_ERR_INVALID, "d_sorted_keys is null");
    GJ_REQUIRE((reinterpret_cast<uintptr_t>(d_sorted_keys) & 15u) == 0, GJ_ERR_INVALID,
               "d_sorted_keys must be 16-byte aligned");
    auto ix = std::make_unique<gj_index>();
    ix->kind = kind;
    ix->key_bytes = key_bytes;
    ix->n = n;
    ix->d_col = d_sorted_keys;
    ix->h_col = h_sorted_keys;
    if (params) ix->params = *params;
    fill_defaults(ix->params);
    ix->cu = device_cu_count();
    if (d_sorted_keys) {
      hipPointerAttribute_t attr{};
      if (hipPointerGetAttributes(&attr, d_sorted_keys) == hipSuccess)
        ix->col_in_host = attr.type == hipMemoryTypeHost;
      else
        (void)hipGetLastError();  // memory HIP does not know: treated as device memory
    }
    const hipStream_t s = S(stream);
    if (d_sorted_keys) {
      ix->min_key = 0;
      ix->max_key = 0;
      GJ_HIP_CHECK(hipMemcpyAsync(&ix->min_key, d_sorted_keys, key_bytes, hipMemcpyDeviceToHost, s));
      GJ_HIP_CHECK(hipMemcpyAsync(&ix->max_key, static_cast<const uint8_t*>(d_sorted_keys) + (n - 1) * key_bytes,
                                  key_bytes, hipMemcpyDeviceToHost, s));
      GJ_HIP_CHECK(hipStreamSynchronize(s));
    }
    GJ_REQUIRE(ix->params.placement <= GJ_PLACE_HOST, GJ_ERR_INVALID, "placement must be 0 or 1");
    // the partitioner's dropped-tuple counter (allocated before the build, which may partition)
    ix->dropped.alloc(16);
    fill_bytes(ix->dropped.p, 0, ix->dropped.bytes, s);
    build_index(*ix, s);
    if (ix->params.placement == GJ_PLACE_HOST) {
      GJ_HIP_CHECK(hipStreamSynchronize(s));
      for (dbuf* b : {&ix->samples, &ix->btree_nodes, &ix->hm_keys, &ix->rs_radix, &ix->rs_knots,
                      &ix->lm_leaves, &ix->lm_mlp, &ix->lm_knots, &ix->lm_seps, &ix->lm_subs, &ix->lm_groups,
                      &ix->lm_lines, &ix->lm_line_leaves, &ix->ht})
        b->to_host();
      ix->stats.placement = GJ_PLACE_HOST;
    }
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    ix->h_col = nullptr;  // the host view is only valid during construct
    ix->stats.kind = kind;
    ix->stats.key_bytes = key_bytes;
    ix->stats.n = n;
    ix->stats.memory_bytes = memory_of(*ix);
    *out = ix.release();
  });
}

int gj_index_lookup(const gj_index* index, const void* d_probe_keys, uint64_t m, uint32_t* d_rowids,
                    const gj_lookup_params* params, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(index, GJ_ERR_INVALID, "index is null");
    if (m == 0) return;
    GJ_REQUIRE(d_probe_keys && d_rowids, GJ_ERR_INVALID, "null probe or row id buffer");
    gj_lookup_params p;
    gj_lookup_params_default(&p);
    if (params) p = *params;
    lookup_dispatch(*index, d_probe_keys, m, d_rowids, p, S(stream));
  });
}

uint64_t gj_index_memory(const gj_index* index) { return index ? memory_of(*index) : 0; }

int gj_index_get_stats(const gj_index* index, gj_index_stats* out) {
  return guarded([&] {
    GJ_REQUIRE(index && out, GJ_ERR_INVALID, "null argument");
    *out = index->stats;
    out->memory_bytes = memory_of(*index);
    out->dropped_tuples = 0;
    if (index->dropped.p) {
      // the counter is written by kernels on any stream: wait for the device, then read it
      GJ_HIP_CHECK(hipDeviceSynchronize());
      GJ_HIP_CHECK(hipMemcpy(&out->dropped_tuples, index->dropped.p, sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
  });
}

void gj_index_destroy(gj_index* index) { delete index; }

}  // extern "C"
