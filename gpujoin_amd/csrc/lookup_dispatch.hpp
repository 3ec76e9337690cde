// lookup_dispatch.hpp - launch helpers and the internal interfaces between the
// libgj translation units.  Included only by *.hip files.
#pragma once

#include "common.hpp"
#include "drivers.cuh"
#include "index_impl.hpp"

namespace gj {

// Byte fill by a kernel (usable under stream capture, unlike hipMemsetAsync nodes).
void fill_bytes(void* d, int value, uint64_t bytes, hipStream_t s);

// Blocks of `block` threads that fit on one CU for kernel `fn` with `lds` bytes of dynamic LDS.
int blocks_per_cu(const void* fn, unsigned block, size_t lds);

inline unsigned lookup_block(const gj_lookup_params& p) {
  unsigned b = p.block_size ? p.block_size : 256u;
  GJ_REQUIRE(b % 64 == 0 && b >= 64 && b <= 256, GJ_ERR_INVALID, "block_size must be 64, 128, 192 or 256");
  return b;
}

template <class Dev, class K>
void launch_plain(const Dev& d, const K* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p, hipStream_t s,
                  int cu) {
  if (m == 0) return;
  const unsigned block = lookup_block(p);
  const size_t lds = d.lds_bytes();
  const void* fn = reinterpret_cast<const void*>(&lookup_plain_kernel<Dev, K>);
  uint64_t blocks = p.grid_blocks ? p.grid_blocks : static_cast<uint64_t>(blocks_per_cu(fn, block, lds)) * cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(m, block));
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL((lookup_plain_kernel<Dev, K>), dim3(static_cast<unsigned>(blocks)), dim3(block), lds, s, d,
                     probes, m, out);
  GJ_HIP_CHECK(hipGetLastError());
}

// Second pass of the plain driver for learned leaves routed to the B-tree: the
// B-tree answers every probe whose result is kFallbackMark.
template <class K>
void launch_marked(const BTreeDev<K>& d, const K* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                   hipStream_t s, int cu) {
  if (m == 0) return;
  const unsigned block = lookup_block(p);
  const size_t lds = d.lds_bytes();
  const void* fn = reinterpret_cast<const void*>(&lookup_marked_kernel<BTreeDev<K>, K>);
  uint64_t blocks = static_cast<uint64_t>(blocks_per_cu(fn, block, lds)) * cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(m, static_cast<uint64_t>(block) * kMarkSlots));
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL((lookup_marked_kernel<BTreeDev<K>, K>), dim3(static_cast<unsigned>(blocks)), dim3(block), lds, s,
                     d, probes, m, out);
  GJ_HIP_CHECK(hipGetLastError());
}

// total_max bounds the tuple count; d_total (optional) holds the exact count on the device.
template <class Dev, class K>
void launch_tuples(const Dev& d, const K* keys, const uint64_t* d_total, uint64_t total_max, uint32_t* res,
                   const gj_lookup_params& p, hipStream_t s, int cu) {
  if (total_max == 0) return;
  const unsigned block = lookup_block(p);
  const size_t lds = d.lds_bytes();
  const void* fn = reinterpret_cast<const void*>(&lookup_tuples_kernel<Dev, K>);
  uint64_t blocks = p.grid_blocks ? p.grid_blocks : static_cast<uint64_t>(blocks_per_cu(fn, block, lds)) * cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(total_max, block));
  if (blocks >= 8) blocks = blocks / 8 * 8;  // whole XCD rounds for the XCD-major chunk map
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL((lookup_tuples_kernel<Dev, K>), dim3(static_cast<unsigned>(blocks)), dim3(block), lds, s, d,
                     keys, d_total, total_max, res);
  GJ_HIP_CHECK(hipGetLastError());
}

constexpr int kBwsTile = 2048;

template <class Dev, class K>
void launch_bws(const Dev& d, const K* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p, hipStream_t s,
                int cu) {
  if (m == 0) return;
  const unsigned block = lookup_block(p);
  const size_t lds = (d.lds_bytes() + 15) / 16 * 16 + kBwsTile * (sizeof(K) + sizeof(uint32_t));
  const void* fn = reinterpret_cast<const void*>(&lookup_bws_kernel<Dev, K, kBwsTile>);
  uint64_t blocks = p.grid_blocks ? p.grid_blocks : static_cast<uint64_t>(blocks_per_cu(fn, block, lds)) * cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(m, kBwsTile));
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL((lookup_bws_kernel<Dev, K, kBwsTile>), dim3(static_cast<unsigned>(blocks)), dim3(block), lds, s,
                     d, probes, m, out);
  GJ_HIP_CHECK(hipGetLastError());
}

void build_index(gj_index& ix, hipStream_t s);
template <class K> void build_btree(gj_index& ix, hipStream_t s);
template <class K> void build_extended(gj_index& ix, hipStream_t s);

// Calls f(descriptor) with the device descriptor of the index (every kind).
template <class K, class F>
void with_index_dev(const gj_index& ix, F&& f) {
  switch (ix.kind) {
    case GJ_BINARY_SEARCH: f(ix.bsearch_dev<K>()); return;
    case GJ_BTREE:
      if (ix.bt.node_bytes == 64)
        f(ix.btree_dev<K, 64>());
      else if (ix.bt.node_bytes == 256)
        f(ix.btree_dev<K, 256>());
      else
        f(ix.btree_dev<K>());
      return;
    case GJ_NO_OP: f(NoOpDev<K>{}); return;
    case GJ_HARMONIA:
      switch (ix.params.harmonia_group_lanes) {
        case 8: f(ix.harmonia_dev<K, 8>()); return;
        case 16: f(ix.harmonia_dev<K, 16>()); return;
        case 32:
          if constexpr (sizeof(K) == 4) {
            f(ix.harmonia_dev<K, 32>());
            return;
          }
          break;
      }
      throw error(GJ_ERR_INVALID, "unsupported harmonia_group_lanes");
    case GJ_RADIX_SPLINE:
      if (ix.rs_lds)
        f(ix.rs_dev<K, true>());
      else
        f(ix.rs_dev<K, false>());
      return;
    case GJ_LEARNED_PLR:
    case GJ_LEARNED_MLP:
      if (ix.lm_layout == GJ_LAYOUT_LINES)
        f(ix.lines_dev<K>());
      else if (ix.lm_compact)
        f(ix.learned_compact_dev<K>());
      else if (ix.lm_fallback_any)
        f(ix.learned_dev<K, kLeafInline>());
      else
        f(ix.learned_dev<K, kLeafNoFallback>());
      return;
    case GJ_HASH: f(ix.hash_dev<K>()); return;
  }
  throw error(GJ_ERR_INVALID, "unknown index kind");
}

// Hash table over the column into `table` (capacity written to cap); temp: working
// memory of the partitioned build (grown as needed).
template <class K>
void build_hash_table(const gj_index& ix, dbuf& table, uint64_t& cap, dbuf& temp, hipStream_t s);
uint64_t hash_capacity(uint64_t n);

// Radix partitioner (partition.hip).  stable: input order kept inside every
// partition; otherwise the order inside a partition is arbitrary.
uint64_t partition_capacity(uint64_t m, int key_bytes, uint32_t bits);
uint64_t partition_scratch(uint64_t m, int key_bytes, uint32_t bits, uint32_t passes);
// A batch of m probes cut into windows of `window`, each window partitioned on its own
// (output partition w * 2^bits + digit).  No payloads: inv1[i] is where probe i went;
// when windows_two_pass(), that is a position in the first pass's output, and
// inv2[inv1[i]] its final position (inv2 is null otherwise).  inv1 and inv2 hold m
// entries each.
uint64_t windows_capacity(uint64_t m, uint64_t window, int key_bytes, uint32_t bits);
uint64_t windows_scratch(uint64_t m, uint64_t window, int key_bytes, uint32_t bits, uint32_t passes);
bool windows_two_pass(uint32_t bits, uint32_t passes);
template <class K>
void partition_windows(const K* keys, uint64_t m, uint64_t window, uint32_t bits, uint32_t ignore, uint32_t passes,
                       K* out_keys, uint32_t* inv1, uint32_t* inv2, uint64_t* start, uint64_t* count,
                       uint8_t* scratch, unsigned long long* dropped, hipStream_t s);
template <class K>
void partition_keys(const K* keys, uint64_t m, uint32_t payload_base, uint32_t bits, uint32_t ignore, uint32_t passes,
                    bool stable, K* out_keys, uint32_t* out_payload, uint64_t* start, uint64_t* count, uint8_t* scratch,
                    unsigned long long* dropped, int cu, hipStream_t s);

// Bytes of index-private device memory held by the extended kinds.
uint64_t extended_memory(const gj_index& ix);

// Drivers (drivers.hip)
void lookup_dispatch(const gj_index& ix, const void* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                     hipStream_t s);

}  // namespace gj
