// index.hip - index construction and the lookup dispatch behind gj_index_build /
// gj_index_lookup.
#include <algorithm>
#include <cstdio>
#include <string>

#include "common.hpp"
#include "drivers.cuh"
#include "index_impl.hpp"
#include "lookup_dispatch.hpp"

namespace gj {

// ---------------------------------------------------------------------------
// Build-time kernels
// ---------------------------------------------------------------------------

// Sets *flag when the column is not ascending (strict: not strictly ascending).
template <class K>
__global__ void check_sorted_kernel(const K* __restrict__ col, uint64_t n, int strict, int* flag) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  bool bad = false;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i + 1 < n; i += stride) {
    const K a = col[i], b = col[i + 1];
    bad |= strict ? (a >= b) : (a > b);
  }
  if (ballot(bad) != 0 && lane() == 0) atomicOr(flag, 1);
}

// samples[j] = max key of bucket j (key_max for empty buckets).
template <class K>
__global__ void bucket_max_kernel(const K* __restrict__ col, uint64_t n, uint64_t step, uint32_t buckets,
                                  K* __restrict__ samples) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= buckets) return;
  const uint64_t lo = static_cast<uint64_t>(j) * step;
  K v = key_max<K>();
  if (lo < n) {
    const uint64_t hi = (lo + step < n) ? lo + step : n;
    v = col[hi - 1];
  }
  samples[j] = v;
}

// One B-tree level: slot t of the level is the max key of child t, whose
// subtree spans `span` column keys; key_max past the last child.
template <class K>
__global__ void btree_fill_kernel(const K* __restrict__ col, uint64_t n, K* __restrict__ level_nodes, uint64_t slots,
                                  uint64_t child_cnt, uint64_t span) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < slots; t += stride) {
    K v = key_max<K>();
    if (t < child_cnt) {
      uint64_t last = (t + 1) * span;
      if (last > n) last = n;
      v = col[last - 1];
    }
    level_nodes[t] = v;
  }
}

BTreeLayout btree_layout(uint64_t n, int key_bytes, uint32_t lds_budget_bytes, uint32_t node_bytes) {
  BTreeLayout L{};
  L.node_bytes = node_bytes;
  const uint64_t NK = node_bytes / static_cast<uint64_t>(key_bytes);
  // C_h = ceil(n / NK^(h+1)): nodes at height h (h = 0: column chunks)
  std::vector<uint64_t> cnt;
  uint64_t span = NK;
  cnt.push_back(ceil_div(n, span));
  do {
    span *= NK;
    cnt.push_back(ceil_div(n, span));
  } while (cnt.back() > 1);
  const uint32_t height = static_cast<uint32_t>(cnt.size() - 1);  // top height L >= 1
  GJ_REQUIRE(height <= static_cast<uint32_t>(kMaxLevels), GJ_ERR_INTERNAL, "B-tree too deep");
  L.levels = height;
  uint64_t off = 0;
  for (uint32_t lvl = 0; lvl < height; ++lvl) {
    L.level_off[lvl] = off;
    L.level_cnt[lvl] = cnt[height - lvl];
    off += L.level_cnt[lvl];
  }
  L.level_cnt[height] = cnt[0];
  L.total_nodes = off;
  uint64_t bytes = 0;
  for (uint32_t lvl = 0; lvl < height; ++lvl) {
    if (bytes + L.level_cnt[lvl] * node_bytes > lds_budget_bytes) break;
    bytes += L.level_cnt[lvl] * node_bytes;
    L.lds_levels = lvl + 1;
    L.lds_nodes = L.level_off[lvl] + L.level_cnt[lvl];
  }
  return L;
}

namespace {

unsigned grid_for(uint64_t work, unsigned block, unsigned cap = 65536) {
  uint64_t g = ceil_div(work, block);
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

template <class K>
void check_sorted(const gj_index& ix, bool strict, hipStream_t s) {
  if (ix.n < 2) return;
  dbuf flag(sizeof(int));
  fill_bytes(flag.p, 0, sizeof(int), s);
  hipLaunchKernelGGL(check_sorted_kernel<K>, dim3(grid_for(ix.n, 256, 8192)), dim3(256), 0, s,
                     static_cast<const K*>(ix.d_col), ix.n, strict ? 1 : 0, flag.as<int>());
  GJ_HIP_CHECK(hipGetLastError());
  int h = 0;
  GJ_HIP_CHECK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  GJ_HIP_CHECK(hipStreamSynchronize(s));
  GJ_REQUIRE(h == 0, GJ_ERR_INVALID,
             strict ? "column must be strictly ascending (unique keys) for this index kind"
                    : "column must be ascending");
}

template <class K>
void build_bsearch(gj_index& ix, hipStream_t s) {
  // bucket maxima in LDS: 8 KiB of keys
  const uint32_t max_bits = (sizeof(K) == 8) ? 10u : 11u;
  uint32_t bits = ilog2_floor(ix.n);
  bits = std::max(2u, std::min(bits, max_bits));
  const uint32_t buckets = 1u << bits;
  ix.sample_bits = bits;
  ix.sample_step = ceil_div(ix.n, buckets);
  ix.samples.alloc(sizeof(K) * buckets);
  hipLaunchKernelGGL(bucket_max_kernel<K>, dim3(grid_for(buckets, 256)), dim3(256), 0, s,
                     static_cast<const K*>(ix.d_col), ix.n, ix.sample_step, buckets, ix.samples.as<K>());
  GJ_HIP_CHECK(hipGetLastError());
  ix.stats.depth = bits + ilog2_ceil(ix.sample_step);
  ix.stats.lds_levels = bits;
}

}  // namespace

template <class K>
void build_btree(gj_index& ix, hipStream_t s) {
  // the node size is a parameter of the B-tree kind; the learned kinds' fallback tree keeps 128
  const uint32_t nb = ix.kind == GJ_BTREE ? ix.params.btree_node_bytes : 128u;
  GJ_REQUIRE(nb == 64 || nb == 128 || nb == 256, GJ_ERR_INVALID, "btree_node_bytes must be 64, 128 or 256");
  ix.bt = btree_layout(ix.n, sizeof(K), ix.params.btree_lds_bytes, nb);
  const uint64_t NK = nb / sizeof(K);
  ix.btree_nodes.alloc(ix.bt.total_nodes * nb);
  uint64_t span = NK;  // keys under one child of a height-1 node
  for (uint32_t h = 1; h <= ix.bt.levels; ++h) {
    const uint32_t lvl = ix.bt.levels - h;
    const uint64_t slots = ix.bt.level_cnt[lvl] * NK;
    const uint64_t child_cnt = ix.bt.level_cnt[lvl + 1];
    hipLaunchKernelGGL(btree_fill_kernel<K>, dim3(grid_for(slots, 256)), dim3(256), 0, s,
                       static_cast<const K*>(ix.d_col), ix.n, ix.btree_nodes.as<K>() + ix.bt.level_off[lvl] * NK,
                       slots, child_cnt, span);
    GJ_HIP_CHECK(hipGetLastError());
    span *= NK;
  }
}
template void build_btree<uint32_t>(gj_index&, hipStream_t);
template void build_btree<uint64_t>(gj_index&, hipStream_t);

void build_index(gj_index& ix, hipStream_t s) {
  dispatch_key(ix.key_bytes, [&]<class K>() {
    const bool strict = !(ix.kind == GJ_BINARY_SEARCH || ix.kind == GJ_NO_OP);
    if (ix.kind != GJ_NO_OP) check_sorted<K>(ix, strict, s);
    switch (ix.kind) {
      case GJ_BINARY_SEARCH: build_bsearch<K>(ix, s); break;
      case GJ_BTREE:
        build_btree<K>(ix, s);
        ix.stats.depth = ix.bt.levels;
        ix.stats.lds_levels = ix.bt.lds_levels;
        break;
      case GJ_NO_OP: break;
      default: build_extended<K>(ix, s); break;
    }
  });
}

}  // namespace gj
