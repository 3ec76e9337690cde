// extended.hip - construction and lookup dispatch of Harmonia, RadixSpline,
// the learned models and the hash table.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "lookup_dispatch.hpp"
#include "rs_builder.hpp"

namespace gj {

template <class K> void build_learned(gj_index& ix, hipStream_t s);  // learned.hip

namespace {

unsigned grid_for(uint64_t work, unsigned block, unsigned cap = 65536) {
  uint64_t g = ceil_div(work, block);
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// ---------------------------------------------------------------------------
// Harmonia construction.  The level sizes follow gather_tree_info
// (src/harmonia.cuh:133-185).  Every node of a level except the last is full
// (32 keys, 33 children), so children and separators have closed forms:
//   children[node i of inner level l] = first(l+1) + 33 i     (store_structure, :324-344)
//   children[leaf node i]             = 32 i                  (populate_leaf_nodes, :297-315)
//   key s of node i of level l        = largest key under child 33 i + s of level l+1
//                                     = col[min(n, (33 i + s + 1) * span(l+1)) - 1]
// with span(l) = 32 * 33^(levels below l) column keys per node
// (get_largest_key / populate_inner_nodes, :245-295).  Only the separators are
// stored; the lookup computes the children from the level starts.
// ---------------------------------------------------------------------------
struct HmLevels {
  uint32_t depth;
  uint64_t first[kMaxLevels];
  uint64_t count[kMaxLevels];
  uint64_t keycnt[kMaxLevels];
  uint64_t span[kMaxLevels];
};

__device__ __forceinline__ uint32_t hm_level_of(const HmLevels& L, uint64_t g) {
  uint32_t l = 0;
  while (l + 1 < L.depth && L.first[l + 1] <= g) ++l;
  return l;
}

template <class K>
__global__ void hm_keys_kernel(const K* __restrict__ col, uint64_t n, HmLevels L, K* __restrict__ keys, uint64_t slots) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < slots; t += stride) {
    const uint64_t g = t / 32, s = t % 32;
    const uint32_t l = hm_level_of(L, g);
    const uint64_t i = g - L.first[l];
    const uint64_t nkc = (i + 1 < L.count[l]) ? 32ull : L.keycnt[l] - i * 32ull;
    K v = key_max<K>();
    if (s < nkc) {
      const uint64_t c = 33ull * i + s;
      const uint64_t span = L.span[l + 1];
      const uint64_t lim = (c + 1 > n / span) ? n : (c + 1) * span;
      v = col[lim - 1];
    }
    keys[t] = v;
  }
}

template <class K>
void build_harmonia(gj_index& ix, hipStream_t s) {
  GJ_REQUIRE(ix.n % 32 == 0, GJ_ERR_INVALID,
             "harmonia: when Reuse_Input_Memory is set the dataset set must be a multiple of the node size (32)");
  // level sizes, bottom-up (gather_tree_info)
  std::vector<uint64_t> kc, pc;
  uint64_t prev = 0, k = ix.n;
  for (;;) {
    uint64_t p = (k + 31) / 32;
    if (33 * p < prev) p += 1;
    kc.push_back(k);
    pc.push_back(p);
    prev = p;
    k = (32 * p) / 33;
    if (p == 1) break;
    GJ_REQUIRE(kc.size() < static_cast<size_t>(kMaxLevels), GJ_ERR_INTERNAL, "harmonia: max depth exceeded");
  }
  HmLevels L{};
  L.depth = static_cast<uint32_t>(kc.size());
  uint64_t first = 0;
  for (uint32_t l = 0; l < L.depth; ++l) {
    const size_t b = L.depth - 1 - l;  // bottom-up index
    L.first[l] = first;
    L.count[l] = pc[b];
    L.keycnt[l] = kc[b];
    first += pc[b];
  }
  uint64_t span = 32;
  for (int l = static_cast<int>(L.depth) - 1; l >= 0; --l) {
    L.span[l] = span;
    span = (span > (~0ull) / 33) ? ~0ull : span * 33;
  }
  ix.hm_depth = L.depth;
  ix.hm_inner_nodes = L.first[L.depth - 1];
  ix.hm_nodes = first;
  for (uint32_t l = 0; l < L.depth; ++l) ix.hm_first[l] = L.first[l];
  GJ_REQUIRE(ix.hm_nodes < 0xFFFFFFFFull && ix.n <= 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "harmonia: too many nodes");
  if (ix.hm_inner_nodes) {
    const uint64_t slots = ix.hm_inner_nodes * 32;
    ix.hm_keys.alloc(slots * sizeof(K));
    hipLaunchKernelGGL(hm_keys_kernel<K>, dim3(grid_for(slots, 256)), dim3(256), 0, s, static_cast<const K*>(ix.d_col),
                       ix.n, L, ix.hm_keys.as<K>(), slots);
    GJ_HIP_CHECK(hipGetLastError());
  }
  // top inner levels into LDS within the budget
  const uint64_t budget = ix.params.btree_lds_bytes;
  uint64_t bytes = 0;
  for (uint32_t l = 0; l + 1 < L.depth; ++l) {
    const uint64_t add = L.count[l] * 32 * sizeof(K);
    if (bytes + add > budget) break;
    bytes += add;
    ix.hm_lds_levels = l + 1;
    ix.hm_lds_nodes = L.first[l] + L.count[l];
  }
  ix.stats.depth = L.depth;
  ix.stats.lds_levels = ix.hm_lds_levels;
  const uint32_t g = ix.params.harmonia_group_lanes;
  const bool ok = (g == 8 || g == 16) || (sizeof(K) == 4 && g == 32);
  GJ_REQUIRE(ok, GJ_ERR_INVALID, "harmonia_group_lanes must be 8 or 16 (or 32 for 4-byte keys)");
}

// ---------------------------------------------------------------------------
// RadixSpline: host builder, device tables, then a device pass that checks
// every build key lies inside its own search window and widens the bound
// until it does (the FP64 interpolation can round across the exact bound).
// ---------------------------------------------------------------------------
template <class K>
__global__ void rs_check_kernel(RadixSplineDev<K, false> d, unsigned long long* bad) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  unsigned long long cnt = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < d.n; i += stride) {
    uint64_t lo, hi;
    d.window(d.estimate(d.col[i], d.radix), lo, hi);
    cnt += (i < lo || i >= hi) ? 1ull : 0ull;
  }
  if (cnt) atomicAdd(bad, cnt);
}

std::vector<uint8_t> host_column(const gj_index& ix, hipStream_t s) {
  std::vector<uint8_t> h(ix.n * ix.key_bytes);
  if (ix.h_col) {
    std::memcpy(h.data(), ix.h_col, h.size());
  } else {
    GJ_HIP_CHECK(hipMemcpyAsync(h.data(), ix.d_col, h.size(), hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
  }
  return h;
}

template <class K>
void build_rs(gj_index& ix, hipStream_t s) {
  const uint32_t bits = ix.params.rs_radix_bits;
  GJ_REQUIRE(bits >= 1 && bits <= 30, GJ_ERR_INVALID, "rs_radix_bits must be in [1, 30]");
  std::vector<uint8_t> hcol;
  const void* hk = ix.h_col;
  if (!hk) {
    hcol = host_column(ix, s);
    hk = hcol.data();
  }
  RadixSplineHost rs = build_radix_spline(hk, ix.n, ix.key_bytes, bits, ix.params.rs_max_error);
  ix.rs_shift = rs.shift_bits;
  ix.rs_points = rs.knot_x.size();
  ix.rs_radix.alloc(rs.radix.size() * sizeof(uint32_t));
  GJ_HIP_CHECK(hipMemcpyAsync(ix.rs_radix.p, rs.radix.data(), ix.rs_radix.bytes, hipMemcpyHostToDevice, s));
  // staged in LDS when the table has at most 2^rs_lds_bits + 2 entries (2^14: 64 KiB, a
  // workgroup's LDS)
  ix.rs_lds = rs.radix.size() <= (1ull << std::min<uint32_t>(ix.params.rs_lds_bits, 14)) + 2;
  ix.stats.lds_levels = ix.rs_lds ? 1 : 0;
  std::vector<RsKnot> kn(rs.knot_x.size());
  for (size_t i = 0; i < kn.size(); ++i) kn[i] = RsKnot{static_cast<uint64_t>(rs.knot_x[i]), rs.knot_y[i]};
  ix.rs_knots.alloc(kn.size() * sizeof(RsKnot));
  GJ_HIP_CHECK(hipMemcpyAsync(ix.rs_knots.p, kn.data(), ix.rs_knots.bytes, hipMemcpyHostToDevice, s));
  ix.rs_err = ix.params.rs_max_error;
  dbuf bad(sizeof(unsigned long long));
  for (int attempt = 0; attempt < 40; ++attempt) {
    fill_bytes(bad.p, 0, bad.bytes, s);
    hipLaunchKernelGGL(rs_check_kernel<K>, dim3(grid_for(ix.n, 256, 8192)), dim3(256), 0, s, ix.rs_dev<K, false>(),
                       bad.as<unsigned long long>());
    GJ_HIP_CHECK(hipGetLastError());
    unsigned long long h = 0;
    GJ_HIP_CHECK(hipMemcpyAsync(&h, bad.p, sizeof(h), hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    if (h == 0) break;
    GJ_REQUIRE(attempt < 39, GJ_ERR_INTERNAL, "radix spline: search bound does not converge");
    ix.rs_err += std::max<uint32_t>(1, ix.rs_err / 4);
  }
  ix.stats.spline_points = ix.rs_points;
  ix.stats.max_window = 2 * ix.rs_err + 2;
}

// ---------------------------------------------------------------------------
// Hash table (linear_probing_hashtable, src/linear_probing_hashtable.cuh:37-93).
// ---------------------------------------------------------------------------
template <class K>
__global__ void ht_insert_kernel(const K* __restrict__ col, uint64_t n, HashEntry* table, uint64_t mask) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const unsigned long long k = static_cast<uint64_t>(col[i]);
    uint64_t s = key_hash<K>(k) & mask;
    for (;;) {
      const unsigned long long prev = atomicCAS(&table[s].key, ~0ull, k);
      if (prev == ~0ull || prev == k) {
        table[s].value = static_cast<uint32_t>(i);
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

// Atomic-free build of the table for large columns.  The hashed keys h = key_hash(k)
// are radix-partitioned on bits [range_bits, cap_bits) of h, that is by the table
// range of 2^range_bits slots that holds each key's home slot h & (cap - 1).  One
// workgroup builds a range in LDS, linear probing with LDS compare-and-swap and the
// key recovered from its hash, then writes the whole range out in full lines, empty
// slots included (so the table needs no fill).  A key whose probe run passes the end of
// its range is moved to the front of the range's tuples and counted; ht_spill_kernel
// then inserts those keys into the finished table.  Every
// key sits at or after its home slot with no empty slot between, which is all a
// lookup (HashDev) relies on; only the order inside a run can differ from an
// insert-built table.
// A range of 2^range_bits slots takes 12 B of LDS per slot (8 B key + 4 B row id): 2^12
// (48 KiB, two workgroups per CU) where the partitioner's 2^16 partitions cover the table,
// 2^13 (96 KiB) for the largest tables.
constexpr uint32_t kHtRangeMinBits = 12;
constexpr uint32_t kHtRangeMaxBits = 13;
constexpr uint32_t kHtRangeBlock = 1024;
constexpr int kHtRangeItems = 4;

template <class K>
__global__ __launch_bounds__(kHtRangeBlock) void ht_range_build_kernel(K* __restrict__ hkeys, uint32_t* __restrict__ rows,
                                                                      const uint64_t* __restrict__ start,
                                                                      const uint64_t* __restrict__ count,
                                                                      uint32_t* __restrict__ spills,
                                                                      HashEntry* __restrict__ table,
                                                                      uint32_t range_bits) {
  const uint32_t S = 1u << range_bits;
  extern __shared__ unsigned long long lk[];  // [S] keys, then [S] row ids
  uint32_t* lv = reinterpret_cast<uint32_t*>(lk + S);
  __shared__ uint32_t nspill;
  const uint32_t b = blockIdx.x;
  for (uint32_t j = threadIdx.x; j < S; j += blockDim.x) {
    lk[j] = ~0ull;
    lv[j] = 0xFFFFFFFFu;
  }
  if (threadIdx.x == 0) nspill = 0;
  __syncthreads();
  const uint64_t lo = start[b], n = count[b];
  // A range holds about 2^range_bits / 2 keys: kHtRangeItems per thread read them in one
  // round of loads.
  const uint64_t round = static_cast<uint64_t>(kHtRangeItems) * blockDim.x;
  for (uint64_t c0 = 0; c0 < n; c0 += round) {
    K h[kHtRangeItems];
    uint32_t r[kHtRangeItems];
#pragma unroll
    for (int i = 0; i < kHtRangeItems; ++i) {
      const uint64_t t = c0 + static_cast<uint64_t>(i) * blockDim.x + threadIdx.x;
      h[i] = t < n ? hkeys[lo + t] : K(0);
      r[i] = t < n ? rows[lo + t] : 0u;
    }
    // every read of this round comes before the round's spill stores, which reuse
    // slots already read (a spill's slot is below the count of tuples read so far)
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kHtRangeItems; ++i) {
      const uint64_t t = c0 + static_cast<uint64_t>(i) * blockDim.x + threadIdx.x;
      if (t >= n) continue;
      const unsigned long long k = key_unhash<K>(static_cast<uint64_t>(h[i]));
      bool placed = false;
      for (uint32_t j = static_cast<uint32_t>(h[i]) & (S - 1); j < S; ++j) {
        const unsigned long long prev = atomicCAS(&lk[j], ~0ull, k);
        if (prev == ~0ull || prev == k) {
          lv[j] = r[i];
          placed = true;
          break;
        }
      }
      if (!placed) {
        const uint32_t q = atomicAdd(&nspill, 1u);
        hkeys[lo + q] = h[i];
        rows[lo + q] = r[i];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) spills[b] = nspill;
  u32x4* dst = reinterpret_cast<u32x4*>(table + (static_cast<uint64_t>(b) << range_bits));
  for (uint32_t j = threadIdx.x; j < S; j += blockDim.x) {
    u32x4 e;
    e.x = static_cast<uint32_t>(lk[j]);
    e.y = static_cast<uint32_t>(lk[j] >> 32);
    e.z = lv[j];
    e.w = 0xFFFFFFFFu;
    store_stream(dst + j, e);
  }
}

// The keys each range could not hold: one workgroup per range inserts them as
// ht_insert_kernel does, starting at the range's end (every slot from such a key's home
// to the end of its range was taken when it spilled).
template <class K>
__global__ __launch_bounds__(256) void ht_spill_kernel(const K* __restrict__ hkeys, const uint32_t* __restrict__ rows,
                                                       const uint64_t* __restrict__ start,
                                                       const uint32_t* __restrict__ spills, HashEntry* table,
                                                       uint64_t mask, uint32_t range_bits) {
  const uint64_t b = blockIdx.x;
  const uint32_t ns = spills[b];
  for (uint32_t q = threadIdx.x; q < ns; q += blockDim.x) {
    const uint64_t h = static_cast<uint64_t>(hkeys[start[b] + q]);
    const uint32_t r = rows[start[b] + q];
    const unsigned long long k = key_unhash<K>(h);
    uint64_t s = ((b + 1) << range_bits) & mask;
    for (;;) {
      const unsigned long long prev = atomicCAS(&table[s].key, ~0ull, k);
      if (prev == ~0ull || prev == k) {
        table[s].value = r;
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

template <class K>
__global__ void ht_hash_kernel(const K* __restrict__ col, uint64_t n, K* __restrict__ h) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    store_stream(h + i, static_cast<K>(key_hash<K>(static_cast<uint64_t>(load_stream(col + i)))));
}

}  // namespace

// calculate_table_size, src/linear_probing_hashtable.cuh:74-79
uint64_t hash_capacity(uint64_t n) {
  const float hint = static_cast<float>(n) / 0.5f;
  const uint64_t bits = static_cast<uint64_t>(std::ceil(std::log2(static_cast<double>(hint))));
  return 1ull << bits;
}

// Tables of 2^13 .. 2^(kHtRangeMaxBits + 16) slots are built by range (the partitioner
// takes up to 2^16 partitions), smaller and larger ones by inserts.
bool hash_build_by_range(uint64_t cap) {
  const uint32_t bits = ilog2_floor(cap);
  return bits > kHtRangeMinBits && bits <= kHtRangeMaxBits + 16;
}
uint32_t hash_range_bits(uint64_t cap) {
  const uint32_t bits = ilog2_floor(cap);
  return bits > kHtRangeMinBits + 16 ? bits - 16 : kHtRangeMinBits;
}

template <class K>
void build_hash_table(const gj_index& ix, dbuf& table, uint64_t& cap, dbuf& temp, hipStream_t s) {
  GJ_REQUIRE(!(sizeof(K) == 8 && ix.max_key == ~0ull), GJ_ERR_INVALID,
             "hash: the key 2^64-1 is the empty-slot marker and cannot be stored");
  cap = hash_capacity(ix.n);
  table.reserve(cap * sizeof(HashEntry));
  if (!hash_build_by_range(cap)) {
    fill_bytes(table.p, 0xFF, cap * sizeof(HashEntry), s);
    hipLaunchKernelGGL(ht_insert_kernel<K>, dim3(grid_for(ix.n, 256)), dim3(256), 0, s,
                       static_cast<const K*>(ix.d_col), ix.n, table.as<HashEntry>(), cap - 1);
    GJ_HIP_CHECK(hipGetLastError());
    return;
  }
  const uint32_t rb = hash_range_bits(cap);
  const uint32_t bits = ilog2_floor(cap) - rb;  // ranges = partitions
  const uint64_t ranges = 1ull << bits, n = ix.n;
  const uint64_t tcap = partition_capacity(n, sizeof(K), bits);
  // temp: hashes | partitioned hashes | row ids | start (ranges + 1) | count | spills | partitioner scratch
  auto up = [](uint64_t b) { return (b + 255) / 256 * 256; };
  const uint64_t o_pk = up(n * sizeof(K)), o_pr = o_pk + up(tcap * sizeof(K)), o_st = o_pr + up(tcap * 4),
                 o_ct = o_st + up((ranges + 1) * 8), o_sp = o_ct + up(ranges * 8), o_ps = o_sp + up(ranges * 4),
                 bytes = o_ps + up(partition_scratch(n, sizeof(K), bits, 1));
  temp.reserve(bytes);
  uint8_t* base = temp.as<uint8_t>();
  K* hk = reinterpret_cast<K*>(base);
  K* pk = reinterpret_cast<K*>(base + o_pk);
  uint32_t* pr = reinterpret_cast<uint32_t*>(base + o_pr);
  uint64_t* st = reinterpret_cast<uint64_t*>(base + o_st);
  uint64_t* ct = reinterpret_cast<uint64_t*>(base + o_ct);
  uint32_t* sp = reinterpret_cast<uint32_t*>(base + o_sp);
  hipLaunchKernelGGL(ht_hash_kernel<K>, dim3(grid_for(n, 256)), dim3(256), 0, s, static_cast<const K*>(ix.d_col), n, hk);
  GJ_HIP_CHECK(hipGetLastError());
  // payload = row id; the digit is the range of the home slot
  partition_keys<K>(hk, n, 0, bits, rb, 1, false, pk, pr, st, ct, base + o_ps,
                    ix.dropped.as<unsigned long long>(), ix.cu, s);
  const size_t lds = (static_cast<size_t>(1) << rb) * (8 + 4);
  const void* fn = reinterpret_cast<const void*>(&ht_range_build_kernel<K>);
  if (lds > 65536)
    GJ_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  hipLaunchKernelGGL(ht_range_build_kernel<K>, dim3(static_cast<unsigned>(ranges)), dim3(kHtRangeBlock), lds, s, pk, pr,
                     st, ct, sp, table.as<HashEntry>(), rb);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(ht_spill_kernel<K>, dim3(static_cast<unsigned>(ranges)), dim3(256), 0, s, pk, pr, st, sp,
                     table.as<HashEntry>(), cap - 1, rb);
  GJ_HIP_CHECK(hipGetLastError());
}
template void build_hash_table<uint32_t>(const gj_index&, dbuf&, uint64_t&, dbuf&, hipStream_t);
template void build_hash_table<uint64_t>(const gj_index&, dbuf&, uint64_t&, dbuf&, hipStream_t);

uint64_t extended_memory(const gj_index& ix) {
  return ix.hm_keys.bytes + ix.rs_radix.bytes + ix.rs_knots.bytes +
         ix.lm_leaves.bytes + ix.lm_mlp.bytes + ix.lm_knots.bytes + ix.lm_seps.bytes + ix.lm_subs.bytes +
         ix.lm_groups.bytes + ix.lm_lines.bytes + ix.lm_line_leaves.bytes + ix.ht.bytes;
}

template <class K>
void build_extended(gj_index& ix, hipStream_t s) {
  switch (ix.kind) {
    case GJ_HARMONIA: build_harmonia<K>(ix, s); break;
    case GJ_RADIX_SPLINE: build_rs<K>(ix, s); break;
    case GJ_LEARNED_PLR:
    case GJ_LEARNED_MLP:
      build_btree<K>(ix, s);
      build_learned<K>(ix, s);
      break;
    case GJ_HASH: {
      dbuf temp;
      build_hash_table<K>(ix, ix.ht, ix.ht_cap, temp, s);
      GJ_HIP_CHECK(hipStreamSynchronize(s));  // temp is freed on return
      break;
    }
    default:
      throw error(GJ_ERR_UNSUPPORTED, std::string("index kind not built: ") + gj_index_type_name(ix.kind));
  }
}
template void build_extended<uint32_t>(gj_index&, hipStream_t);
template void build_extended<uint64_t>(gj_index&, hipStream_t);

}  // namespace gj
