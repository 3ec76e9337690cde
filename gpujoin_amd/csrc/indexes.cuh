// indexes.cuh - device-side index descriptors.
//
// Each descriptor is a small POD passed BY VALUE as a kernel argument, the
// role of the reference's device_index_t (src/indexes.cuh:18-22, 108-124).
// Every descriptor provides
//   size_t   lds_bytes() const                      bytes of LDS the lookup wants resident
//   void     stage(uint8_t* lds) const              block-cooperative fill of that LDS (no barrier)
//   uint32_t lookup(K x, bool active, const uint8_t* lds) const
// lookup() must be called by all 64 lanes of a wave together (like the
// reference's whole-warp cooperative_lookup, src/btree.cuh:598-605,
// src/harmonia.cuh:740-745); inactive lanes pass active = false and get
// kNotFound.  The result is the row id (position in the sorted column) or
// kNotFound when the key is absent.
#pragma once

#include "device.cuh"

namespace gj {

constexpr int kMaxLevels = 16;

// ---------------------------------------------------------------------------
// Binary search (binary_search_index, src/indexes.cuh:336-375 with
// lower_bound, src/search.cuh:26-41).  One probe per lane.  The column is
// cut into 2^sample_bits equal buckets; the max key of every bucket sits in
// LDS, so the first sample_bits steps of the search never leave the CU.  The
// rest is a branch-free lower_bound inside the bucket, then an equality check.
// ---------------------------------------------------------------------------
template <class K>
struct BSearchDev {
  const K* col;
  uint64_t n;
  const K* samples;  // 2^sample_bits bucket maxima (key_max for empty buckets)
  uint64_t step;     // keys per bucket
  uint32_t sample_bits;

  __host__ __device__ size_t lds_bytes() const { return sizeof(K) << sample_bits; }
  __device__ void stage(uint8_t* lds) const { stage_to_lds(lds, samples, lds_bytes()); }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    const K* s = reinterpret_cast<const K*>(lds);
    const uint32_t nb = 1u << sample_bits;
    uint32_t b = 0;
    for (uint32_t h = nb >> 1; h > 0; h >>= 1) b = (s[b + h - 1] < x) ? b + h : b;
    b += (s[b] < x) ? 1u : 0u;
    if (!active || b >= nb) return kNotFound;
    const uint64_t lo = static_cast<uint64_t>(b) * step;
    if (lo >= n) return kNotFound;
    const uint64_t hi = (lo + step < n) ? lo + step : n;
    const K* base = col + lo;
    uint64_t len = hi - lo;
    while (len > 1) {
      const uint64_t half = len >> 1;
      base = (base[half] < x) ? base + half : base;
      len -= half;
    }
    const uint64_t p = static_cast<uint64_t>(base - col) + ((base[0] < x) ? 1u : 0u);
    return (p < hi && col[p] == x) ? static_cast<uint32_t>(p) : kNotFound;
  }
};

// ---------------------------------------------------------------------------
// Cooperative 128-byte node search in groups of G = 8 lanes (8 x 16 B = one
// 128-byte line).  Each lane owns one probe; the group serves its G probes
// in G rounds: in round r all G lanes fetch their 16-byte slice of the node
// of lane (gbase + r), and a ballot counts the keys below that lane's probe.
// All G loads of a level are issued before any compare, so a level costs one
// memory latency.  This is the CDNA form of Harmonia's leader rotation
// (src/harmonia.cuh:561-599) with one full line per load instruction.  NB is the
// node size in bytes (G = NB / 16 lanes per group): 128 by default, 64 or 256 for
// the B-tree's other node sizes.
// ---------------------------------------------------------------------------
template <class K, int G, int NB = 128>
__device__ __forceinline__ uint32_t group_count_less(const K* level_base, uint64_t my_node, K my_x, int gbase,
                                                     int gl) {
  constexpr int NK = NB / static_cast<int>(sizeof(K));
  constexpr int KPL = NK / G;
  static_assert(KPL == Piece<K>::kKeys, "one 16-byte piece per lane");
  Piece<K> pc[G];
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const uint64_t nd = shfl64(my_node, gbase + r);
    pc[r] = load_piece(level_base + nd * NK + gl * KPL);
  }
  // each round's probe is fetched from its lane when compared (while the lines are in
  // flight), so only the G lines stay live across the level: 108 -> 64 VGPRs for u64
  // keys, 4 -> 8 waves per SIMD
  uint32_t result = 0;
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const K xr = shfl(my_x, gbase + r);
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < KPL; ++j) cnt += group_popc<G>(ballot(pc[r].k[j] < xr), gbase);
    if (gl == r) result = static_cast<uint32_t>(cnt);
  }
  return result;
}

// Equality search of each lane's probe in its NB-byte column chunk.
template <class K, int G, int NB = 128>
__device__ __forceinline__ uint32_t group_find_equal(const K* col, uint64_t n, uint64_t my_chunk, K my_x,
                                                     bool my_live, int gbase, int gl) {
  constexpr int NK = NB / static_cast<int>(sizeof(K));
  constexpr int KPL = NK / G;
  Piece<K> pc[G];
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const uint64_t chunk = shfl64(my_chunk, gbase + r);
    const uint64_t start = chunk * NK + static_cast<uint64_t>(gl * KPL);
    const int valid = (start >= n) ? 0 : ((n - start) >= static_cast<uint64_t>(KPL) ? KPL : static_cast<int>(n - start));
    pc[r] = load_piece_guarded(col + start, valid, key_max<K>());
  }
  // as in group_count_less, each round's probe is fetched again when compared; the owning
  // lane forms the match position from its own chunk.  Slots past the column's end hold the
  // key_max fill: the lowest matching slot is taken (the column's keys are unique, so a
  // real match precedes any fill) and a position past the end reads as not found.
  uint64_t pos = ~0ull;
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const K xr = shfl(my_x, gbase + r);
    int best = NK;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
      const int f = group_ffs<G>(ballot(pc[r].k[j] == xr), gbase);
      if (f >= 0 && f * KPL + j < best) best = f * KPL + j;
    }
    if (gl == r && best < NK) pos = my_chunk * NK + static_cast<uint64_t>(best);
  }
  return (my_live && pos < n) ? static_cast<uint32_t>(pos) : kNotFound;
}

// ---------------------------------------------------------------------------
// B+-tree, re-laid out for CDNA4.  Keeps the reference's clustered-index
// semantics (btree_index over the sorted column; leaves are column chunks,
// src/btree.cuh:242-246, 504-531) with a static implicit layout:
//   * a node is NB bytes, NK = NB / sizeof(K) keys: one 128-byte line by default
//     (16 x u64 / 32 x u32), or half a line / two lines (NB = 64 / 256);
//   * the leaves are the column itself, cut into chunks of NK keys (NB bytes each);
//   * slot s of node i holds the max key of child i*NK + s (key_max past the end),
//     so the child index is the count of separators below the probe and no
//     child pointers are stored (the reference stores pointers, src/btree.cuh:58-63);
//   * levels are stored top first; the top lds_levels levels are staged in LDS.
// ---------------------------------------------------------------------------
template <class K, int NB = 128>
struct BTreeDev {
  static_assert(NB == 64 || NB == 128 || NB == 256, "B-tree nodes of 64, 128 or 256 bytes");
  static constexpr int NK = NB / static_cast<int>(sizeof(K));
  static constexpr int G = NB / 16;  // lanes per node search, 16 bytes each
  const K* col;
  uint64_t n;
  const K* nodes;  // all inner levels, top level first, NK keys per node
  uint32_t levels;
  uint32_t lds_levels;
  uint64_t lds_nodes;
  uint64_t level_off[kMaxLevels];      // first node of level l (top = 0)
  uint64_t level_cnt[kMaxLevels + 1];  // nodes of level l; level_cnt[levels] = column chunks

  __host__ __device__ size_t lds_bytes() const { return lds_nodes * NB; }
  __device__ void stage(uint8_t* lds) const { stage_to_lds(lds, nodes, lds_nodes * NB); }

  __device__ __forceinline__ void descend(uint64_t& node, bool& live, uint32_t c, uint32_t lvl) const {
    const uint64_t child = node * NK + c;
    live = live && (c < static_cast<uint32_t>(NK)) && (child < level_cnt[lvl + 1]);
    node = live ? child : 0;
  }

  // Walks the inner levels; returns the column chunk of the probe (or live = false).
  __device__ __forceinline__ uint64_t find_chunk(K x, bool& live, const uint8_t* lds, int gbase, int gl) const {
    uint64_t node = 0;
    uint32_t lvl = 0;
    const K* lds_nodes_p = reinterpret_cast<const K*>(lds);
    for (; lvl < lds_levels; ++lvl) {
      const uint32_t c = group_count_less<K, G, NB>(lds_nodes_p + level_off[lvl] * NK, node, x, gbase, gl);
      descend(node, live, c, lvl);
    }
    for (; lvl < levels; ++lvl) {
      const uint32_t c = group_count_less<K, G, NB>(nodes + level_off[lvl] * NK, node, x, gbase, gl);
      descend(node, live, c, lvl);
    }
    return node;
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    const int ln = lane();
    const int gbase = ln & ~(G - 1);
    const int gl = ln & (G - 1);
    bool live = active;
    const uint64_t chunk = find_chunk(x, live, lds, gbase, gl);
    return group_find_equal<K, G, NB>(col, n, chunk, x, live, gbase, gl);
  }
};

// no_op_index (src/indexes.cuh:384-406): measures probe read + row id write.
template <class K>
struct NoOpDev {
  __host__ __device__ size_t lds_bytes() const { return 0; }
  __device__ void stage(uint8_t*) const {}
  __device__ __forceinline__ uint32_t lookup(K, bool, const uint8_t*) const { return kNotFound; }
};

}  // namespace gj
