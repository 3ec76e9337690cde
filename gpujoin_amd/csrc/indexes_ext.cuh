// indexes_ext.cuh - device descriptors of Harmonia, RadixSpline, the learned
// two-stage models and the hash table (same contract as indexes.cuh).
#pragma once

#include "device.cuh"
#include "indexes.cuh"

namespace gj {

// KPL consecutive keys loaded by one lane (4, 8 or 16 bytes).
template <class K, int KPL>
struct LaneKeys {
  K k[KPL];
};
template <class K, int KPL>
__device__ __forceinline__ LaneKeys<K, KPL> load_lane_keys(const K* p) {
  constexpr int B = KPL * static_cast<int>(sizeof(K));
  LaneKeys<K, KPL> r;
  if constexpr (B == 16) {
    const u32x4 v = *reinterpret_cast<const u32x4*>(p);
    if constexpr (sizeof(K) == 8) {
      r.k[0] = (static_cast<uint64_t>(v.y) << 32) | v.x;
      r.k[1] = (static_cast<uint64_t>(v.w) << 32) | v.z;
    } else {
      r.k[0] = v.x; r.k[1] = v.y; r.k[2] = v.z; r.k[3] = v.w;
    }
  } else if constexpr (B == 8) {
    const u32x2 v = *reinterpret_cast<const u32x2*>(p);
    if constexpr (sizeof(K) == 8) {
      r.k[0] = (static_cast<uint64_t>(v.y) << 32) | v.x;
    } else {
      r.k[0] = v.x; r.k[1] = v.y;
    }
  } else {
    static_assert(B == 4, "4, 8 or 16 bytes per lane");
    r.k[0] = *p;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Harmonia (harmonia_index, src/indexes.cuh:257-307; harmonia_tree,
// src/harmonia.cuh:48-992).  Layout kept from the reference: 32 keys per node
// (fanout 33), inner keys breadth first and padded with key_max, leaf keys =
// the column, next = children[pos] + lb (src/harmonia.cuh:748-766).  Re-laid
// out for CDNA4: the build fills every node of a level but the last, so the
// reference's children[] prefix sums have a closed form,
//   children[node p of inner level l] = first[l + 1] + 33 (p - first[l]),
// which the lookup computes instead of reading a 4-byte entry per level (at
// 2^26 keys a 16.5 MiB array whose bottom-level reads miss L2); the top key
// levels sit in LDS (the reference uses 42 KiB of __constant__,
// src/harmonia.cuh:45-46); a node is searched by a group of G lanes of the
// wave64 that fetch one 128-byte line per load and read a u64 node's second
// line only when every key of the first line is below the probe (the
// reference's early exit, src/harmonia.cuh:581-589).
// ---------------------------------------------------------------------------
template <class K, int G>
struct HarmoniaDev {
  static constexpr int MK = 32;                                   // keys per node
  static constexpr int LK = 128 / static_cast<int>(sizeof(K));    // keys per 128-byte line
  static constexpr int LINES = MK / LK;                           // 2 (u64) or 1 (u32)
  static constexpr int KPL = LK / G;                              // keys per lane per line
  static_assert(KPL >= 1 && KPL * G == LK, "group must tile a line");

  const K* col;
  uint64_t n;
  const K* keys;              // inner nodes x MK
  uint64_t inner_nodes;
  uint32_t depth;             // levels including the leaf level
  uint32_t lds_levels;        // top inner levels staged in LDS
  uint64_t lds_nodes;
  uint64_t first[kMaxLevels];  // first node of level l (top = 0); first[depth - 1] = inner_nodes

  __host__ __device__ size_t lds_bytes() const { return lds_nodes * MK * sizeof(K); }
  __device__ void stage(uint8_t* lds) const {
    if (lds_nodes) stage_to_lds(lds, keys, lds_bytes());
  }

  // Count of node keys below each lane's probe (and whether the key at that
  // position equals the probe), for the G probes of the lane's group.  (Fetching each
  // round's probe and offset from its lane where used, as the B-tree's search does, takes
  // this kernel from 97 to 69 VGPRs and 4 to 7 waves per SIMD but from 18.0 to 14.8 G
  // lookups/s at 2^26 u64 keys; shuffling only the probes, 16.9: the arrays stay.)
  __device__ __forceinline__ uint32_t node_search(const K* base, uint64_t my_off, K my_x, int gbase, int gl,
                                                  bool& my_eq) const {
    uint64_t off[G];
    K xs[G];
    LaneKeys<K, KPL> v[G];
#pragma unroll
    for (int r = 0; r < G; ++r) {
      off[r] = shfl64(my_off, gbase + r);
      xs[r] = shfl(my_x, gbase + r);
      v[r] = load_lane_keys<K, KPL>(base + off[r] + gl * KPL);
    }
    int c[G];
    bool e[G];
#pragma unroll
    for (int r = 0; r < G; ++r) {
      c[r] = 0;
      e[r] = false;
#pragma unroll
      for (int j = 0; j < KPL; ++j) {
        c[r] += group_popc<G>(ballot(v[r].k[j] < xs[r]), gbase);
        e[r] |= group_popc<G>(ballot(v[r].k[j] == xs[r]), gbase) != 0;
      }
    }
    if constexpr (LINES == 2) {
      // every key of the first line is below the probe: read the second line; all such
      // rounds request their second line before any is examined (one more latency, not G)
#pragma unroll
      for (int r = 0; r < G; ++r)
        if (c[r] == LK) v[r] = load_lane_keys<K, KPL>(base + off[r] + LK + gl * KPL);
#pragma unroll
      for (int r = 0; r < G; ++r) {
        if (c[r] != LK) continue;
        int c2 = 0;
        bool e2 = false;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
          c2 += group_popc<G>(ballot(v[r].k[j] < xs[r]), gbase);
          e2 |= group_popc<G>(ballot(v[r].k[j] == xs[r]), gbase) != 0;
        }
        c[r] += c2;
        e[r] = e2;
      }
    }
    uint32_t res = 0;
    my_eq = false;
#pragma unroll
    for (int r = 0; r < G; ++r) {
      if (gl == r) {
        res = static_cast<uint32_t>(c[r]);
        my_eq = e[r];
      }
    }
    return res;
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    const int ln = lane();
    const int gbase = ln & ~(G - 1);
    const int gl = ln & (G - 1);
    bool live = active;
    bool eq = false;
    uint64_t pos = 0;
    uint32_t d = 0;
    const K* lds_keys = reinterpret_cast<const K*>(lds);
    for (; d < lds_levels; ++d) {
      const uint32_t lb = node_search(lds_keys, pos * MK, x, gbase, gl, eq);
      const uint64_t next = first[d + 1] + 33ull * (pos - first[d]) + lb;
      pos = live ? next : 0;
    }
    for (; d + 1 < depth; ++d) {
      const uint32_t lb = node_search(keys, pos * MK, x, gbase, gl, eq);
      const uint64_t next = first[d + 1] + 33ull * (pos - first[d]) + lb;
      pos = live ? next : 0;
    }
    const uint64_t leaf = live ? pos - inner_nodes : 0;
    const uint32_t lb = node_search(col, leaf * MK, x, gbase, gl, eq);
    return (live && eq) ? static_cast<uint32_t>(leaf * MK + lb) : kNotFound;
  }
};

// ---------------------------------------------------------------------------
// Last-mile window search shared by RadixSpline and the learned models.
// Each lane's probe has a predicted position `pred` and a window [lo, hi) of
// the column that holds the key if it is present.  Groups of kWinG = 4 lanes
// read whole 128-byte lines (32 B per lane); a group serves its 4 probes in 4
// rounds.  A round starts at the line holding pred and walks one line left or
// right until the line brackets the probe or leaves the window.  The search
// runs in phases: in each phase every unresolved round of every group requests
// its next line before any line is examined, so a wave pays one memory latency
// per phase (one, occasionally two or three) rather than one per walking round.
//
// A round's line is sorted, so the group only needs two facts about it for the
// owning lane: how many of its keys are below the probe (`below`, which is the
// in-line offset of the probe's lower bound) and whether the key at that offset
// equals the probe.  The owning lane alone holds its probe's state (line,
// window, direction), which keeps the registers per lane, and hence the waves
// per SIMD, where a latency-bound search needs them.  Positions are 32-bit
// (n < 2^32); keys at or past n read as the all-ones fill, which sorts last.
// ---------------------------------------------------------------------------
constexpr int kWinG = 4;

template <class K>
struct LineProbe {
  static constexpr int LK = 128 / static_cast<int>(sizeof(K));  // keys per line
  static constexpr int KPL = LK / kWinG;                         // keys per lane (32 bytes)
  K k[KPL];
  __device__ __forceinline__ void load(const K* col, uint64_t n, uint32_t line, int gl) {
    load(col, n, 0u, gl, static_cast<uint64_t>(line));
  }
  // line given as 64 bits (the line copies of the lines layout can exceed 2^32 slots)
  __device__ __forceinline__ void load(const K* col, uint64_t n, uint32_t, int gl, uint64_t line) {
    const uint64_t start = line * LK + static_cast<uint64_t>(gl * KPL);
    if (start + KPL <= n) {
      const u32x4 a = *reinterpret_cast<const u32x4*>(col + start);
      const u32x4 b = *reinterpret_cast<const u32x4*>(col + start + KPL / 2);
      if constexpr (sizeof(K) == 8) {
        k[0] = (static_cast<uint64_t>(a.y) << 32) | a.x;
        k[1] = (static_cast<uint64_t>(a.w) << 32) | a.z;
        k[2] = (static_cast<uint64_t>(b.y) << 32) | b.x;
        k[3] = (static_cast<uint64_t>(b.w) << 32) | b.z;
      } else {
        k[0] = a.x; k[1] = a.y; k[2] = a.z; k[3] = a.w;
        k[4] = b.x; k[5] = b.y; k[6] = b.z; k[7] = b.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < KPL; ++j) k[j] = (start + j < n) ? col[start + j] : key_max<K>();
    }
  }
  // keys of the group's line below x, and whether the line holds x
  __device__ __forceinline__ void eval(K x, int gbase, uint32_t& below, bool& eq) const {
    below = 0;
    bool e = false;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
      below += static_cast<uint32_t>(group_popc<kWinG>(ballot(k[j] < x), gbase));
      e |= group_bits<kWinG>(ballot(k[j] == x), gbase) != 0;
    }
    eq = e;
  }
};

template <class K>
__device__ __forceinline__ uint32_t group_window_search(const K* col, uint64_t n, uint32_t my_pred, uint32_t my_lo,
                                                        uint32_t my_hi, K my_x, bool my_live) {
  using LP = LineProbe<K>;
  constexpr int G = kWinG;
  constexpr uint32_t LK = LP::LK;
  const int ln = lane();
  const int gbase = ln & ~(G - 1);
  const int gl = ln & (G - 1);
  bool open = my_live && my_lo < my_hi;
  uint32_t line = 0, found = kNotFound;
  if (open) {
    const uint32_t p = my_pred < my_lo ? my_lo : (my_pred >= my_hi ? my_hi - 1 : my_pred);
    line = p / LK;
  }
  const uint32_t first_line = my_lo / LK, last_line = open ? (my_hi - 1) / LK : 0u;
  int dir = 0;  // 0: first line, -1 walking left, +1 walking right
  for (;;) {
    const uint64_t open_mask = ballot(open);
    if (open_mask == 0) break;  // wave-uniform exit
    LP pc[G];
#pragma unroll
    for (int r = 0; r < G; ++r) {
      const uint32_t l = shfl32(line, gbase + r);
      if ((open_mask >> (gbase + r)) & 1ull) pc[r].load(col, n, l, gl);
    }
    uint32_t below = 0;
    bool eq = false;
#pragma unroll
    for (int r = 0; r < G; ++r) {
      if (!((open_mask >> (gbase + r)) & 1ull)) continue;
      uint32_t b;
      bool e;
      pc[r].eval(shfl(my_x, gbase + r), gbase, b, e);
      if (gl == r) {
        below = b;
        eq = e;
      }
    }
    if (open) {
      const uint32_t base = line * LK;
      const uint32_t a = my_lo > base ? my_lo - base : 0u;               // window slots of this line: [a, c)
      const uint32_t c = my_hi - base < LK ? my_hi - base : LK;
      if (eq && static_cast<uint64_t>(base) + below < n) {
        found = base + below;  // the line holds x at its lower-bound offset
        open = false;
      } else if (below >= c && dir >= 0 && line < last_line) {
        ++line;  // every window key of the line is below x: only the right can hold it
        dir = 1;
      } else if (below <= a && dir <= 0 && line > first_line) {
        --line;  // no window key of the line is below x: only the left can hold it
        dir = -1;
      } else {
        open = false;  // x falls between two window keys (or past the window): absent
      }
    }
  }
  return found;
}

// Number of the 16 u64 keys of line `my_line` (of `lines`) that are below each
// lane's probe, for the lanes with my_live; whole-wave call (4-lane groups, one
// 128-byte line per round, all rounds' lines requested before any is examined).
template <class K>
__device__ __forceinline__ uint32_t group_line_count(const uint64_t* lines, uint32_t my_line, K my_x, bool my_live) {
  using LP = LineProbe<uint64_t>;
  constexpr int G = kWinG;
  const int ln = lane();
  const int gbase = ln & ~(G - 1);
  const int gl = ln & (G - 1);
  const uint64_t live_mask = ballot(my_live);
  LP pc[G];
#pragma unroll
  for (int r = 0; r < G; ++r) {
    const uint32_t l = shfl32(my_line, gbase + r);
    if ((live_mask >> (gbase + r)) & 1ull) pc[r].load(lines, ~0ull, l, gl);
  }
  uint32_t below = 0;
#pragma unroll
  for (int r = 0; r < G; ++r) {
    if (!((live_mask >> (gbase + r)) & 1ull)) continue;
    uint32_t b;
    bool e;
    pc[r].eval(shfl64(static_cast<uint64_t>(my_x), gbase + r), gbase, b, e);
    if (gl == r) below = b;
  }
  return below;
}

// ---------------------------------------------------------------------------
// RadixSpline (radix_spline_index, src/indexes.cuh:176-248; estimate
// src/rs.cuh:118-154).  Knots are {u64 key, f64 position} pairs, 16 bytes, like
// the reference's Coord (src/rs.cuh:38-52): the two knots around a probe are 32
// contiguous bytes, read by two 16-byte loads issued together.  The radix table
// is staged in LDS when it fits the rs_lds_bits budget (Lds = true: up to 2^13 + 2
// entries, 32 KiB, by default), else read from global memory (the reference's 18
// bits are 1 MiB).  A coarser table in LDS is possible (every 2^d-th entry of a
// 2^b-entry table is the table for b - d bits, since entry p is the first knot whose
// prefix is >= p), but it widens the knot search: at 2^26 keys, 13 bits in LDS ran
// at 18.6 G lookups/s against 23.8 for 18 bits in global memory.  The estimate is
// the reference's FP64 interpolation with an explicit fma, and the last mile is the
// line-walking window search over [est - err, est + err + 2).
// ---------------------------------------------------------------------------
struct RsKnot {
  uint64_t x;
  double y;
};
static_assert(sizeof(RsKnot) == 16, "16-byte knot");

template <class K, bool Lds = false>
struct RadixSplineDev {
  const K* col;
  uint64_t n;
  const uint32_t* radix;  // radix table over the top bits of key - min_key
  uint32_t radix_count;   // entries
  const RsKnot* knots;
  K min_key, max_key;
  uint32_t shift_bits;    // prefix = (key - min_key) >> shift_bits
  uint32_t err;  // search bound (>= the builder's bound; widened at build until every key is inside)

  __host__ __device__ size_t lds_bytes() const {
    return Lds ? (static_cast<size_t>(radix_count) * 4 + 15) / 16 * 16 : 0;
  }
  __device__ void stage(uint8_t* lds) const {
    if constexpr (Lds) stage_to_lds(lds, radix, lds_bytes());
  }

  // tab: the radix table (LDS during lookups, global memory at build)
  __device__ __forceinline__ double estimate(K x, const uint32_t* tab) const {
    if (x <= min_key) return 0.0;
    if (x >= max_key) return static_cast<double>(n - 1);
    const uint64_t prefix = static_cast<uint64_t>(x - min_key) >> shift_bits;
    uint64_t lower = tab[prefix];
    uint64_t count = static_cast<uint64_t>(tab[prefix + 1]) - lower;
    while (count > 0) {
      const uint64_t step = count >> 1, mid = lower + step;
      if (knots[mid].x < static_cast<uint64_t>(x)) {
        lower = mid + 1;
        count -= step + 1;
      } else {
        count = step;
      }
    }
    const RsKnot a = knots[lower - 1], b = knots[lower];
    const double slope = (b.y - a.y) / static_cast<double>(b.x - a.x);
    return __builtin_fma(static_cast<double>(static_cast<uint64_t>(x) - a.x), slope, a.y);
  }

  __device__ __forceinline__ void window(double est, uint64_t& lo, uint64_t& hi) const {
    const double e = static_cast<double>(err);
    lo = (est < e) ? 0 : static_cast<uint64_t>(est - e);
    hi = (est + e + 2.0 > static_cast<double>(n)) ? n : static_cast<uint64_t>(est + e + 2.0);
    if (lo > hi) lo = hi;
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    uint64_t lo = 0, hi = 0, pred = 0;
    if (active) {
      const double est = estimate(x, Lds ? reinterpret_cast<const uint32_t*>(lds) : radix);
      window(est, lo, hi);
      pred = static_cast<uint64_t>(est);
    }
    return group_window_search<K>(col, n, static_cast<uint32_t>(pred), static_cast<uint32_t>(lo),
                                  static_cast<uint32_t>(hi), x, active);
  }
};

// ---------------------------------------------------------------------------
// Learned two-stage position model (no reference counterpart).
//   root:  cdf = root(x), x = (key - min_key) / (max_key - min_key) in FP32;
//          linear (a*x + b), a monotone 1xH MLP with ramp units
//          clamp(w1*x + b1, 0, 1) (w1, w2 >= 0), or an error-bounded spline of
//          the column's CDF (any key distribution); leaf = floor(cdf * L)
//   leaf:  pred = start + floor(slope * dkey) in FP32, dkey = key - leaf base key
//          (the fit's intercept is folded into start); one 16-byte load.
//   window: [pred + err_lo, pred + err_hi], with err_lo/err_hi the extremes of
//          the error of every build key under this exact FP32 code, over all
//          leaves.  The window only caps the walk: a walk from any line stops
//          at the line that brackets the probe.  Leaves whose own error range
//          exceeds the bound go to the B-tree (slope = NaN).
// Keeping the leaf at 16 bytes halves the leaf table, which keeps more of it
// in L2 while the column lines stream through.
// ---------------------------------------------------------------------------
struct LeafModel {  // 16 bytes
  uint64_t base_key;  // first key routed to the leaf (0 for an empty leaf)
  uint32_t start;     // position of base_key plus the rounded intercept, clamped to [0, n]
  float slope;        // NaN: the leaf's probes are looked up in the B-tree
};
static_assert(sizeof(LeafModel) == 16, "leaf model is 16 bytes");

// Compact leaves (8 bytes) for roots that spread the keys evenly (linear or MLP
// root, every leaf within the error bound, no split leaf; 2^22 keys or more, the
// default leaf count): twice the leaves in
// the same table, so a smaller window and fewer column lines per lookup.  The
// key offset is taken from a group base shared by 2^gshift leaves (at most 512
// groups, 16 bytes each, staged in LDS) instead of the leaf's own base key:
//   pred = group.start + floor(fma(slope, f32(key - group.base_key), icpt))
// The build measures every key's error under exactly this arithmetic, so a
// leaf whose FP32 offsets lose precision only widens the bounds; if any leaf
// then exceeds the bound the build keeps the 16-byte leaves.
struct LeafCompact {  // 8 bytes
  float slope;
  float icpt;  // positions past group.start at key == group.base_key
};
static_assert(sizeof(LeafCompact) == 8, "compact leaf is 8 bytes");

struct LeafGroup {  // 16 bytes
  uint64_t base_key;  // key at start (the last key for a group past the column's end)
  uint32_t start;     // first position routed to the group's first leaf or later
  uint32_t pad;
};
static_assert(sizeof(LeafGroup) == 16, "leaf group is 16 bytes");

constexpr int kMaxHidden = 64;

struct RootModel {
  float a, b;                 // linear root
  float inv_range;            // 1 / (max_key - min_key) in FP32
  uint32_t hidden;            // > 0: MLP root
  const float* mlp;           // [w1(h), b1(h), w2(h), b2] in global memory (staged in LDS)
  uint32_t segments;          // > 0: spline root with this many segments (a power of two; the
                              // knots past the last real one repeat it)
  const uint64_t* knots;      // spline arrays in global memory (staged in LDS):
                              // key[segments + 1] u64 | cdf[segments + 1] f32 | slope[segments] f32

  __host__ __device__ size_t key_bytes() const { return ((static_cast<size_t>(segments) + 1) * 8 + 15) / 16 * 16; }
  __host__ __device__ size_t cdf_bytes() const { return ((static_cast<size_t>(segments) + 1) * 4 + 15) / 16 * 16; }
  __host__ __device__ size_t bytes() const {
    if (hidden) return ((3 * hidden + 1) * 4 + 15) / 16 * 16;
    if (segments) return key_bytes() + cdf_bytes() + (static_cast<size_t>(segments) * 4 + 15) / 16 * 16;
    return 0;
  }
  __host__ __device__ const void* params() const { return hidden ? static_cast<const void*>(mlp) : knots; }
};

// Spline root: the segment s with key[s] <= x < key[s + 1] by a branch-free
// binary search, then cdf = cdf[s] + (x - key[s]) * slope[s], kept inside
// [cdf[s], cdf[s + 1]].  Monotone in the key, so every leaf still owns a
// contiguous key range.  The knots come from an error-bounded greedy spline over
// (key, position / n), so a segment never spans a gap that the keys jump across.
__device__ __forceinline__ float spline_cdf(uint64_t x, const uint8_t* base, const RootModel& r) {
  const uint64_t* key = reinterpret_cast<const uint64_t*>(base);
  const float* cdf = reinterpret_cast<const float*>(base + r.key_bytes());
  const float* slope = reinterpret_cast<const float*>(base + r.key_bytes() + r.cdf_bytes());
  uint32_t s = 0;
  for (uint32_t step = r.segments >> 1; step > 0; step >>= 1) s = (key[s + step] <= x) ? s + step : s;
  const float d = x >= key[s] ? static_cast<float>(x - key[s]) : 0.0f;
  const float c = __builtin_fmaf(d, slope[s], cdf[s]);
  return fminf(fmaxf(c, cdf[s]), cdf[s + 1]);
}

__device__ __forceinline__ float mlp_eval(const float* p, uint32_t h, float x) {
  float f = p[3 * h];
  for (uint32_t j = 0; j < h; ++j) {
    const float a = __builtin_fmaf(p[j], x, p[h + j]);
    f = __builtin_fmaf(p[2 * h + j], fminf(fmaxf(a, 0.0f), 1.0f), f);
  }
  return f;
}

template <class K>
__device__ __forceinline__ float key_unit(K key, K min_key, float inv_range) {
  return static_cast<float>(static_cast<uint64_t>(key - min_key)) * inv_range;
}

__device__ __forceinline__ uint32_t leaf_of(float cdf, uint32_t leaves) {
  float l = cdf * static_cast<float>(leaves);
  l = l < 0.0f ? 0.0f : l;
  uint32_t li = static_cast<uint32_t>(l);
  return li >= leaves ? leaves - 1 : li;
}

// signed FP32 distance of key from the leaf's base key
template <class K>
__device__ __forceinline__ float leaf_dkey(K key, uint64_t base) {
  const uint64_t k = static_cast<uint64_t>(key);
  return k >= base ? static_cast<float>(k - base) : -static_cast<float>(base - k);
}

__device__ __forceinline__ int64_t leaf_pred(const LeafModel& m, float dkey) {
  float p = m.slope * dkey;
  p = fminf(fmaxf(p, -2.0e9f), 2.0e9f);
  return static_cast<int64_t>(m.start) + static_cast<int64_t>(floorf(p));
}

__device__ __forceinline__ int64_t compact_pred(const LeafCompact& m, const LeafGroup& g, float dkey) {
  float p = __builtin_fmaf(m.slope, dkey, m.icpt);
  p = fminf(fmaxf(p, -2.0e9f), 2.0e9f);
  return static_cast<int64_t>(g.start) + static_cast<int64_t>(floorf(p));
}

// The leaf a key is routed to by the root (whose parameters are staged at rp).
template <class K>
__device__ __forceinline__ uint32_t route_leaf(K x, const uint8_t* rp, const RootModel& root, K min_key,
                                               uint32_t leaves) {
  float cdf;
  if (root.segments) {
    cdf = spline_cdf(static_cast<uint64_t>(x), rp, root);
  } else if (root.hidden) {
    cdf = mlp_eval(reinterpret_cast<const float*>(rp), root.hidden, key_unit(x, min_key, root.inv_range));
  } else {
    cdf = __builtin_fmaf(root.a, key_unit(x, min_key, root.inv_range), root.b);
  }
  return leaf_of(cdf, leaves);
}

// How probes of leaves routed to the B-tree are answered:
//   kLeafNoFallback  no such leaf exists; the kernel carries no B-tree code
//   kLeafInline      the B-tree lookup runs inside the same kernel
//   kLeafMark        the probe's result is kFallbackMark; a second kernel of the
//                    plain driver scans the results and answers the marked probes
//                    with the B-tree, so the hot kernel keeps the no-fallback
//                    register count (used while few keys sit in fallback leaves)
constexpr int kLeafNoFallback = 0;
constexpr int kLeafInline = 1;
constexpr int kLeafMark = 2;
constexpr uint32_t kFallbackMark = 0xFFFFFFFEu;  // never a row id: n < 2^32 - 1

// A leaf whose keys no line fits within the bound is split into kSplitWays
// equal-count sub-leaves: the leaf's model holds slope = NaN and start = the
// split's index j; separator line j holds the max key of each sub-leaf (the last
// = all ones), and the number of separators below the probe picks the sub-leaf.
// A sub-leaf that is still too wide (slope = NaN) routes to the B-tree, as does a
// leaf with slope = NaN and start = kNoSplit.
constexpr uint32_t kSplitWays = 16;
constexpr uint32_t kNoSplit = 0xFFFFFFFFu;

template <class K, int Mode>
struct LearnedDev {
  static constexpr bool FB = Mode == kLeafInline;
  const K* col;
  uint64_t n;
  K min_key, max_key;
  RootModel root;
  const LeafModel* leaves;
  uint32_t num_leaves;
  int32_t err_lo, err_hi;  // window [pred + err_lo, pred + err_hi]
  uint32_t num_splits;     // split leaves
  const uint64_t* seps;    // num_splits separator lines of kSplitWays keys
  const LeafModel* subs;   // num_splits x kSplitWays sub-leaf models
  BTreeDev<K> btree;       // fallback (used when FB)

  __host__ __device__ size_t root_offset() const { return FB ? btree.lds_bytes() : 0; }
  __host__ __device__ size_t lds_bytes() const { return root_offset() + root.bytes(); }
  __device__ void stage(uint8_t* lds) const {
    if constexpr (FB) btree.stage(lds);
    if (root.bytes()) stage_to_lds(lds + root_offset(), root.params(), root.bytes());
  }

  __device__ __forceinline__ uint32_t route(K x, const uint8_t* lds) const {
    return route_leaf<K>(x, lds + root_offset(), root, min_key, num_leaves);
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    bool use_tree = false, split = false;
    uint64_t lo = 0, hi = 0, pred = 0;
    bool live = active && x >= min_key && x <= max_key;
    LeafModel m{};
    if (live) {
      m = leaves[route(x, lds)];
      if (__builtin_isnan(m.slope)) {
        if (m.start != kNoSplit) split = true;
        else use_tree = true;  // only with fallback leaves (Mode != kLeafNoFallback)
      }
    }
    if (num_splits && ballot(split) != 0) {  // wave-uniform
      const uint32_t t = group_line_count<K>(seps, split ? m.start : 0u, x, split);
      if (split) {
        m = subs[static_cast<uint64_t>(m.start) * kSplitWays + (t < kSplitWays ? t : kSplitWays - 1)];
        use_tree = __builtin_isnan(m.slope);
      }
    }
    if (live) {
      if (!use_tree) {
        const int64_t p = leaf_pred(m, leaf_dkey(x, m.base_key));
        int64_t a = p + err_lo, b = p + err_hi + 1;
        a = a < 0 ? 0 : a;
        b = b > static_cast<int64_t>(n) ? static_cast<int64_t>(n) : b;
        if (a < b) {
          lo = static_cast<uint64_t>(a);
          hi = static_cast<uint64_t>(b);
          pred = static_cast<uint64_t>(p < a ? a : (p >= b ? b - 1 : p));
        }
      }
    }
    uint32_t r = group_window_search<K>(col, n, static_cast<uint32_t>(pred), static_cast<uint32_t>(lo),
                                        static_cast<uint32_t>(hi), x, live && !use_tree);
    if constexpr (Mode == kLeafInline) {
      if (ballot(use_tree) != 0) {  // wave-uniform
        const uint32_t t = btree.lookup(x, use_tree, lds);
        if (use_tree) r = t;
      }
    } else if constexpr (Mode == kLeafMark) {
      if (use_tree) r = kFallbackMark;
    }
    return r;
  }
};

// The learned index with compact leaves: no fallback and no split leaf exist.
template <class K>
struct LearnedCompactDev {
  const K* col;
  uint64_t n;
  K min_key, max_key;
  RootModel root;
  const LeafCompact* leaves;
  const LeafGroup* groups;  // num_groups entries, staged in LDS after the root
  uint32_t num_leaves, gshift, num_groups;
  int32_t err_lo, err_hi;  // window [pred + err_lo, pred + err_hi]

  __host__ __device__ size_t group_offset() const { return (root.bytes() + 15) / 16 * 16; }
  __host__ __device__ size_t lds_bytes() const {
    return group_offset() + static_cast<size_t>(num_groups) * sizeof(LeafGroup);
  }
  __device__ void stage(uint8_t* lds) const {
    if (root.bytes()) stage_to_lds(lds, root.params(), root.bytes());
    stage_to_lds(lds + group_offset(), groups, static_cast<size_t>(num_groups) * sizeof(LeafGroup));
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    uint64_t lo = 0, hi = 0, pred = 0;
    const bool live = active && x >= min_key && x <= max_key;
    if (live) {
      const uint32_t l = route_leaf<K>(x, lds, root, min_key, num_leaves);
      const LeafCompact m = leaves[l];
      const LeafGroup g = reinterpret_cast<const LeafGroup*>(lds + group_offset())[l >> gshift];
      const int64_t p = compact_pred(m, g, leaf_dkey(x, g.base_key));
      int64_t a = p + err_lo, b = p + err_hi + 1;
      a = a < 0 ? 0 : a;
      b = b > static_cast<int64_t>(n) ? static_cast<int64_t>(n) : b;
      if (a < b) {
        lo = static_cast<uint64_t>(a);
        hi = static_cast<uint64_t>(b);
        pred = static_cast<uint64_t>(p < a ? a : (p >= b ? b - 1 : p));
      }
    }
    return group_window_search<K>(col, n, static_cast<uint32_t>(pred), static_cast<uint32_t>(lo),
                                  static_cast<uint32_t>(hi), x, live);
  }
};

// ---------------------------------------------------------------------------
// Learned index, lines layout (no reference counterpart).  The index keeps its
// own copy of the keys in 128-byte lines.  The model assigns every key a target
// line (monotone in the key); line t holds, in key order, the keys assigned to
// it, after any keys an earlier full line pushed on to it, up to CAP keys; what
// does not fit moves on to line t + 1.  So a key sits in its target line or,
// only behind a run of full lines, a few lines later, and the layout stays in
// global key order.  Each line starts with an 8-byte header {row id of its first
// key, key count}; the row id of a key is header.base + its rank in the line.
//
// Lookup: read the target line; the key is there, or absent, unless every key of
// the line is below the probe and the line is full, in which case the next line
// is read.  A probe below its target line's keys is absent (every key assigned
// to an earlier line is smaller, and no key moves backwards).
//
// Targets:
//   root only   target = umulhi(key - min_key, mult): the keys' range spread
//               linearly over the lines in fixed point (exact integer arithmetic)
//   two-stage   a root model (linear, MLP or CDF spline, staged in LDS) picks a
//               leaf; the leaf maps its key span linearly onto its own run of
//               lines: target = line0 + floor(f32(key - key0) * scale).  Leaves
//               own disjoint, ordered key ranges and line runs, so the target is
//               monotone; the build evaluates the same device code.
// ---------------------------------------------------------------------------
struct LineLeaf {  // 16 bytes
  uint64_t key0;   // first build key routed to the leaf (0 for an empty leaf)
  uint32_t line0;  // first target line of the leaf
  float scale;     // target lines per key unit over the leaf's key span
};
static_assert(sizeof(LineLeaf) == 16, "line leaf is 16 bytes");

// One descriptor serves both targets: a root-only specialisation compiled to one more
// VGPR on gfx950 (73), which costs a wave per SIMD (7 -> 6) and 3 % of the rate.
template <class K>
struct LinesDev {
  static constexpr uint32_t SLOTS = 128 / sizeof(K);  // K-sized slots per line
  static constexpr uint32_t HDR = 8 / sizeof(K);      // slots of the {base, count} header
  static constexpr uint32_t CAP = SLOTS - HDR;        // keys per line: 15 (u64), 30 (u32)
  const K* lines;  // num_lines x SLOTS
  uint64_t num_lines;
  K min_key, max_key;
  uint64_t mult;            // root only: target = umulhi(key - min_key, mult)
  RootModel root;           // two-stage: routes to a leaf
  const LineLeaf* leaves;   // nullptr: root only
  uint32_t num_leaves;

  // The root's parameters (MLP weights, spline knots) are staged in LDS.  Reading them
  // in place instead (a generic pointer that may be global) costs a third of the
  // two-stage lookup rate: flat loads make every memory wait in the kernel conservative.
  __host__ __device__ size_t lds_bytes() const { return leaves ? root.bytes() : 0; }
  __device__ void stage(uint8_t* lds) const {
    if (lds_bytes()) stage_to_lds(lds, root.params(), root.bytes());
  }

  // rp: the root's parameters (LDS during lookups, global memory at build)
  __device__ __forceinline__ uint64_t target(K x, const uint8_t* rp) const {
    const uint64_t off = static_cast<uint64_t>(x - min_key);
    if (!leaves) return __umul64hi(off, mult);
    const LineLeaf f = leaves[route_leaf<K>(x, rp, root, min_key, num_leaves)];
    float p = leaf_dkey(x, f.key0) * f.scale;
    p = fminf(fmaxf(p, 0.0f), 2.0e9f);
    return static_cast<uint64_t>(f.line0) + static_cast<uint64_t>(floorf(p));
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t* lds) const {
    const bool live = active && x >= min_key && x <= max_key;
    const uint64_t t = live ? target(x, lds) : 0;
    return group_line_walk(t, x, live);
  }

  // Groups of kWinG lanes read whole lines (32 bytes per lane, lane 0 holding the
  // header) and serve their probes in kWinG rounds; every open round requests its
  // line before any line is examined, so a wave pays one memory latency per step.
  __device__ __forceinline__ uint32_t group_line_walk(uint64_t my_line, K my_x, bool my_live) const {
    using LP = LineProbe<K>;
    constexpr int G = kWinG;
    constexpr uint32_t KPL = LP::KPL;
    static_assert(LP::LK == SLOTS, "a line probe covers one line");
    const int ln = lane();
    const int gbase = ln & ~(G - 1);
    const int gl = ln & (G - 1);
    const uint64_t all = num_lines * SLOTS;
    bool open = my_live;
    uint64_t line = my_line < num_lines ? my_line : num_lines - 1;
    uint32_t found = kNotFound;
    for (;;) {
      const uint64_t open_mask = ballot(open);
      if (open_mask == 0) break;  // wave-uniform exit
      LP pc[G];
#pragma unroll
      for (int r = 0; r < G; ++r) {
        const uint64_t l = shfl64(line, gbase + r);
        if ((open_mask >> (gbase + r)) & 1ull) pc[r].load(lines, all, static_cast<uint32_t>(0), gl, l);
      }
      uint32_t below = 0, count = 0, base = 0;
      bool eq = false;
#pragma unroll
      for (int r = 0; r < G; ++r) {
        if (!((open_mask >> (gbase + r)) & 1ull)) continue;
        uint64_t w0;
        if constexpr (sizeof(K) == 8) {
          w0 = static_cast<uint64_t>(pc[r].k[0]);
        } else {
          w0 = (static_cast<uint64_t>(pc[r].k[1]) << 32) | static_cast<uint64_t>(pc[r].k[0]);
        }
        const uint64_t hdr = shfl64(w0, gbase);  // group lane 0 holds the header
        const uint32_t c = static_cast<uint32_t>(hdr >> 32);
        const K xr = shfl(my_x, gbase + r);
        uint32_t b = 0;
        bool e = false;
#pragma unroll
        for (uint32_t j = 0; j < KPL; ++j) {
          const uint32_t slot = static_cast<uint32_t>(gl) * KPL + j;
          const bool valid = slot >= HDR && slot < HDR + c;
          b += static_cast<uint32_t>(group_popc<G>(ballot(valid && pc[r].k[j] < xr), gbase));
          e |= group_bits<G>(ballot(valid && pc[r].k[j] == xr), gbase) != 0;
        }
        if (gl == r) {
          below = b;
          eq = e;
          count = c;
          base = static_cast<uint32_t>(hdr);
        }
      }
      if (open) {
        if (eq) {
          found = base + below;
          open = false;
        } else if (below == count && count == CAP && line + 1 < num_lines) {
          ++line;  // a full line below the probe: keys pushed on from it sit in the next line
        } else {
          open = false;
        }
      }
    }
    return found;
  }
};

// ---------------------------------------------------------------------------
// Linear-probing hash table over the column (linear_probing_hashtable,
// src/linear_probing_hashtable.cuh:14-98): capacity 2^ceil(log2(n / 0.5)),
// slot = murmur3(key) & (capacity - 1) with the finaliser of the key's width, step +1,
// empty = all-ones key.
// An entry is 16 bytes {u64 key, u32 row id, pad}.  The reference reads one entry per
// 16-byte load and steps on (ld.global.cs of a longlong2, :62), so a wave waits for its
// longest probe run, one dependent load per slot.  Here groups of kWinG lanes read whole
// 128-byte lines of 8 entries (32 bytes per lane) and serve their probes in kWinG
// rounds, every open round requesting its line before any line is examined: a probe
// resolves within the line of its home slot unless its run passes the line's end.
// Tables below 8 slots take the entry-by-entry walk.
// ---------------------------------------------------------------------------
struct HashEntry {
  unsigned long long key;
  uint32_t value;
  uint32_t pad;
};
static_assert(sizeof(HashEntry) == 16, "16-byte entry");

template <class K>
struct HashDev {
  static constexpr uint32_t kLineEntries = 8;
  static constexpr uint32_t kEntriesPerLane = kLineEntries / kWinG;  // 2
  const HashEntry* table;
  uint64_t mask;  // capacity - 1

  __host__ __device__ size_t lds_bytes() const { return 0; }
  __device__ void stage(uint8_t*) const {}

  __device__ __forceinline__ uint32_t lookup_entries(uint64_t k, bool live) const {
    if (!live) return kNotFound;
    uint64_t s = key_hash<K>(k) & mask;
    for (uint64_t probes = 0; probes <= mask; ++probes) {
      const u32x4 e = *reinterpret_cast<const u32x4*>(table + s);
      const uint64_t ek = (static_cast<uint64_t>(e.y) << 32) | e.x;
      if (ek == k) return e.z;
      if (ek == ~0ull) return kNotFound;
      s = (s + 1) & mask;
    }
    return kNotFound;
  }

  __device__ __forceinline__ uint32_t lookup(K x, bool active, const uint8_t*) const {
    const uint64_t k = static_cast<uint64_t>(x);
    const bool live = active && k != ~0ull;  // all-ones is the empty marker, never a stored key
    if (mask < kLineEntries - 1) return lookup_entries(k, live);  // wave-uniform
    constexpr int G = kWinG;
    const int ln = lane();
    const int gbase = ln & ~(G - 1);
    const int gl = ln & (G - 1);
    uint64_t slot = live ? key_hash<K>(k) & mask : 0;
    uint64_t lines_left = (mask + 1) / kLineEntries;  // a table without an empty slot ends the walk
    bool open = live;
    uint32_t found = kNotFound;
    for (;;) {
      const uint64_t open_mask = ballot(open);
      if (open_mask == 0) break;  // wave-uniform exit
      u32x4 e[G][kEntriesPerLane];
#pragma unroll
      for (int r = 0; r < G; ++r) {
        const uint64_t sr = shfl64(slot, gbase + r);
        if ((open_mask >> (gbase + r)) & 1ull) {
          const u32x4* p = reinterpret_cast<const u32x4*>(table + (sr & ~static_cast<uint64_t>(kLineEntries - 1)) +
                                                          static_cast<uint64_t>(gl) * kEntriesPerLane);
#pragma unroll
          for (uint32_t t = 0; t < kEntriesPerLane; ++t) e[r][t] = p[t];
        }
      }
      bool hit = false, stop = false;
      uint32_t value = 0;
#pragma unroll
      for (int r = 0; r < G; ++r) {
        if (!((open_mask >> (gbase + r)) & 1ull)) continue;
        const uint64_t kr = shfl64(k, gbase + r);
        const uint32_t off = static_cast<uint32_t>(shfl64(slot, gbase + r)) & (kLineEntries - 1);
        uint32_t v = 0;
        bool m = false, em = false;
#pragma unroll
        for (uint32_t t = 0; t < kEntriesPerLane; ++t) {
          const uint32_t j = static_cast<uint32_t>(gl) * kEntriesPerLane + t;  // entry in the line
          const uint64_t ek = (static_cast<uint64_t>(e[r][t].y) << 32) | e[r][t].x;
          const bool at = j >= off;
          if (at && ek == kr) {
            m = true;
            v = e[r][t].z;
          }
          em |= at && ek == ~0ull;
        }
        // a present key comes before every empty slot of its run, so a match decides
        const uint32_t mb = group_bits<G>(ballot(m), gbase);
        const uint32_t eb = group_bits<G>(ballot(em), gbase);
        const uint32_t vr = shfl32(v, gbase + (mb ? __ffs(mb) - 1 : 0));
        if (gl == r) {
          hit = mb != 0;
          stop = hit || eb != 0;
          value = vr;
        }
      }
      if (open) {
        if (hit) {
          found = value;
          open = false;
        } else if (stop || --lines_left == 0) {
          open = false;
        } else {
          slot = ((slot | (kLineEntries - 1)) + 1) & mask;  // the next line's first entry
        }
      }
    }
    return found;
  }
};

}  // namespace gj
