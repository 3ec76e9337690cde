// drivers.hip - the lookup drivers ("approaches") behind gj_index_lookup:
// plain, bws (blockwise sorting), partitioning (windowed radix partitioning)
// and hj (hash join: build over the column, then probe).
#include <algorithm>

#include "common.hpp"
#include "lookup_dispatch.hpp"

namespace gj {

namespace {

// Below one fallback key in this many build keys, the plain driver marks and
// re-looks-up fallback probes; above, it runs the B-tree inline.  A wave with a
// mark costs one B-tree lookup, so marking wins while fewer than about one wave
// in six (64 / 400) holds a fallback probe.
constexpr uint64_t kMarkFallbackRatio = 400;

// Row ids back to probe order after a batch of windows: the partitioner recorded where
// each probe went (inv1; with two passes a first-pass position, which inv2 maps to the
// final one) and the tuple lookup wrote row ids in tuple order, so
// out[i] = res[inv2[inv1[i]]].  Each XCD takes one contiguous eighth of the probes, and
// its workgroups sweep it together, one workgroup tile apart, so the windows in flight
// on an XCD (whose results its L2 must hold for the gathers) are one or two.  A position
// outside the arrays (a tuple the partitioner dropped and counted) reads as not found.
constexpr int kUnpartIn = 4;
__global__ __launch_bounds__(256) void unpartition_kernel(const uint32_t* __restrict__ inv1,
                                                          const uint32_t* __restrict__ inv2,
                                                          const uint32_t* __restrict__ res, uint64_t res_n,
                                                          uint64_t m, uint32_t* __restrict__ out) {
  const uint32_t groups = (gridDim.x % 8 == 0) ? 8u : 1u;  // XCDs
  const uint64_t tile = static_cast<uint64_t>(blockDim.x) * kUnpartIn;
  const uint64_t chunk = ((m + groups - 1) / groups + tile - 1) / tile * tile;
  const uint64_t lo = static_cast<uint64_t>(blockIdx.x % groups) * chunk;
  const uint64_t hi = lo + chunk < m ? lo + chunk : m;
  const uint64_t step = static_cast<uint64_t>(gridDim.x / groups) * tile;
  for (uint64_t i0 = lo + static_cast<uint64_t>(blockIdx.x / groups) * tile + threadIdx.x; i0 < hi; i0 += step) {
    uint32_t t[kUnpartIn];
#pragma unroll
    for (int u = 0; u < kUnpartIn; ++u) {
      const uint64_t i = i0 + static_cast<uint64_t>(u) * blockDim.x;
      t[u] = i < hi ? load_stream(inv1 + i) : 0xFFFFFFFFu;
    }
    if (inv2) {
#pragma unroll
      for (int u = 0; u < kUnpartIn; ++u) t[u] = t[u] < m ? inv2[t[u]] : 0xFFFFFFFFu;
    }
    uint32_t r[kUnpartIn];
#pragma unroll
    for (int u = 0; u < kUnpartIn; ++u) r[u] = t[u] < res_n ? res[t[u]] : kNotFound;
#pragma unroll
    for (int u = 0; u < kUnpartIn; ++u) {
      const uint64_t i = i0 + static_cast<uint64_t>(u) * blockDim.x;
      if (i < hi) store_stream(out + i, r[u]);
    }
  }
}

// out[i] = res[inv2[inv1[i]]] for i < m (inv2 optional); res holds res_n entries.
inline void launch_unpartition(const uint32_t* inv1, const uint32_t* inv2, const uint32_t* res, uint64_t res_n,
                               uint64_t m, uint32_t* out, hipStream_t s, int cu) {
  if (m == 0) return;
  uint64_t blocks = std::min<uint64_t>(static_cast<uint64_t>(cu) * 4, ceil_div(m, 256 * kUnpartIn));
  if (blocks >= 8) blocks = blocks / 8 * 8;  // whole XCD rounds for the XCD-major sweep
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL(unpartition_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, inv1, inv2, res, res_n,
                     m, out);
  GJ_HIP_CHECK(hipGetLastError());
}

// Number of significant bits of the largest key.  The reference derives
// ignore_bits = floor(log2(max key)) - radix_bits (src/index_lookup_new.cu:63-67),
// which leaves the top significant bit out of the digit; using the bit width
// keeps the digit on the top radix_bits significant bits.
uint32_t key_bit_width(uint64_t max_key) { return max_key ? 64u - static_cast<uint32_t>(__builtin_clzll(max_key)) : 0u; }

// Windows are partitioned a batch at a time (each window on its own, all windows of the
// batch in one launch per kernel): one window per launch costs about 36 µs of launches
// and small kernels, which bound 2^16-probe windows at 1.8 G lookups/s.  A batch holds
// up to kBatchProbes probes (or one larger window) and kMaxBatchPartitions partitions.
constexpr uint64_t kBatchProbes = 1ull << 24;
constexpr uint64_t kMaxBatchPartitions = 1ull << 18;

// partitioning_approach (src/index_lookup_partitioning.cu:377-513): the probes are cut
// into windows; per window a radix partition on the top key bits, then a lookup over
// the partitioned tuples, with each probe's row id written to its position in probe
// order.  The reference carries each probe's position as the tuple payload and
// scatters the row ids back (tids[payload] = tid, :218-248), a 4-byte store to a random
// position of the window's output that the memory system writes back as a partial line;
// here the partitioner records, in probe order, where it put each probe, the tuple
// lookup writes the row ids in tuple order (full lines), and unpartition_kernel
// gathers them back (2^26 keys, 2^28 probes, fan-out 2^10, 2^22-probe windows: 27.0 G
// lookups/s against 19.5 with the scatter).  The reference keeps two streams with a
// device-wide sync per window pair; here the windows are stream-ordered on the caller's
// stream and the tuple counts never leave the device.
template <class K>
void lookup_partitioned(const gj_index& ix, const K* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                        hipStream_t s) {
  const uint32_t bits = p.radix_bits ? p.radix_bits : 10u;
  GJ_REQUIRE(bits >= 1 && bits <= 16, GJ_ERR_INVALID, "radix_bits must be in [1, 16]");
  const uint32_t passes = p.passes == 2 ? 2u : 1u;
  GJ_REQUIRE(m < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "partitioning needs fewer than 2^32 - 1 probes");
  uint32_t ignore = p.ignore_bits;
  if (ignore == 0xFFFFFFFFu) {
    const uint32_t w = key_bit_width(ix.max_key);
    ignore = w > bits ? w - bits : 0u;
  }
  GJ_REQUIRE(ignore < 8 * sizeof(K), GJ_ERR_INVALID, "ignore_bits must be below the key width");
  if (m == 0) return;
  const uint64_t window = std::min<uint64_t>(p.window ? p.window : (1ull << 24), m);
  const uint64_t fan = 1ull << bits;
  // a column in host memory: one workgroup per CU for the lookup, so each XCD
  // sweeps a narrow key range and the host pages in flight stay within reach of
  // the GPU's address translation (2x on a 2^26-key column)
  gj_lookup_params tp = p;
  if ((ix.col_in_host || ix.params.placement == GJ_PLACE_HOST) && !tp.grid_blocks)
    tp.grid_blocks = static_cast<uint32_t>(ix.cu);
  const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(kBatchProbes / window, kMaxBatchPartitions >> bits));
  const uint64_t batch = std::min<uint64_t>(per * window, m);
  const uint64_t parts = ceil_div(batch, window) * fan;
  const uint64_t bcap = windows_capacity(batch, window, sizeof(K), bits);
  GJ_REQUIRE(bcap < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "window too large");
  const bool two = windows_two_pass(bits, passes);
  // scratch: keys | row ids in tuple order | inv1 | inv2 (two passes) | start (parts + 1) |
  // count (parts) | partitioner scratch
  auto up = [](uint64_t b) { return (b + 255) / 256 * 256; };
  const uint64_t o_res = up(bcap * sizeof(K)), o_inv1 = o_res + up(bcap * 4), o_inv2 = o_inv1 + up(batch * 4),
                 o_start = o_inv2 + (two ? up(batch * 4) : 0), o_count = o_start + up((parts + 1) * 8),
                 o_part = o_count + up(parts * 8),
                 bytes = o_part + up(windows_scratch(batch, window, sizeof(K), bits, passes));
  dbuf& scratch = ix.scratch_for(s).a;
  scratch.reserve(bytes);
  uint8_t* base = scratch.as<uint8_t>();
  K* tkeys = reinterpret_cast<K*>(base);
  uint32_t* res = reinterpret_cast<uint32_t*>(base + o_res);
  uint32_t* inv1 = reinterpret_cast<uint32_t*>(base + o_inv1);
  uint32_t* inv2 = two ? reinterpret_cast<uint32_t*>(base + o_inv2) : nullptr;
  uint64_t* start = reinterpret_cast<uint64_t*>(base + o_start);
  uint64_t* count = reinterpret_cast<uint64_t*>(base + o_count);
  for (uint64_t b0 = 0; b0 < m; b0 += batch) {
    const uint64_t bm = std::min<uint64_t>(batch, m - b0);
    const uint64_t cap_b = windows_capacity(bm, window, sizeof(K), bits);
    partition_windows<K>(probes + b0, bm, window, bits, ignore, passes, tkeys, inv1, inv2, start, count,
                         base + o_part, ix.dropped.as<unsigned long long>(), s);
    with_index_dev<K>(ix, [&](auto d) {
      launch_tuples(d, tkeys, start + ceil_div(bm, window) * fan, cap_b, res, tp, s, ix.cu);
    });
    launch_unpartition(inv1, inv2, res, cap_b, bm, out + b0, s, ix.cu);
  }
}

// hj_approach (src/index_lookup_new.cu:252-302): build a linear-probing table
// over the column, then probe.  The reference builds on the probe keys and
// writes tids[build_tid] = probe key, which validates only for dense keys; this
// driver returns each probe's row id, like every other driver.
template <class K>
void lookup_hash_join(const gj_index& ix, const K* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                      hipStream_t s) {
  uint64_t cap = 0;
  dbuf& table = ix.scratch_for(s).b;
  build_hash_table<K>(ix, table, cap, ix.scratch_for(s).a, s);
  HashDev<K> d{};
  d.table = table.as<const HashEntry>();
  d.mask = cap - 1;
  launch_plain(d, probes, m, out, p, s, ix.cu);
}

}  // namespace

void lookup_dispatch(const gj_index& ix, const void* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                     hipStream_t s) {
  dispatch_key(ix.key_bytes, [&]<class K>() {
    const K* pk = static_cast<const K*>(probes);
    switch (p.approach) {
      case GJ_PLAIN:
        if ((ix.kind == GJ_LEARNED_PLR || ix.kind == GJ_LEARNED_MLP) && ix.lm_layout == GJ_LAYOUT_COLUMN &&
            ix.lm_fallback_any &&
            ix.stats.fallback_keys * kMarkFallbackRatio <= ix.n) {
          // few keys in leaves routed to the B-tree: the learned kernel marks their
          // probes, then one B-tree pass over the marks (the hot kernel stays free of
          // the B-tree's registers); with more, the B-tree runs inline
          launch_plain(ix.learned_dev<K, kLeafMark>(), pk, m, out, p, s, ix.cu);
          launch_marked(ix.btree_dev<K>(), pk, m, out, p, s, ix.cu);
          return;
        }
        with_index_dev<K>(ix, [&](auto d) { launch_plain(d, pk, m, out, p, s, ix.cu); });
        return;
      case GJ_BWS: with_index_dev<K>(ix, [&](auto d) { launch_bws(d, pk, m, out, p, s, ix.cu); }); return;
      case GJ_PARTITIONING: lookup_partitioned<K>(ix, pk, m, out, p, s); return;
      case GJ_HJ: lookup_hash_join<K>(ix, pk, m, out, p, s); return;
    }
    throw error(GJ_ERR_INVALID, "unknown approach");
  });
}

}  // namespace gj
