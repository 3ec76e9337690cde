// partition.hip - radix partitioner (gj_partition).
//
// Replaces the fast-interconnects pair gpu_contiguous_prefix_sum_int64 +
// gpu_chunked_laswwc_radix_partition_int64_int64 that run_on_stream launches
// (src/index_lookup_partitioning.cu:300-366; declarations
// src/gpu_radix_partition.cuh:8-45), whose sources are absent.  Semantics kept:
// partition p of key k is (k >> ignore_bits) & (2^radix_bits - 1); every
// partition starts on a 128-byte boundary (the reference pads partitions by
// GPU_CACHE_LINE_SIZE, src/partitioned_relation.hpp:7-10); payload = the
// tuple's position in the input.  Structure of arrays instead of Tuple{key,
// payload}.  gj_partition keeps the input order inside each partition; the
// partitioned lookup driver lets it float (the payload carries the position).
//
// Kernels (no cooperative launch, no device-wide barrier):
//   hist     one block per contiguous chunk, LDS histogram of 2^radix_bits u32 bins
//   colscan  one wave per partition: column prefix over the blocks and the partition total
//   starts   one block: aligned exclusive prefix over partitions -> partition starts
//   scatter  one block per chunk (the same chunks as hist), three variants:
//            write-combining (LDS carry buffers, whole aligned segments; unordered,
//              or stable where two blocks fit per CU)
//            register tiles with per-wave cursors (stable; fan-outs the
//              write-combining buffers do not fit)
//            one store per item at an LDS-atomic slot (unordered, where the
//              write-combining buffers do not fit)
#include <algorithm>

#include "common.hpp"
#include "lookup_dispatch.hpp"

namespace gj {

constexpr int kPartBlock = 256;
constexpr int kPartWaves = kPartBlock / kWave;
constexpr int kPartItems = 16;                        // items per lane per tile
constexpr int kPartTile = kPartBlock * kPartItems;    // 4096 items
constexpr uint32_t kMaxRadixBits = 16;

__host__ __device__ __forceinline__ uint32_t part_align(int key_bytes) { return 128u / static_cast<uint32_t>(key_bytes); }

template <class K>
__device__ __forceinline__ uint32_t digit_of(K k, uint32_t shift, uint32_t mask) {
  return static_cast<uint32_t>(static_cast<uint64_t>(k) >> shift) & mask;
}

// A store that would land past the output's capacity is not made; it is counted
// instead (never for a consistent histogram), so callers can report the loss.
__device__ __forceinline__ void drop(unsigned long long* dropped) {
  if (dropped) atomicAdd(dropped, 1ull);
}

// Segment mode of the histogram and stable scatter kernels (the second pass of a
// fan-out above 2^12): the input is cut into segs segments (segment h holds
// [start[h], start[h] + count[h]), the first pass's partitions), each segment into
// chunks of `chunk` items; block j handles chunk j - chunk0[h] of the segment h with
// chunk0[h] <= j < chunk0[h + 1], and partition p of segment h is output partition
// h * 2^bits + p.  Blocks from chunk0[segs] on have no work.  segs = 0: plain mode,
// block c handles items [c * chunk, (c + 1) * chunk).
struct Seg {
  const uint64_t* chunk0;  // segs + 1 entries
  const uint64_t* start;   // segs entries
  const uint64_t* count;   // segs entries
  uint32_t segs;
};

// This block's item range and output partition base; false: the block has no work.
__device__ __forceinline__ bool seg_range(const Seg& sg, uint64_t c, uint64_t chunk, uint64_t m, uint32_t fan,
                                          uint64_t& lo, uint64_t& hi, uint64_t& pbase) {
  if (!sg.segs) {
    lo = c * chunk;
    hi = std::min<uint64_t>(m, lo + chunk);
    pbase = 0;
    return true;
  }
  if (c >= sg.chunk0[sg.segs]) return false;
  uint32_t a = 0, b = sg.segs;  // the segment h with chunk0[h] <= c < chunk0[h + 1]
  while (b - a > 1) {
    const uint32_t mid = (a + b) / 2;
    if (sg.chunk0[mid] <= c) a = mid; else b = mid;
  }
  lo = sg.start[a] + (c - sg.chunk0[a]) * chunk;
  hi = std::min<uint64_t>(sg.start[a] + sg.count[a], lo + chunk);
  pbase = static_cast<uint64_t>(a) * fan;
  return true;
}

// One block per chunk (256 or 1024 threads), four loads in flight per thread.
template <class K>
__global__ __launch_bounds__(1024) void part_hist_kernel(const K* __restrict__ keys, uint64_t m, uint64_t chunk,
                                                         uint32_t shift, uint32_t bits, uint32_t* __restrict__ hist,
                                                         Seg sg) {
  extern __shared__ uint32_t h[];
  const uint32_t fan = 1u << bits;
  // chunk index; the scatter maps blocks the same way
  const uint64_t c = sg.segs ? static_cast<uint64_t>(blockIdx.x) : xcd_logical_block();
  uint64_t lo, hi, pbase;
  if (!seg_range(sg, c, chunk, m, fan, lo, hi, pbase)) return;  // block-uniform, before any barrier
  for (uint32_t p = threadIdx.x; p < fan; p += blockDim.x) h[p] = 0;
  __syncthreads();
  constexpr int kIn = 4;
  for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += kIn * blockDim.x) {
    K v[kIn];
#pragma unroll
    for (int u = 0; u < kIn; ++u) {
      const uint64_t i = i0 + static_cast<uint64_t>(u) * blockDim.x;
      v[u] = i < hi ? load_stream(keys + i) : K(0);
    }
#pragma unroll
    for (int u = 0; u < kIn; ++u)
      if (i0 + static_cast<uint64_t>(u) * blockDim.x < hi) atomicAdd(&h[digit_of(v[u], shift, fan - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < fan; p += blockDim.x) hist[c * fan + p] = h[p];
}

// hist[b][p] <- sum_{b' < b} hist[b'][p]; count[p] <- total.  One wave per
// partition scans its column 64 rows at a time (wave prefix sum by shuffles).
__global__ __launch_bounds__(256) void part_colscan_kernel(uint32_t* __restrict__ hist, uint32_t blocks, uint32_t fan,
                                                           uint64_t* __restrict__ count) {
  const uint32_t p = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (p >= fan) return;  // wave-uniform
  const int ln = lane();
  uint64_t carry = 0;
  for (uint32_t b0 = 0; b0 < blocks; b0 += kWave) {
    const uint32_t b = b0 + static_cast<uint32_t>(ln);
    const uint32_t v = b < blocks ? hist[static_cast<uint64_t>(b) * fan + p] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, kWave);
      if (ln >= off) x += y;
    }
    if (b < blocks) hist[static_cast<uint64_t>(b) * fan + p] = static_cast<uint32_t>(carry) + x - v;
    carry += __shfl(x, kWave - 1, kWave);
  }
  if (ln == 0) count[p] = carry;
}

// start[p] = align_up(start[p-1] + count[p-1], align); one block of 1024 threads.
// Each thread owns `per` consecutive partitions and reads their counts kScanBatch at a
// time, loads first (a thread's loads do not depend on one another, so a batch is in
// flight together: at fan-out 2^16, 64 counts per thread, one at a time took 115 µs).
constexpr uint32_t kScanBatch = 8;
__global__ __launch_bounds__(1024) void part_scan_starts_kernel(const uint64_t* __restrict__ count, uint32_t fan,
                                                                uint32_t align, uint64_t* __restrict__ start) {
  __shared__ uint64_t s[1024];
  const uint32_t per = (fan + blockDim.x - 1) / blockDim.x;
  const uint32_t p0 = threadIdx.x * per;
  uint64_t local = 0;
  for (uint32_t q = 0; q < per; q += kScanBatch) {
    uint64_t c[kScanBatch];
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) {
      const uint32_t p = p0 + q + u;
      c[u] = (q + u < per && p < fan) ? count[p] : 0;
    }
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) local += (c[u] + align - 1) / align * align;
  }
  s[threadIdx.x] = local;
  __syncthreads();
  for (uint32_t off = 1; off < blockDim.x; off <<= 1) {  // Hillis-Steele inclusive scan
    const uint64_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = s[threadIdx.x] - local;
  for (uint32_t q = 0; q < per; q += kScanBatch) {
    uint64_t c[kScanBatch];
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) {
      const uint32_t p = p0 + q + u;
      c[u] = (q + u < per && p < fan) ? count[p] : 0;
    }
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) {
      const uint32_t p = p0 + q + u;
      if (q + u < per && p < fan) start[p] = run;
      run += (c[u] + align - 1) / align * align;
    }
  }
  if (threadIdx.x == blockDim.x - 1) start[fan] = s[threadIdx.x];  // padded total
}

// The same starts over many partitions (a batch of windows brings up to 2^18, where one
// block took 420 µs): block b owns the kScanSpan partitions [b * kScanSpan, ...).
// part_span_sums_kernel writes each span's padded total to sums[b];
// part_span_starts_kernel adds the totals of the spans before its own and scans its span
// through LDS (loads and stores coalesced).
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanSpan = kScanThreads * kScanBatch;  // 2048 partitions per block
constexpr uint32_t kOneBlockScanMax = 8192;                // up to here the one-block kernel
constexpr uint64_t kScanSumsBytes = 8192;                  // sums of up to 1024 spans (2^21 partitions)

__device__ __forceinline__ uint64_t padded_count(uint64_t c, uint32_t align) { return (c + align - 1) / align * align; }

// Sum over the block of one value per thread; every thread gets the total.
__device__ __forceinline__ uint64_t span_block_sum(uint64_t v, uint64_t* s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t off = kScanThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  const uint64_t t = s[0];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(kScanThreads) void part_span_sums_kernel(const uint64_t* __restrict__ count, uint32_t fan,
                                                                     uint32_t align, uint64_t* __restrict__ sums) {
  __shared__ uint64_t s[kScanThreads];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanSpan;
  uint64_t c[kScanBatch];
#pragma unroll
  for (uint32_t u = 0; u < kScanBatch; ++u) {
    const uint64_t p = base + u * kScanThreads + threadIdx.x;
    c[u] = p < fan ? count[p] : 0;
  }
  uint64_t local = 0;
#pragma unroll
  for (uint32_t u = 0; u < kScanBatch; ++u) local += padded_count(c[u], align);
  const uint64_t t = span_block_sum(local, s);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

__global__ __launch_bounds__(kScanThreads) void part_span_starts_kernel(const uint64_t* __restrict__ count,
                                                                       uint32_t fan, uint32_t align,
                                                                       const uint64_t* __restrict__ sums,
                                                                       uint64_t* __restrict__ start) {
  __shared__ uint64_t v[kScanSpan];
  __shared__ uint64_t s[kScanThreads];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanSpan;
  uint64_t before = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kScanThreads) before += sums[b];
#pragma unroll
  for (uint32_t u = 0; u < kScanBatch; ++u) {
    const uint64_t p = base + u * kScanThreads + threadIdx.x;
    v[u * kScanThreads + threadIdx.x] = p < fan ? padded_count(count[p], align) : 0;
  }
  const uint64_t off = span_block_sum(before, s);  // also orders the stores to v before the reads below
  // thread t scans v[t * kScanBatch, (t + 1) * kScanBatch)
  uint64_t local = 0;
#pragma unroll
  for (uint32_t u = 0; u < kScanBatch; ++u) local += v[threadIdx.x * kScanBatch + u];
  s[threadIdx.x] = local;
  __syncthreads();
  for (uint32_t o = 1; o < kScanThreads; o <<= 1) {  // Hillis-Steele inclusive scan
    const uint64_t y = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  uint64_t run = off + s[threadIdx.x] - local;
#pragma unroll
  for (uint32_t u = 0; u < kScanBatch; ++u) {
    const uint64_t x = v[threadIdx.x * kScanBatch + u];
    v[threadIdx.x * kScanBatch + u] = run;
    run += x;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t u = 0; u < kScanBatch; ++u) {
    const uint64_t p = base + u * kScanThreads + threadIdx.x;
    if (p < fan) start[p] = v[u * kScanThreads + threadIdx.x];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) start[fan] = off + s[kScanThreads - 1];
}

// start[0 .. fan] from count[0 .. fan) (start[fan] = padded total); sums: kScanSumsBytes
// of scratch, used above kOneBlockScanMax partitions.
inline void launch_scan_starts(const uint64_t* count, uint64_t fan, uint32_t align, uint64_t* start, uint64_t* sums,
                               hipStream_t s) {
  if (fan <= kOneBlockScanMax) {
    hipLaunchKernelGGL(part_scan_starts_kernel, dim3(1), dim3(1024), 0, s, count, static_cast<uint32_t>(fan), align,
                       start);
    GJ_HIP_CHECK(hipGetLastError());
    return;
  }
  const uint64_t spans = ceil_div(fan, kScanSpan);
  GJ_REQUIRE(spans * 8 <= kScanSumsBytes, GJ_ERR_INVALID, "too many partitions for the starts scan");
  hipLaunchKernelGGL(part_span_sums_kernel, dim3(static_cast<unsigned>(spans)), dim3(kScanThreads), 0, s, count,
                     static_cast<uint32_t>(fan), align, sums);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(part_span_starts_kernel, dim3(static_cast<unsigned>(spans)), dim3(kScanThreads), 0, s, count,
                     static_cast<uint32_t>(fan), align, sums, start);
  GJ_HIP_CHECK(hipGetLastError());
}

// Marks the padding slots after every partition (fewer than `align` each) with payload
// 0xFFFFFFFF and key_max; one thread per partition, grid stride (a batch of windows has
// up to 2^18 partitions).
template <class K>
__global__ void part_pad_kernel(const uint64_t* __restrict__ start, const uint64_t* __restrict__ count, uint32_t fan,
                                uint32_t align, K* __restrict__ keys, uint32_t* __restrict__ payload) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < fan; p += stride) {
    const uint64_t end = start[p] + count[p];
    const uint64_t stop = start[p] + (count[p] + align - 1) / align * align;
    for (uint64_t i = end; i < stop; ++i) {
      keys[i] = key_max<K>();
      if (payload) payload[i] = 0xFFFFFFFFu;
    }
  }
}

inline unsigned pad_grid(uint64_t fan) { return static_cast<unsigned>(std::max<uint64_t>(1, ceil_div(fan, 256))); }

// Stable scatter.  in_payload == nullptr: payload = payload_base + input position.
// A block walks its chunk in tiles of kPartTile items held in registers (wave w
// owns items [w * 1024, (w + 1) * 1024) of the tile, lane-minor).  Per tile:
//   count   each wave counts its items per partition (LDS atomic adds into the
//           wave's row)
//   cursor  per partition: wave cursors = block cursor + counts of earlier waves
//   scatter each wave writes its items at its cursors (rank among equal digits
//           from ballots) and advances them; no ordering between waves is needed
// so a tile costs two block barriers, and the output order is the input order
// inside every partition.
template <class K>
__global__ __launch_bounds__(kPartBlock) void part_scatter_kernel(
    const K* __restrict__ keys, const uint32_t* __restrict__ in_payload, uint32_t payload_base, uint64_t m,
    uint64_t chunk, uint32_t shift, uint32_t bits, const uint32_t* __restrict__ block_off,
    const uint64_t* __restrict__ start, K* __restrict__ out_keys, uint32_t* __restrict__ out_payload,
    uint32_t* __restrict__ inv, uint64_t cap, unsigned long long* __restrict__ dropped, Seg sg) {
  extern __shared__ uint32_t sm[];
  const uint32_t fan = 1u << bits;
  uint32_t* cnt = sm;                        // [kPartWaves][fan]: counts, then cursors
  uint32_t* brun = sm + kPartWaves * fan;    // [fan]: block cursor (absolute position)
  // neighbouring chunks run on one XCD, so the partial lines that neighbouring
  // chunks write into a partition's run meet in the same L2
  const uint64_t c = sg.segs ? static_cast<uint64_t>(blockIdx.x) : xcd_logical_block();
  uint64_t lo, hi, pbase;
  if (!seg_range(sg, c, chunk, m, fan, lo, hi, pbase)) return;  // block-uniform, before any barrier
  for (uint32_t p = threadIdx.x; p < fan; p += blockDim.x) {
    brun[p] = static_cast<uint32_t>(start[pbase + p]) + block_off[c * fan + p];
#pragma unroll
    for (int w = 0; w < kPartWaves; ++w) cnt[w * fan + p] = 0;
  }
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int ln = lane();
  const uint64_t lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));  // lanes below this one
  uint32_t* mine = cnt + w * fan;
  for (uint64_t tile = lo; tile < hi; tile += kPartTile) {
    const uint64_t wbase = tile + static_cast<uint64_t>(w) * (kPartItems * kWave);
    K k[kPartItems];
    uint32_t pay[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const uint64_t i = wbase + static_cast<uint64_t>(j * kWave + ln);
      const bool valid = i < hi;
      k[j] = valid ? load_stream(keys + i) : K(0);
      pay[j] = valid ? (in_payload ? load_stream(in_payload + i) : payload_base + static_cast<uint32_t>(i)) : 0u;
    }
    // lanes of this wave whose item j has the same digit (invalid items match nothing)
    auto peers_of = [&](int j, uint32_t d) {
      const uint64_t i = wbase + static_cast<uint64_t>(j * kWave + ln);
      uint64_t peers = ballot(i < hi);
      for (uint32_t b = 0; b < bits; ++b) {
        const uint64_t set = ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? set : ~set;
      }
      return (i < hi) ? peers : 0ull;
    };
    // counts need no ranks: one LDS atomic add per item
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const uint64_t i = wbase + static_cast<uint64_t>(j * kWave + ln);
      if (i < hi) atomicAdd(&mine[digit_of(k[j], shift, fan - 1)], 1u);
    }
    __syncthreads();
    // a cursor is written only where the wave has items of the partition: those
    // entries are the ones the wave zeroes after its scatter, so every entry is a
    // zero count again when the next tile starts
    for (uint32_t p = threadIdx.x; p < fan; p += blockDim.x) {
      uint32_t r = brun[p];
#pragma unroll
      for (int v = 0; v < kPartWaves; ++v) {
        const uint32_t c = cnt[v * fan + p];
        if (c != 0) cnt[v * fan + p] = r;
        r += c;
      }
      brun[p] = r;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const uint32_t d = digit_of(k[j], shift, fan - 1);
      const uint64_t peers = peers_of(j, d);
      if (peers) {
        const uint32_t at = mine[d];  // every lane of the group reads the cursor before the leader moves it
        const uint32_t pos = at + static_cast<uint32_t>(__popcll(peers & lt));
        if (pos < cap) {  // always true for a consistent histogram
          out_keys[pos] = k[j];
          if (out_payload) out_payload[pos] = pay[j];
        } else {
          drop(dropped);
        }
        if (inv) inv[wbase + static_cast<uint64_t>(j * kWave + ln)] = pos;
        if ((peers & lt) == 0) mine[d] = at + static_cast<uint32_t>(__popcll(peers));
      }
      __builtin_amdgcn_wave_barrier();
    }
    // cursors back to zero counts for the next tile (this wave's row only)
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const uint32_t d = digit_of(k[j], shift, fan - 1);
      const uint64_t i = wbase + static_cast<uint64_t>(j * kWave + ln);
      if (i < hi) mine[d] = 0;
    }
    __syncthreads();  // the next tile's cursor step reads every wave's row
  }
}

// Unordered scatter for callers that carry each tuple's position in its payload
// (the partitioned lookup driver): an item's slot is an LDS atomic add on the
// block's cursor of its partition, so a tile needs no counts, ranks or barriers.
// The block's items of partition p still fill exactly the run the histogram gave
// the block, [start[p] + block_off[c][p], + hist[c][p]); only their order inside
// it is arbitrary.
template <class K>
__global__ __launch_bounds__(kPartBlock) void part_scatter_any_kernel(
    const K* __restrict__ keys, const uint32_t* __restrict__ in_payload, uint32_t payload_base, uint64_t m,
    uint64_t chunk, uint32_t shift, uint32_t bits, const uint32_t* __restrict__ block_off,
    const uint64_t* __restrict__ start, K* __restrict__ out_keys, uint32_t* __restrict__ out_payload, uint64_t cap,
    unsigned long long* __restrict__ dropped) {
  extern __shared__ uint32_t brun[];  // [fan]: block cursor (absolute position)
  const uint32_t fan = 1u << bits;
  const uint64_t c = xcd_logical_block();
  for (uint32_t p = threadIdx.x; p < fan; p += blockDim.x)
    brun[p] = static_cast<uint32_t>(start[p]) + block_off[c * fan + p];
  __syncthreads();
  const uint64_t lo = c * chunk;
  const uint64_t hi = std::min<uint64_t>(m, lo + chunk);
  for (uint64_t tile = lo; tile < hi; tile += kPartTile) {
    K k[kPartItems];
    uint32_t pay[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const uint64_t i = tile + static_cast<uint64_t>(j * kPartBlock + threadIdx.x);
      const bool valid = i < hi;
      k[j] = valid ? load_stream(keys + i) : K(0);
      pay[j] = valid ? (in_payload ? load_stream(in_payload + i) : payload_base + static_cast<uint32_t>(i)) : 0u;
    }
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const uint64_t i = tile + static_cast<uint64_t>(j * kPartBlock + threadIdx.x);
      if (i < hi) {
        const uint32_t pos = atomicAdd(&brun[digit_of(k[j], shift, fan - 1)], 1u);
        if (pos < cap) {  // always true for a consistent histogram
          out_keys[pos] = k[j];
          out_payload[pos] = pay[j];
        } else {
          drop(dropped);
        }
      }
    }
  }
}

// Write-combining scatter.  A random scatter into 2^bits runs leaves most output
// lines partly written when L2 evicts them; each eviction then writes a partial
// line, and a line is written back several times.  Here a block stages its items
// in LDS and stores only whole aligned segments of L = 2^seg_bits slots (64 or
// 128 bytes of keys), written by L consecutive lanes.
//
// LDS per block: a carry buffer of L slots per partition holding the items of the
// partition's open segment, and the current tile of kWcTile items in partition
// order.  Per tile:
//   count   per partition (unordered: LDS atomic add, whose result is the item's
//           rank in the tile; stable: per-wave rows as in part_scatter_kernel)
//   scan    one block scan of (items, complete segments) per partition gives the
//           tile offset of each partition and the list of segments to store
//   place   items into the tile buffer at tile offset + rank
//   store   each complete segment by L lanes: slots below the partition cursor
//           come from the carry buffer, the others from the tile buffer
//   carry   items past the last complete segment move into the carry buffer
// At the end of the chunk the open segments are stored as they are.  The first
// segment of a block's run is shared with the previous block: slots below the
// run's first position (run0) are never stored by this block.
constexpr int kWcTile = 2048;
// Block size: at the larger fan-outs the carry buffer fills most of the LDS and
// one block runs per CU, so the unordered scatter uses 1024 threads (4 waves per
// SIMD) to keep loads in flight; the stable one uses 256 (its per-wave count rows
// grow with the waves) and runs only where two blocks fit per CU.
constexpr int wc_block(bool stable) { return stable ? 256 : 1024; }
// Pair: the unordered kernel's LDS leaves room for two blocks per CU, and the kernel is
// held to 64 VGPRs (8 waves per SIMD) so that both run.  Measured on 2^24-item windows:
// fan-out 2^8 153 -> 129 us per launch with two blocks against one at 73 VGPRs, although
// the cap spills 32 bytes; where the LDS admits one block only (fan-out 2^10) the cap
// only adds the spills (+13 %), so it is not applied there.
constexpr int wc_min_waves(bool stable, bool pair) { return (!stable && pair) ? 8 : 1; }

__host__ __device__ constexpr uint32_t wc_max_segments(uint32_t fan, uint32_t seg_bits) {
  // a tile completes at most one segment per partition beyond kWcTile / L
  return (static_cast<uint32_t>(kWcTile) >> seg_bits) + (fan < static_cast<uint32_t>(kWcTile) ? fan : kWcTile);
}

__host__ __device__ constexpr size_t wc_lds_bytes(uint32_t fan, uint32_t seg_bits, int key_bytes, bool stable) {
  const size_t item = static_cast<size_t>(key_bytes) + 4;
  return static_cast<size_t>(fan) * (1u << seg_bits) * item      // carry keys + payloads
         + static_cast<size_t>(kWcTile) * item                   // tile keys + payloads
         + static_cast<size_t>(fan) * 4 * 4                      // cur, run0, off, endp
         + static_cast<size_t>(fan) * 4 * (stable ? wc_block(true) / kWave : 1)  // counts
         + (static_cast<size_t>(wc_max_segments(fan, seg_bits)) * 2 + 3) / 4 * 4;  // segment list
}

template <class K, bool Stable, bool Pair>
__global__ __launch_bounds__(wc_block(Stable), wc_min_waves(Stable, Pair)) void part_scatter_wc_kernel(
    const K* __restrict__ keys, const uint32_t* __restrict__ in_payload, uint32_t payload_base, uint64_t m,
    uint64_t chunk, uint32_t shift, uint32_t bits, uint32_t seg_bits, const uint32_t* __restrict__ block_off,
    const uint64_t* __restrict__ start, K* __restrict__ out_keys, uint32_t* __restrict__ out_payload,
    uint32_t* __restrict__ inv, uint64_t cap, unsigned long long* __restrict__ dropped, Seg sg) {
  constexpr int kBlock = wc_block(Stable);
  constexpr int kWaves = kBlock / kWave;
  constexpr int kItems = kWcTile / kBlock;  // items per lane per tile
  constexpr int kRows = Stable ? kWaves : 1;
  extern __shared__ uint64_t wsm_[];
  __shared__ uint32_t wsum[kWaves];
  const uint32_t fan = 1u << bits;
  const uint32_t L = 1u << seg_bits;
  const uint32_t lmask = L - 1;
  K* ck = reinterpret_cast<K*>(wsm_);              // [fan][L] carry keys
  K* sk = ck + static_cast<size_t>(fan) * L;        // [kWcTile] tile keys, partition order
  uint32_t* cp = reinterpret_cast<uint32_t*>(sk + kWcTile);  // [fan][L] carry payloads
  uint32_t* sp = cp + static_cast<size_t>(fan) * L;  // [kWcTile] tile payloads
  uint32_t* cur = sp + kWcTile;                      // [fan] next output position of the block's run
  uint32_t* run0 = cur + fan;                        // [fan] first position of the block's run
  uint32_t* off = run0 + fan;                        // [fan] tile offset | first segment << 16
  uint32_t* endp = off + fan;                        // [fan] cursor after the tile
  uint32_t* cnt = endp + fan;                        // [kRows][fan]
  uint16_t* segp = reinterpret_cast<uint16_t*>(cnt + kRows * fan);  // segment -> partition

  const uint64_t c = sg.segs ? static_cast<uint64_t>(blockIdx.x) : xcd_logical_block();  // as part_hist_kernel
  uint64_t lo, hi, pbase;
  if (!seg_range(sg, c, chunk, m, fan, lo, hi, pbase)) return;  // block-uniform, before any barrier
  for (uint32_t p = threadIdx.x; p < fan; p += blockDim.x) {
    const uint32_t r = static_cast<uint32_t>(start[pbase + p]) + block_off[c * fan + p];
    cur[p] = r;
    run0[p] = r;
#pragma unroll
    for (int v = 0; v < kRows; ++v) cnt[v * fan + p] = 0;
  }
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int ln = lane();
  const uint64_t lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  // partitions owned by this thread in the scan: [p0, p0 + per)
  const uint32_t per = fan > static_cast<uint32_t>(kBlock) ? fan / kBlock : 1u;
  const uint32_t p0 = threadIdx.x * per;
  const uint32_t groups = static_cast<uint32_t>(kBlock) >> seg_bits;  // L lanes per segment store
  const uint32_t g = threadIdx.x >> seg_bits;
  const uint32_t q = threadIdx.x & lmask;

  // The next tile's items are loaded while this tile is counted, placed and stored (the
  // barriers wait for LDS operations only, so the loads stay in flight across them).
  K kn[kItems];
  uint32_t pn[kItems];
  auto load_tile = [&](uint64_t t) {
    const uint64_t b = t + static_cast<uint64_t>(w) * (kItems * kWave);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint64_t i = b + static_cast<uint64_t>(j * kWave + ln);
      const bool valid = i < hi;
      kn[j] = valid ? load_stream(keys + i) : K(0);
      pn[j] = valid ? (in_payload ? load_stream(in_payload + i) : payload_base + static_cast<uint32_t>(i)) : 0u;
    }
  };
  load_tile(lo);
  for (uint64_t tile = lo; tile < hi; tile += kWcTile) {
    const uint64_t wbase = tile + static_cast<uint64_t>(w) * (kItems * kWave);
    K k[kItems];
    uint32_t pay[kItems];
    uint32_t rk[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      k[j] = kn[j];
      pay[j] = pn[j];
    }
    if (tile + kWcTile < hi) load_tile(tile + kWcTile);  // block-uniform
    // count
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint64_t i = wbase + static_cast<uint64_t>(j * kWave + ln);
      rk[j] = 0;
      if (i < hi) {
        const uint32_t d = digit_of(k[j], shift, fan - 1);
        if constexpr (Stable)
          atomicAdd(&cnt[w * fan + d], 1u);
        else
          rk[j] = atomicAdd(&cnt[d], 1u);
      }
    }
    __syncthreads();
    // scan: per partition v = items | complete segments << 16 (both sums stay below 2^16)
    uint32_t local = 0;
    for (uint32_t u = 0; u < per; ++u) {
      const uint32_t p = p0 + u;
      if (p < fan) {
        uint32_t n = 0;
#pragma unroll
        for (int v = 0; v < kRows; ++v) n += cnt[v * fan + p];
        const uint32_t a = cur[p] & ~lmask;
        const uint32_t e = cur[p] + n;
        endp[p] = e;
        local += n | (((e & ~lmask) - a) >> seg_bits) << 16;
      }
    }
    uint32_t x = local;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, kWave);
      if (ln >= o) x += y;
    }
    if (ln == kWave - 1) wsum[w] = x;
    __syncthreads();
    uint32_t run = x - local;
    uint32_t total = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      if (v < w) run += wsum[v];
      total += wsum[v];
    }
    const uint32_t nseg_total = total >> 16;
    for (uint32_t u = 0; u < per; ++u) {
      const uint32_t p = p0 + u;
      if (p < fan) {
        off[p] = run;
        uint32_t n = 0;
        if constexpr (Stable) {
          // wave rows become tile cursors: tile offset + items of earlier waves
          uint32_t r = run & 0xFFFFu;
#pragma unroll
          for (int v = 0; v < kWaves; ++v) {
            const uint32_t t = cnt[v * fan + p];
            cnt[v * fan + p] = r;
            r += t;
          }
          n = r - (run & 0xFFFFu);
        } else {
          n = cnt[p];
        }
        const uint32_t ns = ((endp[p] & ~lmask) - (cur[p] & ~lmask)) >> seg_bits;
        for (uint32_t t = 0; t < ns; ++t) segp[(run >> 16) + t] = static_cast<uint16_t>(p);
        run += n | (ns << 16);
      }
    }
    __syncthreads();
    // place
    uint32_t pos[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint64_t i = wbase + static_cast<uint64_t>(j * kWave + ln);
      const bool valid = i < hi;
      const uint32_t d = digit_of(k[j], shift, fan - 1);
      uint32_t slot = 0;
      if constexpr (Stable) {
        uint64_t peers = ballot(valid);
        for (uint32_t b = 0; b < bits; ++b) {
          const uint64_t set = ballot((d >> b) & 1u);
          peers &= ((d >> b) & 1u) ? set : ~set;
        }
        if (valid) {
          const uint32_t at = cnt[w * fan + d];  // read by every lane of the group before the leader moves it
          slot = at + static_cast<uint32_t>(__popcll(peers & lt));
          if ((peers & lt) == 0) cnt[w * fan + d] = at + static_cast<uint32_t>(__popcll(peers));
        }
        __builtin_amdgcn_wave_barrier();
      } else {
        slot = (off[d] & 0xFFFFu) + rk[j];
      }
      pos[j] = 0xFFFFFFFFu;
      if (valid && slot < static_cast<uint32_t>(kWcTile)) {  // slot < kWcTile for a consistent count
        sk[slot] = k[j];
        sp[slot] = pay[j];
        pos[j] = cur[d] + slot - (off[d] & 0xFFFFu);
      } else if (valid) {
        drop(dropped);
      }
      if (inv && valid) inv[i] = pos[j];
    }
    __syncthreads();
    // store the complete segments, L lanes each
    for (uint32_t sg = g; sg < nseg_total; sg += groups) {
      const uint32_t p = segp[sg];
      const uint32_t o = off[p];
      const uint32_t cp0 = cur[p];
      const uint32_t at = (cp0 & ~lmask) + ((sg - (o >> 16)) << seg_bits) + q;
      if (at >= run0[p] && at >= cap) drop(dropped);
      if (at >= run0[p] && at < cap) {
        K kk;
        uint32_t pp;
        if (at < cp0) {
          kk = ck[p * L + (at & lmask)];
          pp = cp[p * L + (at & lmask)];
        } else {
          const uint32_t t = (o & 0xFFFFu) + (at - cp0);
          kk = sk[t];
          pp = sp[t];
        }
        out_keys[at] = kk;
        if (out_payload) out_payload[at] = pp;
      }
    }
    __syncthreads();
    // carry the items of each open segment; cursors move to the tile's end
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      if (pos[j] != 0xFFFFFFFFu) {
        const uint32_t d = digit_of(k[j], shift, fan - 1);
        if (pos[j] >= (endp[d] & ~lmask)) {
          ck[d * L + (pos[j] & lmask)] = k[j];
          cp[d * L + (pos[j] & lmask)] = pay[j];
        }
      }
    }
    for (uint32_t u = 0; u < per; ++u) {
      const uint32_t p = p0 + u;
      if (p < fan) {
        cur[p] = endp[p];
#pragma unroll
        for (int v = 0; v < kRows; ++v) cnt[v * fan + p] = 0;
      }
    }
    __syncthreads();
  }
  // open segments: the slots of the block's run in [segment start, cursor)
  for (uint32_t p = g; p < fan; p += groups) {
    const uint32_t at = (cur[p] & ~lmask) + q;
    if (at >= run0[p] && at < cur[p] && at >= cap) drop(dropped);
    if (at >= run0[p] && at < cur[p] && at < cap) {
      out_keys[at] = ck[p * L + q];
      if (out_payload) out_payload[at] = cp[p * L + q];
    }
  }
}

// Chunks of the segmented pass: chunk0[h] = chunks of the segments before h (each
// segment cut into ceil(count[h] / chunk) chunks), chunk0[segs] = all chunks.  One block
// of 1024 threads, each owning `per` consecutive segments (a batch of windows brings
// thousands of segments), counts read kScanBatch at a time.
__global__ __launch_bounds__(1024) void seg_chunks_kernel(const uint64_t* __restrict__ count, uint32_t segs,
                                                          uint64_t chunk, uint64_t* __restrict__ chunk0) {
  __shared__ uint64_t s[1024];
  const uint32_t per = (segs + blockDim.x - 1) / blockDim.x;
  const uint32_t h0 = threadIdx.x * per;
  uint64_t local = 0;
  for (uint32_t q = 0; q < per; q += kScanBatch) {
    uint64_t c[kScanBatch];
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) {
      const uint32_t h = h0 + q + u;
      c[u] = (q + u < per && h < segs) ? count[h] : 0;
    }
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) local += (c[u] + chunk - 1) / chunk;
  }
  s[threadIdx.x] = local;
  __syncthreads();
  for (uint32_t off = 1; off < blockDim.x; off <<= 1) {  // Hillis-Steele inclusive scan
    const uint64_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = s[threadIdx.x] - local;
  for (uint32_t q = 0; q < per; q += kScanBatch) {
    uint64_t c[kScanBatch];
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) {
      const uint32_t h = h0 + q + u;
      c[u] = (q + u < per && h < segs) ? count[h] : 0;
    }
#pragma unroll
    for (uint32_t u = 0; u < kScanBatch; ++u) {
      const uint32_t h = h0 + q + u;
      if (q + u < per && h < segs) chunk0[h] = run;
      run += (c[u] + chunk - 1) / chunk;
    }
  }
  if (threadIdx.x == blockDim.x - 1) chunk0[segs] = s[threadIdx.x];
}

// Segments of a batch of windows: segment h = probes [h * window, min((h + 1) * window, m)).
__global__ void seg_windows_kernel(uint64_t* __restrict__ start, uint64_t* __restrict__ count, uint32_t segs,
                                   uint64_t window, uint64_t m) {
  const uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (h >= segs) return;
  start[h] = h * window;
  count[h] = std::min<uint64_t>(window, m - h * window);
}

// Column scan of the segmented pass: for output partition g = h * fan + p, the
// histogram rows of segment h's chunks become exclusive prefixes and count[g] the
// total.  One thread per output partition.  A skewed first pass leaves one segment with
// most of the chunks (Zipf(1.0) probes: about 70 % of a window in segment 0, 5700
// chunks), so the rows are read kColBatch at a time, loads first: the loads do not
// depend on the running sum, and one at a time they took 730 µs per window.
constexpr int kColBatch = 16;
__global__ __launch_bounds__(256) void seg_colscan_kernel(uint32_t* __restrict__ hist, const uint64_t* __restrict__ chunk0,
                                                          uint32_t segs, uint32_t fan, uint64_t* __restrict__ count) {
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= static_cast<uint64_t>(segs) * fan) return;
  const uint64_t h = g / fan, p = g % fan;
  uint64_t run = 0;
  uint64_t j = chunk0[h];
  const uint64_t e = chunk0[h + 1];
  for (; j + kColBatch <= e; j += kColBatch) {
    uint32_t v[kColBatch];
#pragma unroll
    for (int u = 0; u < kColBatch; ++u) v[u] = hist[(j + u) * fan + p];
#pragma unroll
    for (int u = 0; u < kColBatch; ++u) {
      hist[(j + u) * fan + p] = static_cast<uint32_t>(run);
      run += v[u];
    }
  }
  for (; j < e; ++j) {
    const uint32_t v = hist[j * fan + p];
    hist[j * fan + p] = static_cast<uint32_t>(run);
    run += v;
  }
  count[g] = run;
}

namespace {

struct PartPlan {
  uint32_t fan, blocks;
  uint64_t chunk;
};

constexpr uint32_t kMaxBlocksPerCu = 4;
constexpr size_t kMaxLds = 160 * 1024;

// Whole tiles per block, about per_cu blocks per CU; each block's histogram row
// lives in scratch.
PartPlan plan(uint64_t m, uint32_t bits, int cu, uint32_t per_cu) {
  PartPlan p;
  p.fan = 1u << bits;
  const uint64_t want = static_cast<uint64_t>(cu) * per_cu;
  uint64_t chunk = std::max<uint64_t>(kPartTile, ceil_div(m, want));
  chunk = ceil_div(chunk, kPartTile) * kPartTile;
  p.chunk = chunk;
  p.blocks = static_cast<uint32_t>(std::max<uint64_t>(1, ceil_div(m, chunk)));
  return p;
}

// The starts scan's span sums, then the histogram rows for the plan with the most
// blocks any pass uses.
uint64_t hist_bytes(uint64_t m, uint32_t bits, int cu) {
  const PartPlan p = plan(m, bits, cu, kMaxBlocksPerCu);
  return kScanSumsBytes + ceil_div(static_cast<uint64_t>(p.blocks) * p.fan * 4 + 256, 256) * 256;
}

// Segment size (log2 of slots) of the write-combining scatter: the largest with
// segments of at most 128 bytes of keys whose LDS fits the block; 0 selects the
// register-tile scatter.  The unordered scatter runs 1024 threads in one block
// per CU and fits up to fan-out 1024.  The stable one runs 256 threads and pays
// off only with two blocks per CU (fan-out up to 128 for u64 keys); with one it
// is slower than the register-tile scatter.
uint32_t wc_seg_bits(uint32_t bits, int key_bytes, bool stable) {
  const uint32_t fan = 1u << bits;
  const uint32_t sb = key_bytes == 4 ? 5u : 4u;
  if (stable) return wc_lds_bytes(fan, sb, key_bytes, true) <= kMaxLds / 2 ? sb : 0u;
  for (uint32_t b = sb; b >= 2; --b)
    if (wc_lds_bytes(fan, b, key_bytes, false) <= kMaxLds) return b;
  return 0;
}

// Above 2^12 partitions: a first pass on the high seg_high_bits() bits, then the
// segmented pass on the rest.
constexpr uint32_t kMaxSinglePassBits = 12;
uint32_t seg_low_bits(uint32_t bits) { return bits / 2; }
uint32_t seg_high_bits(uint32_t bits) { return bits - seg_low_bits(bits); }
uint64_t seg_max_chunks(uint64_t m, uint32_t bits) {
  return ceil_div(m, kPartTile) + (1ull << seg_high_bits(bits));
}

// Histogram region of a partitioning (the span sums, then either pass's rows).
uint64_t hist_region(uint64_t m, uint32_t bits, int cu) {
  if (bits <= kMaxSinglePassBits) return hist_bytes(m, bits, cu);
  const uint64_t seg_rows = seg_max_chunks(m, bits) * (1ull << seg_low_bits(bits)) * 4;
  return std::max<uint64_t>(hist_bytes(m, seg_high_bits(bits), cu),
                            kScanSumsBytes + ceil_div(seg_rows + 256, 256) * 256);
}

// histograms | (2 passes: first-pass keys and payloads [| segment starts, counts, chunk
// starts]) | partition starts for gj_partition
uint64_t scratch_for(uint64_t m, int key_bytes, uint32_t bits, uint32_t passes, int cu) {
  uint64_t b = hist_region(m, bits, cu);
  if (passes == 2 || bits > kMaxSinglePassBits)  // first-pass keys + payloads, each on a 256-byte boundary
    b += ceil_div((m + 32) * key_bytes, 256) * 256 + ceil_div((m + 32) * 4, 256) * 256;
  if (bits > kMaxSinglePassBits) b += 3 * ((1ull << seg_high_bits(bits)) + 2) * 8 + 256;
  b += ((1ull << bits) + 2) * 8 + 16;
  return b;
}

}  // namespace

// One partition pass with digit (k >> shift) & (2^bits - 1), stable (input order
// inside every partition) or unordered.  `align` pads partition starts (1 =
// unpadded).  Writes start[fan] and count[fan].
template <class K>
void partition_pass(const K* keys, const uint32_t* in_payload, uint32_t payload_base, uint64_t m, uint32_t shift,
                    uint32_t bits, uint32_t align, bool stable, K* out_keys, uint32_t* out_payload, uint64_t* start,
                    uint64_t* count, uint32_t* hist, uint64_t* sums, unsigned long long* dropped, int cu,
                    hipStream_t s) {
  const uint32_t sb = wc_seg_bits(bits, sizeof(K), stable);
  size_t slds = 0;
  uint32_t per_cu = bits > 10 ? 2 : 4;  // the register-tile scatter holds 5 x 4 B per partition
  if (sb) {
    slds = wc_lds_bytes(1u << bits, sb, sizeof(K), stable);
    per_cu = static_cast<uint32_t>(std::min<size_t>(kMaxBlocksPerCu, std::max<size_t>(1, kMaxLds / slds)));
  }
  const PartPlan p = plan(m, bits, cu, per_cu);
  const size_t lds = p.fan * sizeof(uint32_t);
  // with one or two chunks per CU a chunk gets 1024 threads
  hipLaunchKernelGGL(part_hist_kernel<K>, dim3(p.blocks), dim3(per_cu <= 2 ? 1024 : kPartBlock), lds, s, keys, m,
                     p.chunk, shift, bits, hist, Seg{});
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(part_colscan_kernel, dim3((p.fan + 3) / 4), dim3(256), 0, s, hist, p.blocks, p.fan, count);
  GJ_HIP_CHECK(hipGetLastError());
  launch_scan_starts(count, p.fan, align, start, sums, s);
  if (align > 1) {
    hipLaunchKernelGGL(part_pad_kernel<K>, dim3(pad_grid(p.fan)), dim3(256), 0, s, start, count, p.fan, align,
                       out_keys, out_payload);
    GJ_HIP_CHECK(hipGetLastError());
  }
  const uint64_t cap = m + (align > 1 ? static_cast<uint64_t>(p.fan) * align : 0ull);
  if (sb) {
    auto launch = [&](auto kernel) {
      if (slds > 65536)
        GJ_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(slds)));
      hipLaunchKernelGGL(kernel, dim3(p.blocks), dim3(wc_block(stable)), slds, s, keys, in_payload, payload_base, m,
                         p.chunk, shift, bits, sb, hist, start, out_keys, out_payload, nullptr, cap, dropped, Seg{});
      GJ_HIP_CHECK(hipGetLastError());
    };
    if (stable)
      launch(&part_scatter_wc_kernel<K, true, false>);
    else if (per_cu >= 2)
      launch(&part_scatter_wc_kernel<K, false, true>);
    else
      launch(&part_scatter_wc_kernel<K, false, false>);
    return;
  }
  if (!stable) {
    hipLaunchKernelGGL(part_scatter_any_kernel<K>, dim3(p.blocks), dim3(kPartBlock), lds, s, keys, in_payload,
                       payload_base, m, p.chunk, shift, bits, hist, start, out_keys, out_payload, cap, dropped);
    GJ_HIP_CHECK(hipGetLastError());
    return;
  }
  const size_t rlds = static_cast<size_t>(kPartWaves + 1) * p.fan * sizeof(uint32_t);  // <= 80 KiB of the 160
  if (rlds > 65536)
    GJ_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&part_scatter_kernel<K>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(rlds)));
  hipLaunchKernelGGL(part_scatter_kernel<K>, dim3(p.blocks), dim3(kPartBlock), rlds, s, keys, in_payload, payload_base,
                     m, p.chunk, shift, bits, hist, start, out_keys, out_payload, nullptr, cap, dropped, Seg{});
  GJ_HIP_CHECK(hipGetLastError());
}

uint64_t partition_capacity(uint64_t m, int key_bytes, uint32_t bits) {
  return m + (static_cast<uint64_t>(1) << bits) * part_align(key_bytes);
}

uint64_t partition_scratch(uint64_t m, int key_bytes, uint32_t bits, uint32_t passes) {
  return scratch_for(m, key_bytes, bits, passes, device_cu_count());
}

// The second pass above 2^12 partitions: the first pass left segment h (its partition
// on the high bits) at [start1[h], start1[h] + count1[h]) of keys / payload; each segment
// is partitioned on the low `bits` bits (digit (k >> shift) & (2^bits - 1)) into output
// partitions h * 2^bits + p, with 128-byte aligned starts.  Chunks never straddle
// segments, so every histogram row and cursor belongs to one segment.  Stable: tiles of
// kPartTile items through the register-tile scatter.  Unordered: chunks of kSegWcChunk
// items through the write-combining scatter where its buffers fit (the register tiles
// write each wave's few items per partition as partial lines: 52 B of writes per 12 B
// tuple at 2^8 partitions per 2^16-item segment).
constexpr uint64_t kSegWcChunk = 8 * kWcTile;

template <class K>
void partition_segmented(const K* keys, const uint32_t* payload, uint64_t m, uint32_t shift, uint32_t bits,
                         uint32_t segs, const uint64_t* start1, const uint64_t* count1, uint64_t* chunk0,
                         uint32_t align, bool stable, K* out_keys, uint32_t* out_payload, uint32_t* inv,
                         uint64_t* start, uint64_t* count, uint32_t* hist, uint64_t* sums,
                         unsigned long long* dropped, hipStream_t s) {
  const uint32_t fan = 1u << bits;
  const uint64_t total_fan = static_cast<uint64_t>(segs) * fan;
  const uint32_t sb = stable ? 0u : wc_seg_bits(bits, sizeof(K), false);
  const uint64_t chunk = sb ? kSegWcChunk : static_cast<uint64_t>(kPartTile);
  const uint64_t blocks = ceil_div(m, chunk) + segs;  // an upper bound on the chunks
  GJ_REQUIRE(blocks < (1ull << 31), GJ_ERR_INVALID, "too many partition chunks");
  hipLaunchKernelGGL(seg_chunks_kernel, dim3(1), dim3(1024), 0, s, count1, segs, chunk, chunk0);
  GJ_HIP_CHECK(hipGetLastError());
  const Seg sg{chunk0, start1, count1, segs};
  hipLaunchKernelGGL(part_hist_kernel<K>, dim3(static_cast<unsigned>(blocks)), dim3(sb ? 1024 : kPartBlock),
                     fan * sizeof(uint32_t), s, keys, m, chunk, shift, bits, hist, sg);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(seg_colscan_kernel, dim3(static_cast<unsigned>(ceil_div(total_fan, 256))), dim3(256), 0, s, hist,
                     chunk0, segs, fan, count);
  GJ_HIP_CHECK(hipGetLastError());
  launch_scan_starts(count, total_fan, align, start, sums, s);
  if (align > 1) {
    hipLaunchKernelGGL(part_pad_kernel<K>, dim3(pad_grid(total_fan)), dim3(256), 0, s, start, count,
                       static_cast<uint32_t>(total_fan), align, out_keys, out_payload);
    GJ_HIP_CHECK(hipGetLastError());
  }
  const uint64_t cap = m + (align > 1 ? total_fan * align : 0ull);
  if (sb) {
    const size_t slds = wc_lds_bytes(fan, sb, sizeof(K), false);
    auto launch = [&](auto kernel) {
      if (slds > 65536)
        GJ_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(slds)));
      hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(blocks)), dim3(wc_block(false)), slds, s, keys, payload,
                         0u, m, chunk, shift, bits, sb, hist, start, out_keys, out_payload, inv, cap, dropped, sg);
      GJ_HIP_CHECK(hipGetLastError());
    };
    if (kMaxLds / slds >= 2)
      launch(&part_scatter_wc_kernel<K, false, true>);
    else
      launch(&part_scatter_wc_kernel<K, false, false>);
    return;
  }
  const size_t rlds = static_cast<size_t>(kPartWaves + 1) * fan * sizeof(uint32_t);  // <= 80 KiB of the 160
  if (rlds > 65536)
    GJ_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&part_scatter_kernel<K>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(rlds)));
  hipLaunchKernelGGL(part_scatter_kernel<K>, dim3(static_cast<unsigned>(blocks)), dim3(kPartBlock), rlds, s, keys,
                     payload, 0u, m, chunk, shift, bits, hist, start, out_keys, out_payload, inv, cap, dropped, sg);
  GJ_HIP_CHECK(hipGetLastError());
}

namespace {

// A batch of windows in one partitioning: the windows are the segments of the
// segmented pass, so each window is partitioned on its own (output partition
// w * 2^bits + digit, window-major) while the batch takes one launch per kernel.  Two
// passes (or a fan-out above 2^12): a segmented pass per window on the high bits, then a
// segmented pass per (window, high digit) on the low bits.  No payloads: each pass
// records, in input order, where it put every item (inv1 for the first pass, inv2 for
// the second), which the lookup driver follows back to probe order.
uint32_t win_high_bits(uint32_t bits, uint32_t passes) {
  if (bits > kMaxSinglePassBits) return seg_high_bits(bits);
  return (passes == 2 && bits >= 2) ? (bits + 1) / 2 : 0u;  // 0: one pass
}

struct WinLayout {
  uint64_t windows, segs_b;  // windows; segments of the last pass
  uint32_t hb, lb;           // digit bits of the first pass (0: one pass) and of the last
  uint64_t o_tk, o_wstart, o_wcount, o_sa, o_ca, o_chunk0, bytes;
};

WinLayout win_layout(uint64_t m, uint64_t window, int key_bytes, uint32_t bits, uint32_t passes) {
  WinLayout L{};
  auto up = [](uint64_t v) { return ceil_div(v, 256) * 256; };
  L.windows = ceil_div(m, window);
  L.hb = win_high_bits(bits, passes);
  L.lb = bits - L.hb;
  L.segs_b = L.hb ? (L.windows << L.hb) : L.windows;
  const uint64_t tiles = ceil_div(m, static_cast<uint64_t>(kPartTile));
  const uint64_t rows_a = L.hb ? (tiles + L.windows) << L.hb : 0;
  const uint64_t rows_b = (tiles + L.segs_b) << L.lb;
  uint64_t o = kScanSumsBytes + up(std::max(rows_a, rows_b) * 4);  // span sums, histogram rows of either pass
  L.o_tk = o;
  if (L.hb) o += up((m + 32) * key_bytes);
  L.o_wstart = o;
  o += up((L.windows + 2) * 8);
  L.o_wcount = o;
  o += up((L.windows + 2) * 8);
  L.o_sa = o;
  if (L.hb) o += up((L.segs_b + 2) * 8);
  L.o_ca = o;
  if (L.hb) o += up((L.segs_b + 2) * 8);
  L.o_chunk0 = o;
  o += up((L.segs_b + 2) * 8);
  L.bytes = o;
  return L;
}

}  // namespace

uint64_t windows_capacity(uint64_t m, uint64_t window, int key_bytes, uint32_t bits) {
  return m + ceil_div(m, window) * (1ull << bits) * part_align(key_bytes);
}

uint64_t windows_scratch(uint64_t m, uint64_t window, int key_bytes, uint32_t bits, uint32_t passes) {
  return win_layout(m, window, key_bytes, bits, passes).bytes;
}

bool windows_two_pass(uint32_t bits, uint32_t passes) { return win_high_bits(bits, passes) != 0; }

template <class K>
void partition_windows(const K* keys, uint64_t m, uint64_t window, uint32_t bits, uint32_t ignore, uint32_t passes,
                       K* out_keys, uint32_t* inv1, uint32_t* inv2, uint64_t* start, uint64_t* count,
                       uint8_t* scratch, unsigned long long* dropped, hipStream_t s) {
  if (m == 0) return;
  const WinLayout L = win_layout(m, window, sizeof(K), bits, passes);
  GJ_REQUIRE(inv1 && (!L.hb || inv2), GJ_ERR_INVALID, "partition_windows: missing position arrays");
  GJ_REQUIRE((L.segs_b << L.lb) < (1ull << 31), GJ_ERR_INVALID, "too many partitions in a batch of windows");
  auto* sums = reinterpret_cast<uint64_t*>(scratch);
  uint32_t* hist = reinterpret_cast<uint32_t*>(scratch + kScanSumsBytes);
  auto* wstart = reinterpret_cast<uint64_t*>(scratch + L.o_wstart);
  auto* wcount = reinterpret_cast<uint64_t*>(scratch + L.o_wcount);
  auto* chunk0 = reinterpret_cast<uint64_t*>(scratch + L.o_chunk0);
  hipLaunchKernelGGL(seg_windows_kernel, dim3(static_cast<unsigned>(ceil_div(L.windows, 256))), dim3(256), 0, s,
                     wstart, wcount, static_cast<uint32_t>(L.windows), window, m);
  GJ_HIP_CHECK(hipGetLastError());
  const uint32_t align = part_align(sizeof(K));
  if (!L.hb) {
    partition_segmented<K>(keys, nullptr, m, ignore, bits, static_cast<uint32_t>(L.windows), wstart, wcount, chunk0,
                           align, false, out_keys, nullptr, inv1, start, count, hist, sums, dropped, s);
    return;
  }
  K* tk = reinterpret_cast<K*>(scratch + L.o_tk);
  auto* sa = reinterpret_cast<uint64_t*>(scratch + L.o_sa);
  auto* ca = reinterpret_cast<uint64_t*>(scratch + L.o_ca);
  partition_segmented<K>(keys, nullptr, m, ignore + L.lb, L.hb, static_cast<uint32_t>(L.windows), wstart, wcount,
                         chunk0, 1, false, tk, nullptr, inv1, sa, ca, hist, sums, dropped, s);
  partition_segmented<K>(tk, nullptr, m, ignore, L.lb, static_cast<uint32_t>(L.segs_b), sa, ca, chunk0, align, false,
                         out_keys, nullptr, inv2, start, count, hist, sums, dropped, s);
}
template void partition_windows<uint32_t>(const uint32_t*, uint64_t, uint64_t, uint32_t, uint32_t, uint32_t,
                                          uint32_t*, uint32_t*, uint32_t*, uint64_t*, uint64_t*, uint8_t*,
                                          unsigned long long*, hipStream_t);
template void partition_windows<uint64_t>(const uint64_t*, uint64_t, uint64_t, uint32_t, uint32_t, uint32_t,
                                          uint64_t*, uint32_t*, uint32_t*, uint64_t*, uint64_t*, uint8_t*,
                                          unsigned long long*, hipStream_t);

// Full partitioning into `out` (1 or 2 passes; fan-outs above 2^12 always take two, the
// second one segmented); scratch must hold partition_scratch() bytes; start must hold
// 2^bits + 1 entries (start[2^bits] = padded total).
template <class K>
void partition_keys(const K* keys, uint64_t m, uint32_t payload_base, uint32_t bits, uint32_t ignore, uint32_t passes,
                    bool stable, K* out_keys, uint32_t* out_payload, uint64_t* start, uint64_t* count, uint8_t* scratch,
                    unsigned long long* dropped, int cu, hipStream_t s) {
  auto* sums = reinterpret_cast<uint64_t*>(scratch);
  uint32_t* hist = reinterpret_cast<uint32_t*>(scratch + kScanSumsBytes);
  const uint32_t align = part_align(sizeof(K));
  if (bits > kMaxSinglePassBits) {
    const uint32_t hb = seg_high_bits(bits), lb = seg_low_bits(bits);
    const uint32_t segs = 1u << hb;
    uint8_t* tmp = scratch + hist_region(m, bits, cu);
    K* tk = reinterpret_cast<K*>(tmp);
    uint32_t* tp = reinterpret_cast<uint32_t*>(tmp + ceil_div((m + 32) * sizeof(K), 256) * 256);
    uint8_t* small = tmp + ceil_div((m + 32) * sizeof(K), 256) * 256 + ceil_div((m + 32) * 4, 256) * 256;
    auto* start1 = reinterpret_cast<uint64_t*>(small);
    auto* count1 = start1 + segs + 2;
    auto* chunk0 = count1 + segs + 2;
    partition_pass<K>(keys, nullptr, payload_base, m, ignore + lb, hb, 1, stable, tk, tp, start1, count1, hist,
                      sums, dropped, cu, s);
    partition_segmented<K>(tk, tp, m, ignore, lb, segs, start1, count1, chunk0, align, stable, out_keys, out_payload,
                           nullptr, start, count, hist, sums, dropped, s);
    return;
  }
  if (passes == 1 || bits < 2) {
    partition_pass<K>(keys, nullptr, payload_base, m, ignore, bits, align, stable, out_keys, out_payload, start, count,
                      hist, sums, dropped, cu, s);
    return;
  }
  // pass 1: the high half of the digit, unpadded, into scratch; pass 2: the full digit.
  const uint32_t b_hi = (bits + 1) / 2;
  uint8_t* tmp = scratch + hist_bytes(m, bits, cu);
  K* tk = reinterpret_cast<K*>(tmp);
  uint32_t* tp = reinterpret_cast<uint32_t*>(tmp + ceil_div((m + 32) * sizeof(K), 256) * 256);
  partition_pass<K>(keys, nullptr, payload_base, m, ignore + (bits - b_hi), b_hi, 1, stable, tk, tp, start, count,
                    hist, sums, dropped, cu, s);
  partition_pass<K>(tk, tp, 0, m, ignore, bits, align, stable, out_keys, out_payload, start, count, hist, sums,
                    dropped, cu, s);
}
template void partition_keys<uint32_t>(const uint32_t*, uint64_t, uint32_t, uint32_t, uint32_t, uint32_t, bool,
                                       uint32_t*, uint32_t*, uint64_t*, uint64_t*, uint8_t*, unsigned long long*, int,
                                       hipStream_t);
template void partition_keys<uint64_t>(const uint64_t*, uint64_t, uint32_t, uint32_t, uint32_t, uint32_t, bool,
                                       uint64_t*, uint32_t*, uint64_t*, uint64_t*, uint8_t*, unsigned long long*, int,
                                       hipStream_t);

}  // namespace gj

using namespace gj;

extern "C" {

uint64_t gj_partition_capacity(uint64_t m, int key_bytes, uint32_t radix_bits) {
  if ((key_bytes != 4 && key_bytes != 8) || radix_bits > kMaxRadixBits) return 0;
  return partition_capacity(m, key_bytes, radix_bits);
}

uint64_t gj_partition_scratch_bytes(uint64_t m, int key_bytes, const gj_partition_params* p) {
  if (!p || (key_bytes != 4 && key_bytes != 8) || p->radix_bits > kMaxRadixBits) return 0;
  uint64_t r = 0;
  const int st = guarded([&] { r = partition_scratch(m, key_bytes, p->radix_bits, p->passes == 2 ? 2 : 1); });
  return st == GJ_OK ? r : 0;
}

int gj_partition(const void* d_keys, uint64_t m, int key_bytes, const gj_partition_params* params,
                 gj_partitioned* out, void* d_scratch, uint64_t scratch_bytes, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(params && out, GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(key_bytes == 4 || key_bytes == 8, GJ_ERR_INVALID, "key_bytes must be 4 or 8");
    const uint32_t bits = params->radix_bits;
    GJ_REQUIRE(bits >= 1 && bits <= kMaxRadixBits, GJ_ERR_INVALID, "radix_bits must be in [1, 16]");
    GJ_REQUIRE(params->ignore_bits < static_cast<uint32_t>(8 * key_bytes), GJ_ERR_INVALID,
               "ignore_bits must be below the key width");
    GJ_REQUIRE(params->passes == 1 || params->passes == 2, GJ_ERR_INVALID, "passes must be 1 or 2");
    GJ_REQUIRE(m < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "partitioning needs m < 2^32 - 1 (u32 payloads)");
    GJ_REQUIRE(out->d_offsets && out->d_counts, GJ_ERR_INVALID, "null offsets / counts");
    const uint64_t cap = partition_capacity(m, key_bytes, bits);
    GJ_REQUIRE(out->capacity >= cap, GJ_ERR_INVALID, "gj_partitioned.capacity < gj_partition_capacity()");
    GJ_REQUIRE(cap < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "partitioned relation exceeds 2^32 tuples");
    const int cu = device_cu_count();
    const uint64_t need = scratch_for(m, key_bytes, bits, params->passes, cu);
    GJ_REQUIRE(d_scratch && scratch_bytes >= need, GJ_ERR_INVALID, "scratch smaller than gj_partition_scratch_bytes()");
    const hipStream_t s = S(stream);
    if (m == 0) {
      fill_bytes(out->d_offsets, 0, (1ull << bits) * 8, s);
      fill_bytes(out->d_counts, 0, (1ull << bits) * 8, s);
      return;
    }
    GJ_REQUIRE(d_keys && out->d_keys && out->d_payloads, GJ_ERR_INVALID, "null buffer");
    // partition starts (2^bits + 1 entries, the last is the padded total) live at the end of scratch
    uint8_t* sc = static_cast<uint8_t*>(d_scratch);
    uint64_t* start = reinterpret_cast<uint64_t*>(sc + need - ((1ull << bits) + 1) * 8 - 8);
    start = reinterpret_cast<uint64_t*>(reinterpret_cast<uintptr_t>(start) & ~static_cast<uintptr_t>(7));
    dispatch_key(key_bytes, [&]<class K>() {
      partition_keys<K>(static_cast<const K*>(d_keys), m, 0, bits, params->ignore_bits, params->passes, true,
                        static_cast<K*>(out->d_keys), out->d_payloads, start, out->d_counts, sc,
                        reinterpret_cast<unsigned long long*>(out->d_dropped), cu, s);
    });
    GJ_HIP_CHECK(hipMemcpyAsync(out->d_offsets, start, (1ull << bits) * 8, hipMemcpyDeviceToDevice, s));
  });
}

}  // extern "C"
