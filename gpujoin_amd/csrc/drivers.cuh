// drivers.cuh - lookup drivers shared by every index descriptor.
#pragma once

#include "device.cuh"
#include "indexes.cuh"
#include "indexes_ext.cuh"

namespace gj {

// lookup_plain_kernel is the successor of lookup_kernel
// (src/index_lookup_common.cuh:217-245): a grid-stride loop in whole waves
// (the reference rounds its loop bound to a warp multiple for the same
// reason, :235) with 64-bit indices (the reference uses int, which limits it
// to 2^31 probes).
template <class Dev, class K>
__global__ __launch_bounds__(256) void lookup_plain_kernel(Dev d, const K* __restrict__ probes, uint64_t m,
                                                           uint32_t* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  __syncthreads();
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) / kWave;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  for (uint64_t base = wave * kWave; base < m; base += waves * kWave) {
    const uint64_t i = base + static_cast<uint64_t>(lane());
    const bool active = i < m;
    const K x = active ? load_stream(probes + i) : K(0);
    const uint32_t r = d.lookup(x, active, lds);
    if (active) store_stream(out + i, r);
  }
}

// Second pass of the plain driver for learned models with a few leaves routed
// to the B-tree: every probe whose result holds kFallbackMark is looked up again
// with index d.  A wave scans kMarkSlots x 64 results per step (independent,
// coalesced loads) and runs a cooperative lookup only for the slots holding a mark.
constexpr int kMarkSlots = 16;

template <class Dev, class K>
__global__ __launch_bounds__(256) void lookup_marked_kernel(Dev d, const K* __restrict__ probes, uint64_t m,
                                                            uint32_t* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  __syncthreads();
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) / kWave;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t span = static_cast<uint64_t>(kMarkSlots) * kWave;
  for (uint64_t base = wave * span; base < m; base += waves * span) {
    uint32_t v[kMarkSlots];
#pragma unroll
    for (int k = 0; k < kMarkSlots; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k * kWave + lane());
      v[k] = i < m ? out[i] : 0u;
    }
    uint32_t mk = 0;  // bit k: slot k holds a mark
#pragma unroll
    for (int k = 0; k < kMarkSlots; ++k) mk |= (v[k] == kFallbackMark ? 1u : 0u) << k;
    if (ballot(mk != 0) == 0) continue;  // wave-uniform
#pragma unroll 1
    for (int k = 0; k < kMarkSlots; ++k) {
      const bool marked = (mk >> k) & 1u;
      if (ballot(marked) == 0) continue;  // wave-uniform
      const uint64_t i = base + static_cast<uint64_t>(k * kWave + lane());
      const K x = marked ? probes[i] : K(0);
      const uint32_t r = d.lookup(x, marked, lds);
      if (marked) out[i] = r;
    }
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  This maps the dispatch index to a logical index so that each XCD
// receives one contiguous eighth of the logical range.
__device__ __forceinline__ uint32_t xcd_logical_block() {
  const uint32_t nb = gridDim.x;
  if (nb % 8 != 0) return blockIdx.x;
  return (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
}

// Lookup over a partitioned relation (partitioned_lookup_kernel,
// src/index_lookup_partitioning.cu:218-248): tuple t has key keys[t], and its row id
// goes to res[t], in tuple order (full-line stores); the driver then gathers the row ids
// back to probe order.  Padding tuples (key_max) are looked up too, their row ids never
// read.  Each XCD takes one contiguous eighth of the tuples, and the workgroups of an
// XCD sweep it together, one workgroup-wide step apart, so the key range in flight on
// an XCD (hence the index nodes and column lines its L2 must hold) spans a few
// partitions rather than one partition per workgroup.  The tuple count is read from
// *d_total when given (it is known only on the device).
template <class Dev, class K>
__global__ __launch_bounds__(256) void lookup_tuples_kernel(Dev d, const K* __restrict__ keys,
                                                            const uint64_t* __restrict__ d_total, uint64_t total_max,
                                                            uint32_t* __restrict__ res) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  __syncthreads();
  uint64_t total = total_max;
  if (d_total) total = *d_total < total_max ? *d_total : total_max;
  const uint32_t groups = (gridDim.x % 8 == 0) ? 8u : 1u;  // XCDs
  const uint32_t per = gridDim.x / groups;                 // workgroups per XCD
  const uint64_t chunk = ((total + groups - 1) / groups + kWave - 1) / kWave * kWave;
  const uint64_t lo = static_cast<uint64_t>(blockIdx.x % groups) * chunk;
  const uint64_t hi = (lo + chunk < total) ? lo + chunk : total;
  const uint64_t w = threadIdx.x / kWave;
  const uint64_t step = static_cast<uint64_t>(per) * blockDim.x;
  for (uint64_t base = lo + static_cast<uint64_t>(blockIdx.x / groups) * blockDim.x + w * kWave; base < hi;
       base += step) {
    const uint64_t i = base + static_cast<uint64_t>(lane());
    const bool active = i < hi;
    const K x = active ? load_stream(keys + i) : K(0);
    const uint32_t r = d.lookup(x, active, lds);
    if (active) store_stream(res + i, r);
  }
}

// Blockwise-sorting driver (bws_lookup, src/index_lookup_common.cuh:397-549):
// each workgroup loads a tile of TILE probes with their positions, sorts the
// tile by key in LDS (bitonic network; the reference uses CUB's block radix
// sort on key bits [4, max_bits)), then looks the tile up in key order so the
// lanes of a wave walk neighbouring index paths, and scatters the row ids back.
template <class Dev, class K, int TILE>
__global__ __launch_bounds__(256) void lookup_bws_kernel(Dev d, const K* __restrict__ probes, uint64_t m,
                                                         uint32_t* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  const size_t ilds = (d.lds_bytes() + 15) / 16 * 16;
  K* sk = reinterpret_cast<K*>(lds + ilds);
  uint32_t* si = reinterpret_cast<uint32_t*>(lds + ilds + TILE * sizeof(K));
  __syncthreads();
  for (uint64_t tile = static_cast<uint64_t>(blockIdx.x) * TILE; tile < m;
       tile += static_cast<uint64_t>(gridDim.x) * TILE) {
    for (int t = threadIdx.x; t < TILE; t += blockDim.x) {
      const uint64_t i = tile + t;
      sk[t] = i < m ? load_stream(probes + i) : key_max<K>();
      si[t] = i < m ? static_cast<uint32_t>(i - tile) : 0xFFFFFFFFu;
    }
    __syncthreads();
    for (int k = 2; k <= TILE; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < TILE / 2; t += blockDim.x) {
          const int a = 2 * t - (t & (j - 1));  // lower element of the pair (a, a + j)
          const int b = a + j;
          const bool up = (a & k) == 0;
          const K ka = sk[a], kb = sk[b];
          if ((ka > kb) == up) {
            sk[a] = kb;
            sk[b] = ka;
            const uint32_t ia = si[a];
            si[a] = si[b];
            si[b] = ia;
          }
        }
        __syncthreads();
      }
    }
    for (int base = (threadIdx.x / kWave) * kWave; base < TILE; base += blockDim.x) {
      const int t = base + lane();
      const uint32_t idx = si[t];
      const bool active = idx != 0xFFFFFFFFu;
      const uint32_t r = d.lookup(sk[t], active, lds);
      if (active) out[tile + idx] = r;
    }
    __syncthreads();
  }
}

}  // namespace gj
