// drivers.cuh - lookup drivers shared by every index descriptor.
//
// lookup_plain_kernel is the successor of lookup_kernel
// (src/index_lookup_common.cuh:217-245): a grid-stride loop in whole waves
// (the reference rounds its loop bound to a warp multiple for the same
// reason, :235) with 64-bit indices (the reference uses int, which limits it
// to 2^31 probes).
#pragma once

#include "device.cuh"
#include "indexes.cuh"

namespace gj {

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

// Lookup over a partitioned relation: tuple t of the window has key keys[t]
// and original probe index payload[t]; the row id is scattered back to the
// original position (partitioned_lookup_kernel, src/index_lookup_partitioning.cu:218-248).
// Tuples are consumed in whole waves over [0, total); padding slots between
// partitions carry key_max and payload 0xFFFFFFFF and are skipped.
template <class Dev, class K>
__global__ __launch_bounds__(256) void lookup_tuples_kernel(Dev d, const K* __restrict__ keys,
                                                            const uint32_t* __restrict__ payload, uint64_t total,
                                                            uint32_t* __restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  __syncthreads();
  const uint64_t waves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) / kWave;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  for (uint64_t base = wave * kWave; base < total; base += waves * kWave) {
    const uint64_t i = base + static_cast<uint64_t>(lane());
    uint32_t dst = 0xFFFFFFFFu;
    K x = K(0);
    if (i < total) {
      dst = load_stream(payload + i);
      x = load_stream(keys + i);
    }
    const bool active = dst != 0xFFFFFFFFu;
    const uint32_t r = d.lookup(x, active, lds);
    if (active) out[dst] = r;
  }
}

}  // namespace gj
