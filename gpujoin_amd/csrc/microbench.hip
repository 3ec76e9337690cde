// microbench.hip - the random-gather roofline probe, a dependent pointer chase
// and a streaming copy, measured on the same GPU as the lookups.
//
// gather_kernel is the successor of random_access<512>
// (src/random_access.cu:53-66): the reference gathers 4 B through a murmur3-32
// permutation of the loop index from zero-copy host memory; here each slot
// reads its 8-byte probe (streamed), gathers 8 B from an HBM-resident array at
// murmur3_64(probe) & (n - 1) and writes 4 B, so per slot it moves the same
// algorithmic bytes as one index lookup (probe in, indexed key, row id out).
#include "common.hpp"
#include "device.cuh"
#include "lookup_dispatch.hpp"

namespace gj {

template <int ITEMS>
__global__ __launch_bounds__(256) void gather_kernel(const uint64_t* __restrict__ arr, uint64_t mask,
                                                     const uint64_t* __restrict__ probes, uint64_t m,
                                                     uint32_t* __restrict__ out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; base < m;
       base += stride * ITEMS) {
    uint64_t idx[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t i = base + j * stride;
      idx[j] = (i < m) ? (murmur3_64(load_stream(probes + i)) & mask) : 0;
    }
    uint64_t v[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) v[j] = arr[idx[j]];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t i = base + j * stride;
      if (i < m) store_stream(out + i, static_cast<uint32_t>(v[j]));
    }
  }
}

// start may alias out: each lane reads its start before it writes its end.
__global__ void chase_kernel(const uint32_t* next, uint64_t n, uint32_t chains, uint32_t steps,
                             const uint32_t* start, uint32_t* out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= chains) return;
  uint32_t cur = start ? start[c] : static_cast<uint32_t>((static_cast<uint64_t>(c) * 0x9E3779B97F4A7C15ull) % n);
  for (uint32_t s = 0; s < steps; ++s) cur = next[cur];
  out[c] = cur;
}

// Streaming copy: grid stride over 16-byte units, one load and one store per step
// (4.9-5.0 TB/s read + write; four loads in flight before their stores gave 4.6).
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                   uint64_t units) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < units; i += stride)
    store_stream(dst + i, load_stream(src + i));
}

// Streaming read: grid stride over 16-byte units, kStreamUnroll independent loads per
// thread in flight before their reduction (6.4 TB/s).
constexpr int kStreamUnroll = 4;
constexpr int kReadBlock = 1024;

// XOR of every 8-byte word of src into *out: one atomic per workgroup (a few hundred
// workgroups), so the read, not the reduction, sets the time.
__global__ __launch_bounds__(kReadBlock) void read_kernel(const u32x4* __restrict__ src, uint64_t units,
                                                          unsigned long long* __restrict__ out) {
  __shared__ unsigned long long wx[kReadBlock / kWave];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long acc = 0;
  for (; i + (kStreamUnroll - 1) * stride < units; i += kStreamUnroll * stride) {
    u32x4 v[kStreamUnroll];
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) v[u] = load_stream(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u)
      acc ^= ((static_cast<unsigned long long>(v[u].y) << 32) | v[u].x) ^
             ((static_cast<unsigned long long>(v[u].w) << 32) | v[u].z);
  }
  for (; i < units; i += stride) {
    const u32x4 v = load_stream(src + i);
    acc ^= ((static_cast<unsigned long long>(v.y) << 32) | v.x) ^ ((static_cast<unsigned long long>(v.w) << 32) | v.z);
  }
  for (int off = 32; off > 0; off >>= 1) acc ^= __shfl_xor(acc, off, kWave);
  if (lane() == 0) wx[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long x = 0;
    for (int w = 0; w < kReadBlock / kWave; ++w) x ^= wx[w];
    atomicXor(out, x);
  }
}

}  // namespace gj

using namespace gj;

extern "C" {

int gj_gather(const uint64_t* d_array, uint64_t n_pow2, const uint64_t* d_probes, uint64_t m, uint32_t* d_out,
              uint32_t block_size, uint32_t items_per_thread, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_array && d_probes && d_out, GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(n_pow2 && (n_pow2 & (n_pow2 - 1)) == 0, GJ_ERR_INVALID, "n_pow2 must be a power of two");
    if (m == 0) return;
    const unsigned block = block_size ? block_size : 256u;
    GJ_REQUIRE(block % 64 == 0 && block <= 256, GJ_ERR_INVALID, "block_size must be a multiple of 64, <= 256");
    const unsigned items = items_per_thread ? items_per_thread : 4u;
    const hipStream_t s = S(stream);
    const int cu = device_cu_count();
    auto launch = [&](auto kernel) {
      const int nb = blocks_per_cu(reinterpret_cast<const void*>(kernel), block, 0);
      uint64_t blocks = static_cast<uint64_t>(nb) * cu;
      blocks = std::min<uint64_t>(blocks, ceil_div(m, static_cast<uint64_t>(block) * items));
      blocks = std::max<uint64_t>(blocks, 1);
      hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(blocks)), dim3(block), 0, s, d_array, n_pow2 - 1,
                         d_probes, m, d_out);
    };
    switch (items) {
      case 1: launch(gather_kernel<1>); break;
      case 2: launch(gather_kernel<2>); break;
      case 4: launch(gather_kernel<4>); break;
      case 8: launch(gather_kernel<8>); break;
      default: throw error(GJ_ERR_INVALID, "items_per_thread must be 1, 2, 4 or 8");
    }
    GJ_HIP_CHECK(hipGetLastError());
  });
}

int gj_pointer_chase(const uint32_t* d_next, uint64_t n, uint32_t chains, uint32_t steps, const uint32_t* d_start,
                     uint32_t* d_out, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_next && d_out && n > 0 && n <= 0xFFFFFFFFull, GJ_ERR_INVALID, "bad argument");
    if (chains == 0) return;
    hipLaunchKernelGGL(chase_kernel, dim3((chains + 63) / 64), dim3(64), 0, S(stream), d_next, n, chains, steps,
                       d_start, d_out);
    GJ_HIP_CHECK(hipGetLastError());
  });
}

int gj_stream_read(const void* d_src, uint64_t bytes, uint64_t* d_xor, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_src && d_xor, GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(bytes % 16 == 0, GJ_ERR_INVALID, "bytes must be a multiple of 16");
    GJ_REQUIRE((reinterpret_cast<uintptr_t>(d_src) & 15u) == 0, GJ_ERR_INVALID, "buffer must be 16-byte aligned");
    if (bytes == 0) return;
    const uint64_t units = bytes / 16;
    const int cu = device_cu_count();
    const int nb = blocks_per_cu(reinterpret_cast<const void*>(read_kernel), kReadBlock, 0);
    uint64_t blocks = std::min<uint64_t>(static_cast<uint64_t>(std::max(nb, 1)) * cu, ceil_div(units, kReadBlock));
    blocks = std::max<uint64_t>(blocks, 1);
    hipLaunchKernelGGL(read_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kReadBlock), 0, S(stream),
                       static_cast<const u32x4*>(d_src), units, reinterpret_cast<unsigned long long*>(d_xor));
    GJ_HIP_CHECK(hipGetLastError());
  });
}

int gj_stream_copy(const void* d_src, void* d_dst, uint64_t bytes, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_src && d_dst, GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(bytes % 16 == 0, GJ_ERR_INVALID, "bytes must be a multiple of 16");
    GJ_REQUIRE(((reinterpret_cast<uintptr_t>(d_src) | reinterpret_cast<uintptr_t>(d_dst)) & 15u) == 0,
               GJ_ERR_INVALID, "buffers must be 16-byte aligned");
    if (bytes == 0) return;
    const uint64_t units = bytes / 16;
    const int cu = device_cu_count();
    const int nb = blocks_per_cu(reinterpret_cast<const void*>(copy_kernel), 256, 0);
    uint64_t blocks = std::min<uint64_t>(static_cast<uint64_t>(nb) * cu * 2, ceil_div(units, 256));
    blocks = std::max<uint64_t>(blocks, 1);
    hipLaunchKernelGGL(copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, S(stream),
                       static_cast<const u32x4*>(d_src), static_cast<u32x4*>(d_dst), units);
    GJ_HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
