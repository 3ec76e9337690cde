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

template <class Dev, class K>
void launch_tuples(const Dev& d, const K* keys, const uint32_t* payload, uint64_t total, uint32_t* out,
                   const gj_lookup_params& p, hipStream_t s, int cu) {
  if (total == 0) return;
  const unsigned block = lookup_block(p);
  const size_t lds = d.lds_bytes();
  const void* fn = reinterpret_cast<const void*>(&lookup_tuples_kernel<Dev, K>);
  uint64_t blocks = p.grid_blocks ? p.grid_blocks : static_cast<uint64_t>(blocks_per_cu(fn, block, lds)) * cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(total, block));
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL((lookup_tuples_kernel<Dev, K>), dim3(static_cast<unsigned>(blocks)), dim3(block), lds, s, d,
                     keys, payload, total, out);
  GJ_HIP_CHECK(hipGetLastError());
}

void build_index(gj_index& ix, hipStream_t s);
template <class K> void build_btree(gj_index& ix, hipStream_t s);
template <class K> void build_extended(gj_index& ix, hipStream_t s);

template <class K>
void lookup_plain(const gj_index& ix, const K* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                  hipStream_t s);
template <class K>
void lookup_tuples(const gj_index& ix, const K* keys, const uint32_t* payload, uint64_t total, uint32_t* out,
                   const gj_lookup_params& p, hipStream_t s);
template <class K>
void dispatch_extended_plain(const gj_index& ix, const K* probes, uint64_t m, uint32_t* out,
                             const gj_lookup_params& p, hipStream_t s);
template <class K>
void dispatch_extended_tuples(const gj_index& ix, const K* keys, const uint32_t* payload, uint64_t total,
                              uint32_t* out, const gj_lookup_params& p, hipStream_t s);

// Bytes of index-private device memory held by the extended kinds.
uint64_t extended_memory(const gj_index& ix);

// Drivers (drivers.hip)
void lookup_dispatch(const gj_index& ix, const void* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                     hipStream_t s);

}  // namespace gj
