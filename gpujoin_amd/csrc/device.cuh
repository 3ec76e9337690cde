// device.cuh - wave64 helpers for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gj {

constexpr int kWave = 64;
constexpr uint32_t kNotFound = 0xFFFFFFFFu;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int lane() { return static_cast<int>(__lane_id()); }

// Cross-lane read through the LDS crossbar (no LDS storage used).
__device__ __forceinline__ uint32_t shfl32(uint32_t v, int src_lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_lane << 2, static_cast<int>(v)));
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src_lane) {
  const uint32_t lo = shfl32(static_cast<uint32_t>(v), src_lane);
  const uint32_t hi = shfl32(static_cast<uint32_t>(v >> 32), src_lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
template <class T>
__device__ __forceinline__ T shfl(T v, int src_lane) {
  if constexpr (sizeof(T) == 8) {
    return static_cast<T>(shfl64(static_cast<uint64_t>(v), src_lane));
  } else {
    return static_cast<T>(shfl32(static_cast<uint32_t>(v), src_lane));
  }
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Bits of `mask` belonging to the G-lane group that starts at lane gbase.
template <int G>
__device__ __forceinline__ uint32_t group_bits(uint64_t mask, int gbase) {
  if constexpr (G == 64) {
    return 0;  // not used with G == 64 (see group_popc / group_ffs)
  } else {
    return static_cast<uint32_t>((mask >> gbase) & ((1ull << G) - 1ull));
  }
}
template <int G>
__device__ __forceinline__ int group_popc(uint64_t mask, int gbase) {
  if constexpr (G == 64) {
    return __popcll(mask);
  } else {
    return __popc(group_bits<G>(mask, gbase));
  }
}
// index (0-based, within the group) of the lowest set bit, or -1
template <int G>
__device__ __forceinline__ int group_ffs(uint64_t mask, int gbase) {
  if constexpr (G == 64) {
    return mask ? __ffsll(static_cast<unsigned long long>(mask)) - 1 : -1;
  } else {
    const uint32_t b = group_bits<G>(mask, gbase);
    return b ? __ffs(b) - 1 : -1;
  }
}

// Streaming (non-temporal) load of one probe key.
template <class K>
__device__ __forceinline__ K load_stream(const K* p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void store_stream(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// 16-byte load viewed as 16 / sizeof(K) keys.
template <class K>
struct Piece {
  static constexpr int kKeys = 16 / static_cast<int>(sizeof(K));
  K k[kKeys];
};
template <class K>
__device__ __forceinline__ Piece<K> load_piece(const K* p) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  Piece<K> r;
  if constexpr (sizeof(K) == 8) {
    r.k[0] = (static_cast<uint64_t>(v.y) << 32) | v.x;
    r.k[1] = (static_cast<uint64_t>(v.w) << 32) | v.z;
  } else {
    r.k[0] = v.x;
    r.k[1] = v.y;
    r.k[2] = v.z;
    r.k[3] = v.w;
  }
  return r;
}
// Same, for the tail of a column: keys at or beyond `valid` are replaced by `fill`.
template <class K>
__device__ __forceinline__ Piece<K> load_piece_guarded(const K* p, int valid, K fill) {
  Piece<K> r;
  if (valid >= Piece<K>::kKeys) return load_piece(p);
#pragma unroll
  for (int j = 0; j < Piece<K>::kKeys; ++j) r.k[j] = (j < valid) ? p[j] : fill;
  return r;
}

template <class K>
__device__ __forceinline__ constexpr K key_max() {
  return static_cast<K>(~static_cast<K>(0));
}

// murmur3 64-bit finaliser (src/cuda_utils.cuh:85-100).
__host__ __device__ __forceinline__ uint64_t murmur3_64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// murmur3 32-bit finaliser (src/cuda_utils.cuh:65-79).
__host__ __device__ __forceinline__ uint32_t murmur3_32(uint32_t k) {
  k ^= k >> 16;
  k *= 0x85ebca6bu;
  k ^= k >> 13;
  k *= 0xc2b2ae35u;
  k ^= k >> 16;
  return k;
}

// The hash of a K-wide key: the reference hashes with the finaliser of the key's own width
// (murmur3_hash's 32- and 64-bit overloads, src/cuda_utils.cuh:65-100).
template <class K>
__host__ __device__ __forceinline__ uint64_t key_hash(uint64_t k) {
  if constexpr (sizeof(K) == 4) {
    return murmur3_32(static_cast<uint32_t>(k));
  } else {
    return murmur3_64(k);
  }
}

// Inverses of the finalisers (each step undone in reverse order; the multipliers'
// inverses mod 2^64 and 2^32): the hash-table build partitions hashed keys and
// recovers each key from its hash.
__host__ __device__ __forceinline__ uint64_t murmur3_64_inverse(uint64_t h) {
  h ^= h >> 33;
  h *= 0x9cb4b2f8129337dbull;  // inverse of 0xc4ceb9fe1a85ec53
  h ^= h >> 33;
  h *= 0x4f74430c22a54005ull;  // inverse of 0xff51afd7ed558ccd
  h ^= h >> 33;
  return h;
}
__host__ __device__ __forceinline__ uint32_t murmur3_32_inverse(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7ed1b41du;  // inverse of 0xc2b2ae35
  h ^= (h >> 13) ^ (h >> 26);
  h *= 0xa5cb9243u;  // inverse of 0x85ebca6b
  h ^= h >> 16;
  return h;
}
template <class K>
__host__ __device__ __forceinline__ uint64_t key_unhash(uint64_t h) {
  if constexpr (sizeof(K) == 4) {
    return murmur3_32_inverse(static_cast<uint32_t>(h));
  } else {
    return murmur3_64_inverse(h);
  }
}

// Block-wide copy of `count` 16-byte units from global memory into LDS.
__device__ __forceinline__ void stage_to_lds(void* lds, const void* src, uint64_t bytes) {
  const u32x4* s = reinterpret_cast<const u32x4*>(src);
  u32x4* d = reinterpret_cast<u32x4*>(lds);
  const uint64_t units = bytes / 16;
  for (uint64_t i = threadIdx.x; i < units; i += blockDim.x) d[i] = s[i];
}

}  // namespace gj
