// runtime.hip - device memory, streams, events, the fill kernel and the
// thread-local error message of libgj.
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>

#include "common.hpp"
#include "device.cuh"
#include "lookup_dispatch.hpp"

namespace gj {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

int device_cu_count() {
  int dev = 0;
  GJ_HIP_CHECK(hipGetDevice(&dev));
  int cu = 0;
  GJ_HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  return cu > 0 ? cu : 1;
}

int blocks_per_cu(const void* fn, unsigned block, size_t lds) {
  static std::mutex mu;
  static std::unordered_map<std::string, int> cache;
  int dev = 0;
  GJ_HIP_CHECK(hipGetDevice(&dev));
  char key[96];
  std::snprintf(key, sizeof(key), "%p/%u/%zu/%d", fn, block, lds, dev);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int nb = 0;
  GJ_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, static_cast<int>(block), lds));
  if (nb < 1) nb = 1;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = nb;
  return nb;
}

__global__ void fill_kernel(uint8_t* __restrict__ d, uint8_t v, uint64_t bytes) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (reinterpret_cast<uintptr_t>(d) & 15u) {  // unaligned target: byte stores
    for (uint64_t i = i0; i < bytes; i += stride) d[i] = v;
    return;
  }
  const uint64_t words = bytes / 16;
  const uint32_t w = 0x01010101u * v;
  const u32x4 pattern = {w, w, w, w};
  u32x4* d16 = reinterpret_cast<u32x4*>(d);
  for (uint64_t i = i0; i < words; i += stride) d16[i] = pattern;
  for (uint64_t i = words * 16 + i0; i < bytes; i += stride) d[i] = v;
}

void fill_bytes(void* d, int value, uint64_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  uint64_t blocks = ceil_div(bytes / 16 + 1, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, static_cast<uint8_t*>(d),
                     static_cast<uint8_t>(value), bytes);
  GJ_HIP_CHECK(hipGetLastError());
}

}  // namespace gj

using namespace gj;

struct gj_event_s {
  hipEvent_t ev;
};

extern "C" {

const char* gj_last_error(void) { return g_last_error.c_str(); }
const char* gj_version(void) { return "gpujoin-amd 0.1 (gfx950)"; }

int gj_device_count(int* count) {
  return guarded([&] {
    GJ_REQUIRE(count, GJ_ERR_INVALID, "count is null");
    GJ_HIP_CHECK(hipGetDeviceCount(count));
  });
}

int gj_device_info(char* name, size_t len, int* compute_units, int* gcn_arch_major) {
  return guarded([&] {
    int dev = 0;
    GJ_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GJ_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    if (name && len) {
      std::snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (gcn_arch_major) *gcn_arch_major = prop.major;
  });
}

int gj_malloc(void** d_ptr, uint64_t bytes) {
  return guarded([&] {
    GJ_REQUIRE(d_ptr, GJ_ERR_INVALID, "d_ptr is null");
    *d_ptr = nullptr;
    if (bytes == 0) return;
    GJ_HIP_CHECK(hipMalloc(d_ptr, bytes));
  });
}

int gj_free(void* d_ptr) {
  return guarded([&] {
    if (d_ptr) GJ_HIP_CHECK(hipFree(d_ptr));
  });
}

int gj_host_malloc(void** h_ptr, uint64_t bytes) {
  return guarded([&] {
    GJ_REQUIRE(h_ptr, GJ_ERR_INVALID, "h_ptr is null");
    *h_ptr = nullptr;
    if (bytes == 0) return;
    GJ_HIP_CHECK(hipHostMalloc(h_ptr, bytes, hipHostMallocMapped | hipHostMallocPortable));
    // kernels use the host address directly: require the GPU mapping at that address
    void* d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, *h_ptr, 0);
    if (e != hipSuccess || d != *h_ptr) {
      (void)hipHostFree(*h_ptr);
      *h_ptr = nullptr;
      GJ_REQUIRE(false, GJ_ERR_UNSUPPORTED, "page-locked host memory is not mapped at the same address on the GPU");
    }
  });
}

int gj_host_free(void* h_ptr) {
  return guarded([&] {
    if (h_ptr) GJ_HIP_CHECK(hipHostFree(h_ptr));
  });
}

int gj_memcpy_htod(void* d_dst, const void* h_src, uint64_t bytes, gj_stream_t stream) {
  return guarded([&] {
    if (bytes == 0) return;
    GJ_REQUIRE(d_dst && h_src, GJ_ERR_INVALID, "null pointer");
    GJ_HIP_CHECK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, S(stream)));
    GJ_HIP_CHECK(hipStreamSynchronize(S(stream)));
  });
}

int gj_memcpy_dtoh(void* h_dst, const void* d_src, uint64_t bytes, gj_stream_t stream) {
  return guarded([&] {
    if (bytes == 0) return;
    GJ_REQUIRE(h_dst && d_src, GJ_ERR_INVALID, "null pointer");
    GJ_HIP_CHECK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, S(stream)));
    GJ_HIP_CHECK(hipStreamSynchronize(S(stream)));
  });
}

int gj_memcpy_dtod(void* d_dst, const void* d_src, uint64_t bytes, gj_stream_t stream) {
  return guarded([&] {
    if (bytes == 0) return;
    GJ_REQUIRE(d_dst && d_src, GJ_ERR_INVALID, "null pointer");
    GJ_HIP_CHECK(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, S(stream)));
  });
}

int gj_memset(void* d_dst, int value, uint64_t bytes, gj_stream_t stream) {
  return guarded([&] {
    if (bytes == 0) return;
    GJ_REQUIRE(d_dst, GJ_ERR_INVALID, "null pointer");
    fill_bytes(d_dst, value, bytes, S(stream));
  });
}

int gj_stream_create(gj_stream_t* stream) {
  return guarded([&] {
    GJ_REQUIRE(stream, GJ_ERR_INVALID, "stream is null");
    hipStream_t s;
    GJ_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<gj_stream_t>(s);
  });
}

int gj_stream_destroy(gj_stream_t stream) {
  return guarded([&] {
    if (stream) GJ_HIP_CHECK(hipStreamDestroy(S(stream)));
  });
}

int gj_stream_synchronize(gj_stream_t stream) {
  return guarded([&] { GJ_HIP_CHECK(hipStreamSynchronize(S(stream))); });
}

int gj_device_synchronize(void) {
  return guarded([&] { GJ_HIP_CHECK(hipDeviceSynchronize()); });
}

int gj_event_create(gj_event_t* ev) {
  return guarded([&] {
    GJ_REQUIRE(ev, GJ_ERR_INVALID, "ev is null");
    auto* e = new gj_event_s;
    hipError_t r = hipEventCreate(&e->ev);
    if (r != hipSuccess) {
      delete e;
      GJ_HIP_CHECK(r);
    }
    *ev = e;
  });
}

int gj_event_destroy(gj_event_t ev) {
  return guarded([&] {
    if (!ev) return;
    hipError_t r = hipEventDestroy(ev->ev);
    delete ev;
    GJ_HIP_CHECK(r);
  });
}

int gj_event_record(gj_event_t ev, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(ev, GJ_ERR_INVALID, "ev is null");
    GJ_HIP_CHECK(hipEventRecord(ev->ev, S(stream)));
  });
}

int gj_event_elapsed_ms(gj_event_t start, gj_event_t stop, float* ms) {
  return guarded([&] {
    GJ_REQUIRE(start && stop && ms, GJ_ERR_INVALID, "null argument");
    GJ_HIP_CHECK(hipEventSynchronize(stop->ev));
    GJ_HIP_CHECK(hipEventElapsedTime(ms, start->ev, stop->ev));
  });
}

}  // extern "C"
