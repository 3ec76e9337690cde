// common.hpp - error handling, device buffers and small helpers shared by the
// host side of libgj.  Host-only: kernels live in the *.hip / *.cuh files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "gj.h"
#include "gj_runtime.h"

namespace gj {

// Every failure inside the library is a gj::error carrying a gj_status; the
// C entry points translate it into a return code and the thread-local message.
struct error : std::runtime_error {
  int status;
  error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& msg);

#define GJ_HIP_CHECK(expr)                                                                          \
  do {                                                                                              \
    hipError_t gj_e_ = (expr);                                                                      \
    if (gj_e_ != hipSuccess)                                                                        \
      throw ::gj::error(gj_e_ == hipErrorOutOfMemory ? GJ_ERR_NOMEM : GJ_ERR_HIP,                   \
                        std::string(#expr) + " failed: " + hipGetErrorString(gj_e_));               \
  } while (0)

#define GJ_REQUIRE(cond, status, msg)                 \
  do {                                                \
    if (!(cond)) throw ::gj::error((status), (msg));  \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return GJ_OK;
  } catch (const error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return GJ_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return GJ_ERR_INTERNAL;
  } catch (...) {
    set_last_error("unknown exception");
    return GJ_ERR_INTERNAL;
  }
}

// Owning buffer (the role of device_array_wrapper, src/device_array.hpp:132-236): HBM,
// or page-locked host memory mapped into the GPU's address space after to_host().
struct dbuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool host = false;
  dbuf() = default;
  explicit dbuf(size_t b) { alloc(b); }
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  dbuf(dbuf&& o) noexcept : p(o.p), bytes(o.bytes), host(o.host) { o.p = nullptr; o.bytes = 0; o.host = false; }
  dbuf& operator=(dbuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      host = o.host;
      o.p = nullptr;
      o.bytes = 0;
      o.host = false;
    }
    return *this;
  }
  // Moves the contents to page-locked host memory (the index placement of the reference's
  // host allocators, src/index_lookup_config.hpp:52-66); kernels then read it over the host link.
  void to_host() {
    if (!p || host) return;
    void* h = nullptr;
    GJ_HIP_CHECK(hipHostMalloc(&h, bytes, hipHostMallocDefault));
    const hipError_t e = hipMemcpy(h, p, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      (void)hipHostFree(h);
      GJ_HIP_CHECK(e);
    }
    (void)hipFree(p);
    p = h;
    host = true;
  }
  ~dbuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    GJ_HIP_CHECK(hipMalloc(&p, b));
    bytes = b;
  }
  // grow-only reallocation for reusable scratch
  void reserve(size_t b) {
    if (b > bytes) alloc(b);
  }
  void release() {
    if (p) (void)(host ? hipHostFree(p) : hipFree(p));
    p = nullptr;
    bytes = 0;
    host = false;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline hipStream_t S(gj_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint32_t ilog2_floor(uint64_t v) { return 63u - static_cast<uint32_t>(__builtin_clzll(v)); }
inline uint32_t ilog2_ceil(uint64_t v) { return v <= 1 ? 0u : 64u - static_cast<uint32_t>(__builtin_clzll(v - 1)); }

int device_cu_count();

// Key-width dispatch: calls f.template operator()<uint32_t>() or <uint64_t>().
template <class F>
decltype(auto) dispatch_key(int key_bytes, F&& f) {
  if (key_bytes == 4) return f.template operator()<uint32_t>();
  if (key_bytes == 8) return f.template operator()<uint64_t>();
  throw error(GJ_ERR_INVALID, "key_bytes must be 4 or 8");
}

}  // namespace gj
