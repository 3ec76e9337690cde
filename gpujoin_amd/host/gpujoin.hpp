// gpujoin.hpp - C++ host mirror of gpujoin's index interface over the C ABI.
//
// Mirrors the reference's host-side concepts with the same names and meaning,
// so a caller written against them needs only an include change:
//   abstract_index<Key>::construct(h_column, d_column), memory_consumption()   src/indexes.cuh:30-39
//   btree_index / harmonia_index / binary_search_index / radix_spline_index /
//   no_op_index (+ learned_plr_index, learned_mlp_index, hash_index)           src/indexes.cuh:93-413
//   parse_index_type                                                           src/indexes.cpp:5-19
//   abstract_approach::{initialize, run}, plain / bws / partitioning / hj      src/index_lookup.cuh:29-35,
//                                                                              src/index_lookup_new.cu:180-310
//   measuring::measure (3 warm-ups + repetitions, validator, YAML records)     src/measuring.{hpp,cpp}
// Errors surface as std::runtime_error, as in the reference; every device
// operation is a libgj call (gj.h), so there is no CPU fallback anywhere.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gj.h"
#include "gj_runtime.h"

namespace gpujoin {

inline void check(int status, const char* what) {
  if (status != GJ_OK) throw std::runtime_error(std::string(what) + ": " + gj_last_error());
}

template <class T>
struct vector_view {
  const T* ptr = nullptr;
  size_t n = 0;
  vector_view() = default;
  vector_view(const T* p, size_t s) : ptr(p), n(s) {}
  const T* data() const { return ptr; }
  size_t size() const { return n; }
  const T& operator[](size_t i) const { return ptr[i]; }
};

// Owning HBM array (the role of device_array_wrapper, src/device_array.hpp:132-236).
template <class T>
class device_array {
 public:
  device_array() = default;
  explicit device_array(size_t n) : n_(n) { check(gj_malloc(reinterpret_cast<void**>(&p_), sizeof(T) * (n ? n : 1)), "gj_malloc"); }
  device_array(const std::vector<T>& h) : device_array(h.size()) {
    check(gj_memcpy_htod(p_, h.data(), sizeof(T) * h.size(), nullptr), "gj_memcpy_htod");
  }
  device_array(const device_array&) = delete;
  device_array& operator=(const device_array&) = delete;
  device_array(device_array&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  device_array& operator=(device_array&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~device_array() {
    if (p_) gj_free(p_);
  }
  T* data() const { return p_; }
  size_t size() const { return n_; }
  std::vector<T> to_host() const {
    std::vector<T> h(n_);
    check(gj_memcpy_dtoh(h.data(), p_, sizeof(T) * n_, nullptr), "gj_memcpy_dtoh");
    return h;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

template <class Key>
struct abstract_index {
  using key_t = Key;
  virtual ~abstract_index() = default;
  virtual void construct(const vector_view<key_t>& h_column, const key_t* d_column) = 0;
  virtual size_t memory_consumption() const = 0;
  // batched lookup (the device side of cooperative_lookup, driven by an approach)
  virtual void lookup(const key_t* d_keys, uint64_t m, uint32_t* d_tids, const gj_lookup_params& p,
                      gj_stream_t stream) const = 0;
};

// One class per index kind, all backed by gj_index_build / gj_index_lookup.
template <class Key, gj_index_kind Kind>
struct gj_index_t : abstract_index<Key> {
  static constexpr uint32_t invalid_tid = GJ_NOT_FOUND;
  gj_index_params params{};
  gj_index* handle = nullptr;

  gj_index_t() { gj_index_params_default(&params); }
  explicit gj_index_t(const gj_index_params& p) : params(p) {}
  ~gj_index_t() override {
    if (handle) gj_index_destroy(handle);
  }
  void construct(const vector_view<Key>& h_column, const Key* d_column) override {
    if (handle) gj_index_destroy(handle);
    handle = nullptr;
    check(gj_index_build(Kind, h_column.data(), d_column, h_column.size(), sizeof(Key), &params, nullptr, &handle),
          "construct");
  }
  size_t memory_consumption() const override { return handle ? gj_index_memory(handle) : 0; }
  void lookup(const Key* d_keys, uint64_t m, uint32_t* d_tids, const gj_lookup_params& p,
              gj_stream_t stream) const override {
    check(gj_index_lookup(handle, d_keys, m, d_tids, &p, stream), "lookup");
  }
  gj_index_stats stats() const {
    gj_index_stats s{};
    check(gj_index_get_stats(handle, &s), "stats");
    return s;
  }
};

template <class Key> using binary_search_index = gj_index_t<Key, GJ_BINARY_SEARCH>;
template <class Key> using btree_index = gj_index_t<Key, GJ_BTREE>;
template <class Key> using harmonia_index = gj_index_t<Key, GJ_HARMONIA>;
template <class Key> using radix_spline_index = gj_index_t<Key, GJ_RADIX_SPLINE>;
template <class Key> using learned_plr_index = gj_index_t<Key, GJ_LEARNED_PLR>;
template <class Key> using learned_mlp_index = gj_index_t<Key, GJ_LEARNED_MLP>;
template <class Key> using hash_index = gj_index_t<Key, GJ_HASH>;
template <class Key> using no_op_index = gj_index_t<Key, GJ_NO_OP>;

inline gj_index_kind parse_index_type(const std::string& name) {
  gj_index_kind k;
  if (gj_parse_index_type(name.c_str(), &k) != GJ_OK) throw std::runtime_error("unknown index type");
  return k;
}

// build_index<Key, IndexT> (src/index_lookup_common.cuh:208-215), by name.
template <class Key>
std::unique_ptr<abstract_index<Key>> build_index(const std::string& name, const std::vector<Key>& h_keys,
                                                 const Key* d_keys, const gj_index_params* p = nullptr) {
  std::unique_ptr<abstract_index<Key>> ix;
  gj_index_params params;
  gj_index_params_default(&params);
  if (p) params = *p;
  switch (parse_index_type(name)) {
    case GJ_BINARY_SEARCH: ix = std::make_unique<binary_search_index<Key>>(params); break;
    case GJ_BTREE: ix = std::make_unique<btree_index<Key>>(params); break;
    case GJ_HARMONIA: ix = std::make_unique<harmonia_index<Key>>(params); break;
    case GJ_RADIX_SPLINE: ix = std::make_unique<radix_spline_index<Key>>(params); break;
    case GJ_LEARNED_PLR: ix = std::make_unique<learned_plr_index<Key>>(params); break;
    case GJ_LEARNED_MLP: ix = std::make_unique<learned_mlp_index<Key>>(params); break;
    case GJ_HASH: ix = std::make_unique<hash_index<Key>>(params); break;
    case GJ_NO_OP: ix = std::make_unique<no_op_index<Key>>(params); break;
  }
  ix->construct(vector_view<Key>(h_keys.data(), h_keys.size()), d_keys);
  return ix;
}

// ---- measuring (src/measuring.hpp:10-76, src/measuring.cpp:33-63) ----------
namespace measuring {

static unsigned warm_up_rounds = 3;

struct measuring_settings {
  std::string dest_file;
  unsigned repetitions = 10;
  bool stdout_only = true;
};

inline measuring_settings& get_settings() {
  static measuring_settings s;
  return s;
}

struct experiment_description {
  std::string name;
  std::string approach;
  std::vector<std::pair<std::string, std::string>> other;
};

struct measurement {
  std::vector<std::chrono::high_resolution_clock::time_point> timestamps;
  double device_ms = 0.0;  // event time on the lookup stream
};

inline void record_timestamp(measurement& m) { m.timestamps.push_back(std::chrono::high_resolution_clock::now()); }

inline void write_out_measurement(const experiment_description& d, const measurement& m,
                                  const std::vector<std::pair<std::string, std::string>>& extra = {}) {
  std::ostringstream dt;
  dt << "[";
  for (size_t i = 1; i < m.timestamps.size(); ++i) {
    const double ms =
        std::chrono::duration_cast<std::chrono::microseconds>(m.timestamps[i] - m.timestamps.front()).count() / 1000.;
    dt << (i > 1 ? ", " : "") << ms;
  }
  dt << "]";
  const uint64_t seconds =
      std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now().time_since_epoch()).count();
  std::ostringstream e;
  e << "-\n  name: " << d.name << "\n  approach: " << d.approach << "\n  dt_prefix_sum: " << dt.str()
    << "\n  timestamp: !int64 " << seconds << "\n";
  for (const auto& kv : d.other) e << "  " << kv.first << ": " << kv.second << "\n";
  for (const auto& kv : extra) e << "  " << kv.first << ": " << kv.second << "\n";
  if (get_settings().stdout_only) {
    std::cout << e.str() << std::endl;
    return;
  }
  std::ofstream out(get_settings().dest_file, std::ios_base::app);
  out << e.str();
}

// warm-ups, then repetitions, each validated; invalid measurements are dropped
template <class Func, class Validator, class Extra>
void measure(const experiment_description& d, Func func, Validator validator, Extra extra) {
  for (unsigned i = 0; i < warm_up_rounds; ++i) {
    measurement m;
    func(m);
  }
  for (unsigned i = 0; i < get_settings().repetitions; ++i) {
    measurement m;
    record_timestamp(m);
    func(m);
    record_timestamp(m);
    if (validator()) {
      write_out_measurement(d, m, extra(m));
    } else {
      std::cerr << "invalid measurement" << std::endl;
    }
  }
}

}  // namespace measuring

// ---- approaches (src/index_lookup.cuh:11-35, src/index_lookup_new.cu:180-310) ----

template <class Key>
struct query_data {
  std::unique_ptr<abstract_index<Key>> index_structure;
  std::vector<Key> indexed, lookup_keys;
  device_array<Key> d_indexed, d_lookup_keys;
  device_array<uint32_t> d_tids;
  unsigned dataset_max_bits = 0;
  gj_stream_t stream = nullptr;

  // tid < n and indexed[tid] == key for every lookup (validate_results, src/index_lookup_new.cu:108-145)
  bool validate_results() const {
    const std::vector<uint32_t> tids = d_tids.to_host();
    for (size_t i = 0; i < lookup_keys.size(); ++i) {
      if (tids[i] >= indexed.size()) {
        std::cerr << "invalid tid: " << tids[i] << ", at " << i << " from " << lookup_keys.size() << std::endl;
        return false;
      }
      if (indexed[tids[i]] != lookup_keys[i]) {
        std::cerr << "mismatch at " << i << std::endl;
        return false;
      }
    }
    return true;
  }
};

template <class Key>
struct abstract_approach {
  gj_lookup_params params{};
  explicit abstract_approach(gj_approach a) {
    gj_lookup_params_default(&params);
    params.approach = a;
  }
  virtual ~abstract_approach() = default;
  virtual void initialize(query_data<Key>&) {}
  virtual void run(query_data<Key>& d, measuring::measurement& m) {
    gj_event_t e0, e1;
    check(gj_event_create(&e0), "event");
    check(gj_event_create(&e1), "event");
    check(gj_event_record(e0, d.stream), "event");
    d.index_structure->lookup(d.d_lookup_keys.data(), d.lookup_keys.size(), d.d_tids.data(), params, d.stream);
    check(gj_event_record(e1, d.stream), "event");
    float ms = 0;
    check(gj_event_elapsed_ms(e0, e1, &ms), "event");
    m.device_ms = ms;
    gj_event_destroy(e0);
    gj_event_destroy(e1);
  }
};

template <class Key> struct plain_approach : abstract_approach<Key> { plain_approach() : abstract_approach<Key>(GJ_PLAIN) {} };
template <class Key> struct blockwise_sorting_approach : abstract_approach<Key> { blockwise_sorting_approach() : abstract_approach<Key>(GJ_BWS) {} };
template <class Key> struct partitioning_approach : abstract_approach<Key> { partitioning_approach() : abstract_approach<Key>(GJ_PARTITIONING) {} };
template <class Key> struct hj_approach : abstract_approach<Key> { hj_approach() : abstract_approach<Key>(GJ_HJ) {} };

template <class Key>
std::unique_ptr<abstract_approach<Key>> make_approach(const std::string& name) {
  if (name == "plain") return std::make_unique<plain_approach<Key>>();
  if (name == "bws") return std::make_unique<blockwise_sorting_approach<Key>>();
  if (name == "partitioning") return std::make_unique<partitioning_approach<Key>>();
  if (name == "hj") return std::make_unique<hj_approach<Key>>();
  throw std::runtime_error("unknown approach: " + name);
}

}  // namespace gpujoin
