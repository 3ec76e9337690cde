// tpch.hpp - TPC-H .tbl loading for query 14 (lineitem, part), host side.
//
// Reads the dbgen text tables that the reference loads in load_tables
// (src/tpch/common.cpp:115-216, parse<> src/tpch/parser.hpp:499-578), keeping
// only the columns Q14 touches:
//   lineitem  l_partkey (field 1), l_extendedprice (5), l_discount (6), l_shipdate (10)
//   part      p_partkey (field 0), p_type (4) -> a PROMO flag
// Value conventions follow the reference's input parsers:
//   numeric(15,2)  hundredths as int64 (input_parser<numeric>, src/tpch/parser.hpp:99-119);
//                  a fraction with fewer than two digits is scaled (the reference
//                  adds the digits as written; dbgen always writes two)
//   date           Julian day number (input_parser<date>::to_julian_day,
//                  src/tpch/parser.hpp:124-129), so 1995-09-01 is 2449962
// The file is memory-mapped and split into one byte range per OpenMP thread at
// line boundaries; the per-thread columns are concatenated in file order.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace gpujoin::tpch {

constexpr uint32_t to_julian_day(uint32_t day, uint32_t month, uint32_t year) {
  const uint32_t a = (14 - month) / 12;
  const uint32_t y = year + 4800 - a;
  const uint32_t m = month + 12 * a - 3;
  return day + (153 * m + 2) / 5 + y * 365 + y / 4 - y / 100 + y / 400 - 32045;
}

// the Q14 predicate window [1995-09-01, 1995-10-01) (src/tpch/config.hpp:36-37)
constexpr uint32_t q14_lower_shipdate = to_julian_day(1, 9, 1995);
constexpr uint32_t q14_upper_shipdate = to_julian_day(1, 10, 1995);
static_assert(q14_lower_shipdate == 2449962 && q14_upper_shipdate == 2449992);

struct q14_lineitem {
  std::vector<int32_t> partkey;
  std::vector<int32_t> shipdate;       // Julian day
  std::vector<int64_t> extendedprice;  // hundredths
  std::vector<int8_t> discount;        // hundredths (0.05 -> 5)
  size_t size() const { return partkey.size(); }
};

struct q14_part {
  std::vector<uint32_t> partkey;  // ascending
  std::vector<uint8_t> promo;     // p_type starts with "PROMO"
  size_t size() const { return partkey.size(); }
};

class mapped_file {
 public:
  explicit mapped_file(const std::string& path) {
    fd_ = ::open(path.c_str(), O_RDONLY);
    if (fd_ < 0) throw std::runtime_error("cannot open " + path);
    struct stat st {};
    if (::fstat(fd_, &st) != 0) {
      ::close(fd_);
      throw std::runtime_error("cannot stat " + path);
    }
    size_ = static_cast<size_t>(st.st_size);
    if (size_) {
      void* p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) {
        ::close(fd_);
        throw std::runtime_error("cannot map " + path);
      }
      data_ = static_cast<const char*>(p);
    }
  }
  mapped_file(const mapped_file&) = delete;
  mapped_file& operator=(const mapped_file&) = delete;
  ~mapped_file() {
    if (data_) ::munmap(const_cast<char*>(data_), size_);
    if (fd_ >= 0) ::close(fd_);
  }
  const char* data() const { return data_; }
  size_t size() const { return size_; }

 private:
  int fd_ = -1;
  const char* data_ = nullptr;
  size_t size_ = 0;
};

// ---- field parsers: [b, e) is the field text without the separator ----
inline bool parse_uint(const char* b, const char* e, uint64_t& v) {
  if (b == e) return false;
  uint64_t x = 0;
  for (; b != e; ++b) {
    if (*b < '0' || *b > '9') return false;
    x = x * 10 + static_cast<uint64_t>(*b - '0');
  }
  v = x;
  return true;
}

inline bool parse_numeric2(const char* b, const char* e, int64_t& v) {
  bool neg = false;
  if (b != e && *b == '-') {
    neg = true;
    ++b;
  }
  const char* dot = static_cast<const char*>(std::memchr(b, '.', static_cast<size_t>(e - b)));
  uint64_t whole = 0, frac = 0;
  if (dot != b && !parse_uint(b, dot ? dot : e, whole)) return false;
  if (dot) {
    const char* f = dot + 1;
    if (e - f > 2 || (e != f && !parse_uint(f, e, frac))) return false;
    if (e - f == 1) frac *= 10;
  }
  const int64_t x = static_cast<int64_t>(whole * 100 + frac);
  v = neg ? -x : x;
  return true;
}

inline bool parse_date(const char* b, const char* e, uint32_t& jd) {
  if (e - b != 10 || b[4] != '-' || b[7] != '-') return false;
  uint64_t y, m, d;
  if (!parse_uint(b, b + 4, y) || !parse_uint(b + 5, b + 7, m) || !parse_uint(b + 8, b + 10, d)) return false;
  if (m < 1 || m > 12 || d < 1 || d > 31) return false;
  jd = to_julian_day(static_cast<uint32_t>(d), static_cast<uint32_t>(m), static_cast<uint32_t>(y));
  return true;
}

// Splits line [b, e) at '|' into at most `want` fields (dbgen ends each line with '|').
inline int split_fields(const char* b, const char* e, const char** fb, const char** fe, int want) {
  int k = 0;
  while (k < want && b <= e) {
    const char* bar = static_cast<const char*>(std::memchr(b, '|', static_cast<size_t>(e - b)));
    fb[k] = b;
    fe[k] = bar ? bar : e;
    ++k;
    if (!bar) break;
    b = bar + 1;
  }
  return k;
}

// Calls row(thread, fields_begin, fields_end, line_no_hint) for every non-empty line,
// each thread over its own byte range; returns the thread count used.  A row
// callback returns false on a malformed line, which makes the load throw.
template <class Row>
int for_each_line(const mapped_file& f, int max_fields, const std::string& what, Row row) {
  const char* data = f.data();
  const size_t size = f.size();
  int threads = 1;
#ifdef _OPENMP
  threads = std::max(1, omp_get_max_threads());
#endif
  if (size < (1u << 20)) threads = 1;
  std::vector<size_t> cut(static_cast<size_t>(threads) + 1, size);
  cut[0] = 0;
  for (int t = 1; t < threads; ++t) {
    size_t p = size * static_cast<size_t>(t) / static_cast<size_t>(threads);
    const char* nl = static_cast<const char*>(std::memchr(data + p, '\n', size - p));
    cut[static_cast<size_t>(t)] = nl ? static_cast<size_t>(nl - data) + 1 : size;
  }
  for (int t = 1; t <= threads; ++t) cut[static_cast<size_t>(t)] = std::max(cut[static_cast<size_t>(t)], cut[static_cast<size_t>(t) - 1]);
  std::vector<int> bad(static_cast<size_t>(threads), 0);
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int t = 0; t < threads; ++t) {
    const char* fb[32];
    const char* fe[32];
    const char* p = data + cut[static_cast<size_t>(t)];
    const char* end = data + cut[static_cast<size_t>(t) + 1];
    while (p < end) {
      const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
      const char* le = nl ? nl : end;
      const char* lt = le;
      if (lt > p && lt[-1] == '\r') --lt;
      if (lt > p) {
        const int k = split_fields(p, lt, fb, fe, max_fields);
        if (k < max_fields || !row(t, fb, fe)) {
          bad[static_cast<size_t>(t)] = 1;
          break;
        }
      }
      p = le + 1;
    }
  }
  for (int t = 0; t < threads; ++t)
    if (bad[static_cast<size_t>(t)]) throw std::runtime_error("malformed line in " + what);
  return threads;
}

template <class T>
void append_all(std::vector<T>& dst, std::vector<std::vector<T>>& parts) {
  size_t n = dst.size();
  for (const auto& p : parts) n += p.size();
  dst.reserve(n);
  for (auto& p : parts) {
    dst.insert(dst.end(), p.begin(), p.end());
    std::vector<T>().swap(p);
  }
}

inline q14_lineitem load_lineitem(const std::string& path) {
  mapped_file f(path);
  const int max_threads =
#ifdef _OPENMP
      std::max(1, omp_get_max_threads());
#else
      1;
#endif
  std::vector<std::vector<int32_t>> pk(max_threads), sd(max_threads);
  std::vector<std::vector<int64_t>> ep(max_threads);
  std::vector<std::vector<int8_t>> dc(max_threads);
  const int used = for_each_line(f, 11, path, [&](int t, const char** fb, const char** fe) {
    uint64_t key;
    int64_t price, disc;
    uint32_t day;
    if (!parse_uint(fb[1], fe[1], key) || key > 0x7FFFFFFFull) return false;
    if (!parse_numeric2(fb[5], fe[5], price) || !parse_numeric2(fb[6], fe[6], disc)) return false;
    if (disc < -128 || disc > 127) return false;
    if (!parse_date(fb[10], fe[10], day)) return false;
    pk[static_cast<size_t>(t)].push_back(static_cast<int32_t>(key));
    ep[static_cast<size_t>(t)].push_back(price);
    dc[static_cast<size_t>(t)].push_back(static_cast<int8_t>(disc));
    sd[static_cast<size_t>(t)].push_back(static_cast<int32_t>(day));
    return true;
  });
  pk.resize(static_cast<size_t>(used));
  sd.resize(static_cast<size_t>(used));
  ep.resize(static_cast<size_t>(used));
  dc.resize(static_cast<size_t>(used));
  q14_lineitem li;
  append_all(li.partkey, pk);
  append_all(li.shipdate, sd);
  append_all(li.extendedprice, ep);
  append_all(li.discount, dc);
  return li;
}

// p_partkey ascending (dbgen writes it so; sorted here otherwise), flag = "PROMO" prefix of p_type
inline q14_part load_part(const std::string& path) {
  mapped_file f(path);
  const int max_threads =
#ifdef _OPENMP
      std::max(1, omp_get_max_threads());
#else
      1;
#endif
  std::vector<std::vector<uint32_t>> pk(max_threads);
  std::vector<std::vector<uint8_t>> pr(max_threads);
  const int used = for_each_line(f, 5, path, [&](int t, const char** fb, const char** fe) {
    uint64_t key;
    if (!parse_uint(fb[0], fe[0], key) || key > 0xFFFFFFFEull) return false;
    const bool promo = fe[4] - fb[4] >= 5 && std::memcmp(fb[4], "PROMO", 5) == 0;
    pk[static_cast<size_t>(t)].push_back(static_cast<uint32_t>(key));
    pr[static_cast<size_t>(t)].push_back(promo ? 1 : 0);
    return true;
  });
  pk.resize(static_cast<size_t>(used));
  pr.resize(static_cast<size_t>(used));
  q14_part part;
  append_all(part.partkey, pk);
  append_all(part.promo, pr);
  if (!std::is_sorted(part.partkey.begin(), part.partkey.end())) {
    std::vector<size_t> order(part.size());
    std::iota(order.begin(), order.end(), size_t{0});
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return part.partkey[a] < part.partkey[b]; });
    q14_part s;
    s.partkey.reserve(part.size());
    s.promo.reserve(part.size());
    for (size_t i : order) {
      s.partkey.push_back(part.partkey[i]);
      s.promo.push_back(part.promo[i]);
    }
    part = std::move(s);
  }
  if (std::adjacent_find(part.partkey.begin(), part.partkey.end()) != part.partkey.end())
    throw std::runtime_error("duplicate p_partkey in " + path);
  return part;
}

// Q14 on the host, the reference's CPU loop (query_14_part_build,
// src/tpch/cpu_only.cpp:230-275) with a binary search over the sorted part keys
// in place of its hash table: num/den in hundredths squared.
struct q14_sums {
  int64_t numerator = 0;
  int64_t denominator = 0;
  uint64_t matches = 0;
};

inline q14_sums q14_host(const q14_lineitem& li, const q14_part& part, uint32_t lo = q14_lower_shipdate,
                         uint32_t hi = q14_upper_shipdate) {
  int64_t num = 0, den = 0;
  uint64_t hits = 0;
  const int64_t n = static_cast<int64_t>(li.size());
#pragma omp parallel for reduction(+ : num, den, hits) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t d = static_cast<uint32_t>(li.shipdate[static_cast<size_t>(i)]);
    if (d < lo || d >= hi) continue;
    const uint32_t key = static_cast<uint32_t>(li.partkey[static_cast<size_t>(i)]);
    const auto it = std::lower_bound(part.partkey.begin(), part.partkey.end(), key);
    if (it == part.partkey.end() || *it != key) continue;
    const int64_t s = li.extendedprice[static_cast<size_t>(i)] * (100 - li.discount[static_cast<size_t>(i)]);
    den += s;
    if (part.promo[static_cast<size_t>(it - part.partkey.begin())]) num += s;
    ++hits;
  }
  return {num, den, hits};
}

// process_results (src/tpch/tpch_14_common.cu:116-132): percent x 1e6
inline int64_t q14_result(int64_t numerator, int64_t denominator) {
  const int64_t d = denominator / 1000;
  return d ? 100 * (numerator * 1000) / d : 0;
}

// known answers by lineitem row count (expected_results, src/tpch/tpch_14_common.cu:95-99)
inline bool q14_known_answer(size_t lineitem_rows, int64_t& expected) {
  switch (lineitem_rows) {
    case 6001215: expected = 16380778; return true;    // SF 1
    case 59986052: expected = 16647594; return true;   // SF 10
    case 59986050: expected = 16647594; return true;
    default: return false;
  }
}

}  // namespace gpujoin::tpch
