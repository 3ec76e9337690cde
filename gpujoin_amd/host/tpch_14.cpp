// tpch_14.cpp - TPC-H query 14 over dbgen tables, the counterpart of the
// reference's tpch_14 harness (main src/tpch/tpch_14_new.cu:7-26, options
// src/tpch/config.cpp:19-30, execute_approach src/tpch/tpch_14_common.cu:790-807).
//
//   tpch_14 [-a approach] [-i index] [-b blocksize] [-d destfile] [--cpu-only] <db_path>
//
//   -a,--approach   hj | ij_plain | ij_np_lf_plain | ij_np_lf_lr_plain | ij_partitioning | ij_p_bws
//                   (default hj; the reference's six approach names)
//   -i,--index      index over p_partkey for the ij_* approaches      (default btree)
//   -b,--blocksize  accepted for compatibility (the kernel's workgroups are fixed)
//   -d,--destfile   append YAML records here (3 repetitions); default stdout, 1 repetition
//   --cpu-only      load and answer on the host only (no GPU calls)
//
// <db_path> holds lineitem.tbl and part.tbl.  Each approach runs gj_join_aggregate
// (date filter, index lookup, exact int64 sums) with its strategy:
//   hj                 the linear-probing hash table over p_partkey, probed row by row (plain)
//   ij_plain           the chosen index, each lane looking up its own rows (plain)
//   ij_np_lf_plain     lookups first, writing (fact row, dimension row) join entries; then a join
//                      kernel gathers and sums them
//   ij_np_lf_lr_plain  the chosen index with per-wave lane refill (fused with the sums)
//   ij_partitioning    filter, radix-partition the qualifying rows on the fkey, then look up
//   ij_p_bws           each workgroup sorts its buffered qualifying rows by fkey, then looks up
// Each repetition is checked against the host restatement of the query and,
// at SF 1 and SF 10, against the reference's known answers.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "gpujoin.hpp"
#include "tpch.hpp"

using namespace gpujoin;

namespace {

struct config {
  std::string approach = "hj";
  std::string index_type = "btree";
  std::string dest_file;
  std::string db_path;
  unsigned block_size = 128;
  bool cpu_only = false;
};

void usage() {
  std::cout << "tpch_14 [-a hj|ij_plain|ij_np_lf_plain|ij_np_lf_lr_plain|ij_partitioning|ij_p_bws] [-i index]\n"
               "        [-b blocksize] [-d destfile] [--cpu-only] <db_path>\n";
}

config parse(int argc, char** argv) {
  config c;
  const std::map<std::string, std::string> alias = {{"a", "approach"}, {"i", "index"}, {"b", "blocksize"},
                                                    {"d", "destfile"}, {"h", "help"}};
  std::vector<std::string> positional;
  for (int k = 1; k < argc; ++k) {
    std::string a = argv[k];
    if (a.size() < 2 || a[0] != '-') {
      positional.push_back(a);
      continue;
    }
    std::string key = a.substr(a[1] == '-' ? 2 : 1);
    std::string val;
    const auto eq = key.find('=');
    if (eq != std::string::npos) {
      val = key.substr(eq + 1);
      key = key.substr(0, eq);
    }
    if (alias.count(key)) key = alias.at(key);
    if (key == "help") {
      usage();
      std::exit(0);
    }
    if (key == "cpu-only") {
      c.cpu_only = true;
      continue;
    }
    if (eq == std::string::npos) {
      if (k + 1 >= argc) throw std::runtime_error("missing value for -" + key);
      val = argv[++k];
    }
    if (key == "approach") c.approach = val;
    else if (key == "index") c.index_type = val;
    else if (key == "blocksize") c.block_size = static_cast<unsigned>(std::stoul(val));
    else if (key == "destfile") c.dest_file = val;
    else throw std::runtime_error("unknown option -" + key);
  }
  if (positional.size() != 1) {
    usage();
    throw std::runtime_error("expected one <db_path>");
  }
  c.db_path = positional[0];
  if (!c.db_path.empty() && c.db_path.back() != '/') c.db_path += '/';
  static const char* approaches[] = {"hj", "ij_plain", "ij_np_lf_plain", "ij_np_lf_lr_plain", "ij_partitioning",
                                     "ij_p_bws"};
  if (std::find(std::begin(approaches), std::end(approaches), c.approach) == std::end(approaches))
    throw std::runtime_error("unknown approach " + c.approach);
  return c;
}

std::string fixed6(int64_t r) {
  char b[64];
  std::snprintf(b, sizeof(b), "%ld.%06ld", static_cast<long>(r / 1000000), static_cast<long>(r % 1000000));
  return b;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const config c = parse(argc, argv);
    auto t0 = std::chrono::steady_clock::now();
    const tpch::q14_lineitem li = tpch::load_lineitem(c.db_path + "lineitem.tbl");
    const tpch::q14_part part = tpch::load_part(c.db_path + "part.tbl");
    const double load_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("lineitem_size: %zu part_size: %zu (loaded in %.2f s)\n", li.size(), part.size(), load_s);

    const tpch::q14_sums host = tpch::q14_host(li, part);
    const int64_t host_result = tpch::q14_result(host.numerator, host.denominator);
    std::printf("host numerator: %ld denominator: %ld matches: %lu\n", static_cast<long>(host.numerator),
                static_cast<long>(host.denominator), static_cast<unsigned long>(host.matches));
    std::printf("host query result: %s\n", fixed6(host_result).c_str());
    int64_t known = 0;
    const bool have_known = tpch::q14_known_answer(li.size(), known);
    if (have_known) {
      std::printf("known answer: %s (%s)\n", fixed6(known).c_str(), known == host_result ? "match" : "MISMATCH");
      if (known != host_result) return 2;
    }
    if (c.cpu_only) return 0;
    if (part.size() == 0) throw std::runtime_error("part.tbl is empty");

    auto& ms = measuring::get_settings();
    if (!c.dest_file.empty()) {
      ms.dest_file = c.dest_file;
      ms.repetitions = 3;
      ms.stdout_only = false;
    } else {
      ms.repetitions = 1;
      ms.stdout_only = true;
    }

    gj_stream_t stream = nullptr;
    check(gj_stream_create(&stream), "stream");
    device_array<uint32_t> d_pkey(part.partkey);
    device_array<uint8_t> d_flag(part.promo);
    device_array<int32_t> d_fkey(li.partkey), d_date(li.shipdate);
    device_array<int64_t> d_price(li.extendedprice);
    device_array<int8_t> d_disc(li.discount);
    device_array<gj_agg_result> d_res(1);

    const gj_index_kind kind = c.approach == "hj" ? GJ_HASH : parse_index_type(c.index_type);
    gj_index_params ip;
    gj_index_params_default(&ip);
    gj_index* ix = nullptr;
    check(gj_index_build(kind, part.partkey.data(), d_pkey.data(), part.size(), 4, &ip, stream, &ix), "index build");
    std::printf("index size: %lu bytes\n", static_cast<unsigned long>(gj_index_memory(ix)));

    gj_fact_batch batch{d_fkey.data(), d_date.data(), d_price.data(), d_disc.data(), li.size()};
    const uint32_t strategy = c.approach == "ij_np_lf_lr_plain" ? GJ_AGG_LANE_REFILL
                              : c.approach == "ij_partitioning" ? GJ_AGG_PARTITIONED
                              : c.approach == "ij_p_bws"        ? GJ_AGG_SORTED
                              : c.approach == "ij_np_lf_plain"  ? GJ_AGG_LOOKUP_FIRST
                                                                : GJ_AGG_PLAIN;
    gj_agg_params ap{static_cast<int32_t>(tpch::q14_lower_shipdate), static_cast<int32_t>(tpch::q14_upper_shipdate),
                     d_flag.data(), strategy, 0};
    gj_event_t e0 = nullptr, e1 = nullptr;
    check(gj_event_create(&e0), "event");
    check(gj_event_create(&e1), "event");
    gj_agg_result last{};
    measuring::experiment_description d;
    d.name = "tpch_query14";
    d.approach = c.approach;
    d.other = {{"db_path", c.db_path}, {"sort_indexed_relation", "true"}, {"block_size", std::to_string(c.block_size)}};
    if (c.approach != "hj") d.other.emplace_back("index_type", c.index_type);
    float device_ms = 0.f;
    auto run = [&](measuring::measurement& m) {
      check(gj_memset(d_res.data(), 0, sizeof(gj_agg_result), stream), "memset");
      check(gj_event_record(e0, stream), "event");
      check(gj_join_aggregate(ix, &batch, &ap, d_res.data(), stream), "join_aggregate");
      check(gj_event_record(e1, stream), "event");
      check(gj_stream_synchronize(stream), "sync");
      check(gj_event_elapsed_ms(e0, e1, &device_ms), "elapsed");
      m.device_ms = device_ms;
      check(gj_memcpy_dtoh(&last, d_res.data(), sizeof(last), stream), "copy result");
      check(gj_stream_synchronize(stream), "sync");
    };
    auto validator = [&]() {
      const int64_t r = gj_promo_revenue(last.numerator, last.denominator);
      std::printf("numerator: %ld denominator: %ld\n", static_cast<long>(last.numerator),
                  static_cast<long>(last.denominator));
      std::printf("query result: %s\n", fixed6(r).c_str());
      const bool ok = last.numerator == host.numerator && last.denominator == host.denominator &&
                      last.matches == host.matches && (!have_known || r == known);
      if (!ok) std::cerr << "invalid query result" << std::endl;
      return ok;
    };
    auto extra = [&](const measuring::measurement& m) {
      return std::vector<std::pair<std::string, std::string>>{
          {"device_ms", std::to_string(m.device_ms)},
          {"rows_per_s", std::to_string(static_cast<double>(li.size()) / (m.device_ms * 1e-3))},
          {"lineitem_size", std::to_string(li.size())}};
    };
    bool all_ok = true;
    auto checked = [&]() {
      const bool ok = validator();
      all_ok = all_ok && ok;
      return ok;
    };
    measuring::measure(d, run, checked, extra);
    gj_event_destroy(e0);
    gj_event_destroy(e1);
    gj_index_destroy(ix);
    gj_stream_destroy(stream);
    return all_ok ? 0 : 1;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
