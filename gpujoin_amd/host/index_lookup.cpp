// index_lookup.cpp - command-line benchmark driver with the reference's flags
// and YAML records (index_lookup_new.cu main, src/index_lookup_new.cu:375-434;
// flags src/index_lookup_config.cpp:35-97; defaults src/index_lookup_config.hpp:18-34).
//
//   -a,--approach        plain | bws | partitioning | hj           (default plain)
//   -i,--index           btree | harmonia | binary_search | radix_spline | no_op
//                        | learned_plr | learned_mlp | hash          (default btree)
//   -e,--elements        keys in the index                         (default 1e6)
//   -l,--lookups         probe keys                                (default 1e8)
//   -b,--blocksize       threads per block                         (default 128)
//   -m,--maxbits         sparse keys are drawn below 2^(maxbits-1); 0 = the full key width
//   -z,--zipf            Zipf exponent for -p zipf                 (default 1.25)
//   -d,--dataset         dense | sparse                            (default sparse)
//   -p,--lookup_pattern  uniform | uniform_unique | zipf           (default uniform)
//   -s,--sorted_lookups  true | false                              (default false)
//   -o,--output          append YAML records to this file (default: stdout)
//   -w,--window          partitioning window in probes             (default 10000)
//   -r,--repetitions     timed repetitions after 3 warm-ups        (default 10)
//
// Keys are uint64 and row ids uint32.  Every repetition is validated
// (tid < n and indexed[tid] == lookup key) before its record is written.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <map>
#include <numeric>
#include <string>

#include "gpujoin.hpp"

using namespace gpujoin;
using probe_key_t = uint64_t;

namespace {

struct experiment_config {
  std::string approach = "plain";
  std::string index_type = "btree";
  std::string output_file;
  uint64_t num_lookups = 100000000;
  uint64_t num_elements = 1000000;
  unsigned max_bits = 0;
  double zipf_factor = 1.25;
  unsigned block_size = 128;
  std::string dataset = "sparse";
  std::string lookup_pattern = "uniform";
  bool sorted_lookups = false;
  uint64_t window = 10000;
  unsigned repetitions = 10;
  uint64_t seed = 0x11;
};

void usage() {
  std::cout << "index_lookup -a approach -i index -e elements -l lookups -b blocksize -m maxbits -z zipf\n"
               "             -d dense|sparse -p uniform|uniform_unique|zipf -s true|false -o file -w window\n";
}

bool parse_bool(const std::string& v) { return v == "1" || v == "true" || v == "yes"; }

experiment_config parse_options(int argc, char** argv) {
  experiment_config c;
  const std::map<std::string, std::string> alias = {
      {"a", "approach"}, {"i", "index"}, {"e", "elements"}, {"l", "lookups"}, {"b", "blocksize"},
      {"m", "maxbits"}, {"z", "zipf"}, {"d", "dataset"}, {"p", "lookup_pattern"}, {"s", "sorted_lookups"},
      {"o", "output"}, {"w", "window"}, {"r", "repetitions"}, {"h", "help"}};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i], key, val;
    if (a.rfind("--", 0) == 0) {
      key = a.substr(2);
      const auto eq = key.find('=');
      if (eq != std::string::npos) {
        val = key.substr(eq + 1);
        key = key.substr(0, eq);
      }
    } else if (a.size() >= 2 && a[0] == '-') {
      key = a.substr(1, 1);
      if (a.size() > 2) val = a.substr(2);
    } else {
      throw std::runtime_error("unexpected argument: " + a);
    }
    if (alias.count(key)) key = alias.at(key);
    if (key == "help") {
      usage();
      std::exit(0);
    }
    if (val.empty()) {
      if (i + 1 >= argc) throw std::runtime_error("missing value for --" + key);
      val = argv[++i];
    }
    if (key == "approach") c.approach = val;
    else if (key == "index") c.index_type = val;
    else if (key == "elements") c.num_elements = static_cast<uint64_t>(std::stod(val));
    else if (key == "lookups") c.num_lookups = static_cast<uint64_t>(std::stod(val));
    else if (key == "blocksize") c.block_size = static_cast<unsigned>(std::stoul(val));
    else if (key == "maxbits") c.max_bits = static_cast<unsigned>(std::stoul(val));
    else if (key == "zipf") c.zipf_factor = std::stod(val);
    else if (key == "dataset") c.dataset = val;
    else if (key == "lookup_pattern") c.lookup_pattern = val;
    else if (key == "sorted_lookups") c.sorted_lookups = parse_bool(val);
    else if (key == "output") c.output_file = val;
    else if (key == "window") c.window = static_cast<uint64_t>(std::stod(val));
    else if (key == "repetitions") c.repetitions = static_cast<unsigned>(std::stoul(val));
    else throw std::runtime_error("unknown option --" + key);
  }
  if (c.dataset != "dense" && c.dataset != "sparse") throw std::runtime_error("dataset must be dense or sparse");
  if (c.lookup_pattern != "uniform" && c.lookup_pattern != "uniform_unique" && c.lookup_pattern != "zipf")
    throw std::runtime_error("lookup_pattern must be uniform, uniform_unique or zipf");
  return c;
}

// generate_datasets (src/index_lookup_common.cuh:155-206), seeded
void generate_datasets(const experiment_config& c, std::vector<probe_key_t>& keys, std::vector<probe_key_t>& lookups) {
  const uint64_t n = keys.size();
  if (c.dataset == "sparse" && c.max_bits > 0) {
    const double limit = std::ldexp(1.0, static_cast<int>(c.max_bits) - 1);
    if (static_cast<double>(n - 1) > limit)
      throw std::runtime_error("resulting dataset would exceed the provided limit defined by 'max_bits'");
    if (c.max_bits < std::log2(2.0 * static_cast<double>(n))) throw std::runtime_error("index dataset not sparse enough");
  }
  if (c.lookup_pattern == "uniform_unique" && 2 * n < lookups.size())
    throw std::runtime_error("lookup dataset not sparse enough");
  const gj_key_dist dist = c.dataset == "dense" ? GJ_KEYS_DENSE : GJ_KEYS_UNIFORM;
  const uint32_t upper = (c.dataset == "sparse" && c.max_bits > 0) ? c.max_bits - 1 : 0;
  check(gj_gen_build_keys(keys.data(), n, 8, dist, upper, c.seed, 0), "generate keys");
  if (c.lookup_pattern == "uniform_unique") {
    // distinct positions: a seeded partial Fisher-Yates shuffle of [0, n)
    std::vector<uint64_t> pos(n);
    std::iota(pos.begin(), pos.end(), 0);
    for (uint64_t i = 0; i < lookups.size() && i < n; ++i) {
      const uint64_t r = gj_rng_u64(c.seed + 2, i);
      const uint64_t j = i + static_cast<uint64_t>((static_cast<unsigned __int128>(r) * (n - i)) >> 64);
      std::swap(pos[i], pos[j]);
      lookups[i] = keys[pos[i]];
    }
  } else {
    const gj_probe_dist pd = c.lookup_pattern == "zipf" ? GJ_PROBES_ZIPF : GJ_PROBES_UNIFORM;
    check(gj_gen_probes(keys.data(), n, 8, lookups.data(), lookups.size(), pd, c.zipf_factor, 0.0, upper,
                        c.seed + 1, nullptr, 0),
          "generate lookups");
  }
}

}  // namespace

int main(int argc, char** argv) {
  try {
    experiment_config c = parse_options(argc, argv);
    auto& ms = measuring::get_settings();
    ms.repetitions = c.repetitions;
    ms.stdout_only = c.output_file.empty();
    ms.dest_file = c.output_file;
    // enforce the Harmonia size constraint for every index (src/index_lookup_new.cu:43-49)
    if (c.num_elements % 32 != 0) {
      c.num_elements = (c.num_elements & ~uint64_t(31)) + 32;
      std::printf("adjusting dataset size to: %lu\n", static_cast<unsigned long>(c.num_elements));
    }
    query_data<probe_key_t> qd;
    qd.indexed.resize(c.num_elements);
    qd.lookup_keys.resize(c.num_lookups);
    generate_datasets(c, qd.indexed, qd.lookup_keys);
    if (c.sorted_lookups) std::sort(qd.lookup_keys.begin(), qd.lookup_keys.end());
    qd.dataset_max_bits = static_cast<unsigned>(std::log2(static_cast<double>(qd.indexed.back())));
    qd.d_indexed = device_array<probe_key_t>(qd.indexed);
    qd.d_lookup_keys = device_array<probe_key_t>(qd.lookup_keys);
    qd.d_tids = device_array<uint32_t>(qd.lookup_keys.size());
    check(gj_stream_create(&qd.stream), "stream");
    qd.index_structure = build_index<probe_key_t>(c.index_type, qd.indexed, qd.d_indexed.data());
    std::printf("index size: %zu bytes\n", qd.index_structure->memory_consumption());

    auto approach = make_approach<probe_key_t>(c.approach);
    approach->params.block_size = c.block_size;
    approach->params.window = c.window;
    approach->initialize(qd);

    // G: independent random 8 B gathers per second over the indexed column (its
    // power-of-two prefix) with the same probe stream, the roofline every record's
    // lookups/s is reported against
    double gathers_per_s = 0;
    {
      uint64_t n2 = 1;
      while (n2 * 2 <= qd.indexed.size()) n2 *= 2;
      gj_event_t a = nullptr, b = nullptr;
      check(gj_event_create(&a), "event");
      check(gj_event_create(&b), "event");
      auto gather = [&] {
        check(gj_gather(qd.d_indexed.data(), n2, qd.d_lookup_keys.data(), c.num_lookups, qd.d_tids.data(), 256, 8,
                        qd.stream),
              "gather");
      };
      gather();
      check(gj_event_record(a, qd.stream), "event");
      for (int r = 0; r < 3; ++r) gather();
      check(gj_event_record(b, qd.stream), "event");
      check(gj_stream_synchronize(qd.stream), "sync");
      float ms = 0;
      check(gj_event_elapsed_ms(a, b, &ms), "elapsed");
      gathers_per_s = ms > 0 ? 3.0 * static_cast<double>(c.num_lookups) / (ms * 1e-3) : 0;
      gj_event_destroy(a);
      gj_event_destroy(b);
    }

    char dev[256] = {0};
    int cu = 0, major = 0;
    check(gj_device_info(dev, sizeof(dev), &cu, &major), "device");
    measuring::experiment_description d;
    d.name = "plain_lookup";
    d.approach = c.approach;
    auto& o = d.other;
    o.emplace_back("device", dev);
    o.emplace_back("index_type", c.index_type);
    o.emplace_back("dataset", c.dataset);
    o.emplace_back("lookup_pattern", c.lookup_pattern);
    o.emplace_back("num_elements", "!int64 " + std::to_string(c.num_elements));
    o.emplace_back("num_lookups", "!int64 " + std::to_string(c.num_lookups));
    o.emplace_back("sorted_lookups", std::to_string(c.sorted_lookups));
    o.emplace_back("host_allocator", "std_allocator");
    o.emplace_back("device_index_allocator", "hbm");
    o.emplace_back("indexed_allocator", "hbm");
    o.emplace_back("lookup_keys_allocator", "hbm");
    if (c.dataset == "sparse") o.emplace_back("max_bits", std::to_string(c.max_bits));
    if (c.lookup_pattern == "zipf") o.emplace_back("zipf_factor", std::to_string(c.zipf_factor));
    if (c.approach == "partitioning") {
      o.emplace_back("partitioning_approach_ignore_bits", "dynamic");
      o.emplace_back("partitioning_approach_dynamic_bit_range", "1");
      o.emplace_back("partitioning_approach_window_size", std::to_string(c.window));
    }
    o.emplace_back("index_size", "!int64 " + std::to_string(qd.index_structure->memory_consumption()));

    measuring::measure(
        d, [&](measuring::measurement& m) { approach->run(qd, m); }, [&] { return qd.validate_results(); },
        [&](const measuring::measurement& m) {
          std::vector<std::pair<std::string, std::string>> x;
          x.emplace_back("device_ms", std::to_string(m.device_ms));
          const double lps = c.num_lookups / (m.device_ms * 1e-3);
          x.emplace_back("lookups_per_s", std::to_string(lps));
          x.emplace_back("gather_per_s", std::to_string(gathers_per_s));
          x.emplace_back("gather_roofline_frac", std::to_string(gathers_per_s > 0 ? lps / gathers_per_s : 0.0));
          // 8 B probe key + 8 B indexed key + 4 B row id per lookup; HBM traffic and the
          // L2 hit rate need the profiler's counters
          x.emplace_back("algorithmic_GBps", std::to_string(lps * 20.0 / 1e9));
          return x;
        });
    gj_stream_destroy(qd.stream);
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
