/*
 * gj_runtime.h - host-side plumbing exported by libgj.so next to gj.h:
 * device memory, streams, events and the seeded workload generators.
 *
 * Device memory replaces the reference's owning device arrays
 * (device_array_wrapper, src/device_array.hpp:132-236; cuda_allocator,
 * src/cuda_allocator.hpp:41-70).  All data lives in HBM; the reference's
 * host-memory / NUMA placement (src/mmap_allocator.hpp) has no counterpart.
 *
 * The generators replace generate_datasets and its helpers
 * (src/index_lookup_common.cuh:44-206) and the Zipf sampler
 * (src/zipf.hpp:13-102).  The reference seeds std::mt19937 from
 * std::random_device, so its data is not reproducible; these generators are
 * counter-based (splitmix64 of seed and element index), so the output depends
 * only on the arguments, not on the thread count.
 */
#ifndef GJ_RUNTIME_H
#define GJ_RUNTIME_H

#include <stddef.h>
#include <stdint.h>

#include "gj.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Device / memory / streams / events --------------------------------- */
int gj_device_count(int* count);
/* Fills name (NUL-terminated, at most len bytes), compute units and the device's max workgroups-in-flight hint. */
int gj_device_info(char* name, size_t len, int* compute_units, int* gcn_arch_major);
int gj_malloc(void** d_ptr, uint64_t bytes);
int gj_free(void* d_ptr);
/* Page-locked host memory (hipHostMalloc), mapped into every GPU's address space at the
 * same address: kernels may read it in place (e.g. an index column kept in host memory). */
int gj_host_malloc(void** h_ptr, uint64_t bytes);
int gj_host_free(void* h_ptr);
int gj_memcpy_htod(void* d_dst, const void* h_src, uint64_t bytes, gj_stream_t stream);
int gj_memcpy_dtoh(void* h_dst, const void* d_src, uint64_t bytes, gj_stream_t stream);
int gj_memcpy_dtod(void* d_dst, const void* d_src, uint64_t bytes, gj_stream_t stream);
/* Fill with a byte value by a kernel (safe inside graph capture). */
int gj_memset(void* d_dst, int value, uint64_t bytes, gj_stream_t stream);
int gj_stream_create(gj_stream_t* stream);
int gj_stream_destroy(gj_stream_t stream);
int gj_stream_synchronize(gj_stream_t stream);
int gj_device_synchronize(void);

typedef struct gj_event_s* gj_event_t;
int gj_event_create(gj_event_t* ev);
int gj_event_destroy(gj_event_t ev);
int gj_event_record(gj_event_t ev, gj_stream_t stream);
int gj_event_elapsed_ms(gj_event_t start, gj_event_t stop, float* ms);

/* ---- Seeded workload generators (host, OpenMP) ---------------------------
 * threads <= 0 uses every hardware thread. */

/* n unique keys, ascending.  Key distribution:
 *   GJ_KEYS_DENSE      0, 1, ..., n-1 (generate_dense_dataset, src/index_lookup_common.cuh:44-49)
 *   GJ_KEYS_UNIFORM    uniform without replacement from [0, upper) (generate_unique_uniform_sorted_dataset,
 *                      src/index_lookup_common.cuh:73-138), upper = 2^upper_bits (0 = full key width)
 *   GJ_KEYS_LOGNORMAL  lognormal(mu=0, sigma=2) scaled to [0, upper), deduplicated and topped up
 *   GJ_KEYS_CLUSTERED  64 uniform centres in [0, upper) with Gaussian spread 2^20, deduplicated and topped up */
typedef enum gj_key_dist { GJ_KEYS_DENSE = 0, GJ_KEYS_UNIFORM = 1, GJ_KEYS_LOGNORMAL = 2, GJ_KEYS_CLUSTERED = 3 } gj_key_dist;
int gj_gen_build_keys(void* h_out, uint64_t n, int key_bytes, gj_key_dist dist, uint32_t upper_bits, uint64_t seed,
                      int threads);

/* m probes over a sorted build column of n keys.
 *   GJ_PROBES_UNIFORM  rank uniform in [0, n) with replacement (generate_uniform_dataset, src/index_lookup_common.cuh:51-66)
 *   GJ_PROBES_ZIPF     rank = k - 1 where k ~ Zipf(theta) on {1..n} by rejection-inversion
 *                      (zipf_distribution, src/zipf.hpp:13-102; the reference draws k in [1, n-1] and
 *                       never rank 0, src/index_lookup_common.cuh:147 - here every rank is reachable)
 * miss_fraction in [0, 1): that fraction of probes are keys drawn uniformly from [0, upper) that are
 * absent from the column (the sorted column is searched to reject present keys).
 * h_expected (optional) receives the expected row id per probe (rank, or GJ_NOT_FOUND for misses). */
typedef enum gj_probe_dist { GJ_PROBES_UNIFORM = 0, GJ_PROBES_ZIPF = 1 } gj_probe_dist;
int gj_gen_probes(const void* h_sorted_keys, uint64_t n, int key_bytes, void* h_out, uint64_t m, gj_probe_dist dist,
                  double zipf_theta, double miss_fraction, uint32_t upper_bits, uint64_t seed, uint32_t* h_expected,
                  int threads);

/* Dimension/fact tables of the join-aggregate workload.
 * Dimension: key = random permutation of 1..dim_rows, flag = 1 with probability 1/3.
 * Fact: fkey uniform in [1, dim_rows], date uniform in [0, days), price uniform in
 * [price_lo, price_hi], discount uniform in [0, disc_max]. */
int gj_gen_dimension(int32_t* h_key, uint8_t* h_flag, uint64_t dim_rows, uint64_t seed, int threads);
int gj_gen_fact(int32_t* h_fkey, int32_t* h_date, int64_t* h_price, int8_t* h_discount, uint64_t rows,
                uint64_t dim_rows, int32_t days, int64_t price_lo, int64_t price_hi, int32_t disc_max, uint64_t seed,
                int threads);

/* Sort the dimension by key (as sort_relation does for part, src/tpch/common.cpp:223-226):
 * writes the sorted keys and the flags permuted alongside. */
int gj_sort_dimension(const int32_t* h_key, const uint8_t* h_flag, uint64_t rows, uint32_t* h_sorted_key,
                      uint8_t* h_sorted_flag);

/* Parallel ascending sort of a host key array (used to sort generated data). */
int gj_sort_keys(void* h_keys, uint64_t n, int key_bytes, int threads);

/* splitmix64 finaliser of (seed, index); the counter-based stream every generator draws from. */
uint64_t gj_rng_u64(uint64_t seed, uint64_t index);

/* One rejection-inversion trial of the Zipf(theta) sampler on {1..n} that GJ_PROBES_ZIPF draws
 * from, fed the uniform u01 in [0, 1): the rank k, or 0 when the trial is rejected
 * (zipf_distribution::operator(), src/zipf.hpp:30-42). */
uint64_t gj_zipf_trial(uint64_t n, double theta, double u01);

#ifdef __cplusplus
}
#endif

#endif /* GJ_RUNTIME_H */
