/*
 * gj.h - C ABI of the MI355X index-join probe library (libgj.so).
 *
 * This is the drop-in boundary for gpujoin's index-based join probe path.
 * Each entry point cites the reference interface it replaces.  The reference
 * has no C ABI: its boundary is the C++ template concept
 *   abstract_index<Key>::construct(h_column, d_column) / memory_consumption()
 *   (src/indexes.cuh:30-39) plus a by-value device_index_t with
 *   cooperative_lookup(active, key) (src/indexes.cuh:120-123, 205-215,
 *   278-282, 357-361, 398), driven by abstract_approach::{initialize,run}
 *   (src/index_lookup.cuh:29-35) and looked up by name
 *   (src/index_lookup_new.cu:305-310, src/indexes.cpp:5-19).
 *
 * Conventions
 *   - Pointers prefixed d_ are device (HBM) pointers, h_ are host pointers.
 *   - Every entry point returns GJ_OK (0) or a negative gj_status; no C++
 *     exception crosses this boundary.  gj_last_error() returns a
 *     thread-local message for the last failure on the calling thread.
 *   - Work is stream-ordered on the gj_stream_t passed in (NULL = the null
 *     stream).  No entry point synchronises the device unless it says so.
 *   - Row id = position of the key in the sorted build column.  An absent
 *     key yields GJ_NOT_FOUND for every index kind (the reference's
 *     binary_search and radix_spline return the lower-bound position without
 *     an equality check, src/indexes.cuh:357-361, 205-215; this library
 *     checks equality everywhere, which agrees on every hit).
 *   - The build column is borrowed, never copied or freed: it must stay
 *     alive and unchanged while the index exists (as in the reference,
 *     src/indexes.cuh:223, 365; src/btree.cuh:447-451).
 */
#ifndef GJ_H
#define GJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t, declared without pulling in the HIP headers. */
typedef struct ihipStream_t* gj_stream_t;

#define GJ_NOT_FOUND 0xFFFFFFFFu /* reference: numeric_limits<value_t>::max(), src/indexes.cuh:99; src/tpch/config.hpp:38 */

typedef enum gj_status {
  GJ_OK = 0,
  GJ_ERR_INVALID = -1,     /* bad argument (null pointer, bad size, unsorted column, ...) */
  GJ_ERR_UNSUPPORTED = -2, /* valid request the library does not implement */
  GJ_ERR_HIP = -3,         /* a HIP runtime call failed */
  GJ_ERR_NOMEM = -4,       /* device or host allocation failed */
  GJ_ERR_OVERFLOW = -5,    /* row ids do not fit in uint32 ("'value_t' does not suffice", src/indexes.cuh:219-221, 367-369) */
  GJ_ERR_INTERNAL = -6
} gj_status;

/* Index kinds.  Names accepted by gj_parse_index_type follow
 * parse_index_type (src/indexes.cpp:5-19): "binary_search", "btree",
 * "harmonia", "radix_spline", "no_op"; plus "learned_plr", "learned_mlp"
 * and "hash" which have no reference counterpart. */
typedef enum gj_index_kind {
  GJ_BINARY_SEARCH = 0, /* binary_search_index, src/indexes.cuh:336-375 */
  GJ_BTREE = 1,         /* btree_index, src/indexes.cuh:93-147 */
  GJ_HARMONIA = 2,      /* harmonia_index, src/indexes.cuh:257-307 */
  GJ_RADIX_SPLINE = 3,  /* radix_spline_index, src/indexes.cuh:176-248 */
  GJ_LEARNED_PLR = 4,   /* two-stage FP32 model: linear root + linear leaves, B-tree fallback */
  GJ_LEARNED_MLP = 5,   /* two-stage FP32 model: 1xH MLP root + linear leaves, B-tree fallback */
  GJ_NO_OP = 6,         /* no_op_index, src/indexes.cuh:384-406 */
  GJ_HASH = 7           /* linear_probing_hashtable over the column, src/linear_probing_hashtable.cuh:14-98 */
} gj_index_kind;

/* Lookup drivers ("approaches"), src/index_lookup_new.cu:305-310. */
typedef enum gj_approach {
  GJ_PLAIN = 0,        /* plain_approach, src/index_lookup_new.cu:180-200 */
  GJ_BWS = 1,          /* blockwise_sorting_approach, src/index_lookup_new.cu:202-250 */
  GJ_PARTITIONING = 2, /* partitioning_approach, src/index_lookup_partitioning.cu:377-513 */
  GJ_HJ = 3            /* hj_approach: build a hash table over the column, then probe, src/index_lookup_new.cu:252-302 */
} gj_approach;

/* Index construction parameters.  Zero in any field selects its default
 * (gj_index_params_default fills the defaults explicitly). */
typedef struct gj_index_params {
  uint32_t rs_radix_bits;        /* RadixSpline radix bits, default 18 */
  uint32_t rs_max_error;         /* RadixSpline error bound, default 32 */
  uint32_t harmonia_group_lanes; /* lanes per Harmonia node search (8, 16, or 32 for u32 keys), default 8 */
  uint32_t btree_lds_bytes;      /* LDS budget for resident upper B-tree levels, default 16384 */
  uint32_t learned_leaf_bits;    /* log2 of the number of linear leaves; 0 (default) = max(16, ceil(log2 n) - 10)
                                    16-byte leaves, or twice as many 8-byte leaves from 2^22 keys when the
                                    root spreads the keys evenly (linear or MLP root, no fallback leaf) */
  uint32_t learned_max_error;    /* last-mile bound; leaves with a wider measured window use the B-tree, default 128 */
  uint32_t mlp_hidden;           /* hidden width of the MLP root (<= 64), default 32 */
  uint32_t mlp_steps;            /* training steps of the MLP root, default 400 */
  uint32_t mlp_batch;            /* training minibatch (keys per step), default 65536 */
  float mlp_lr;                  /* Adam learning rate, default 3e-3 */
  uint64_t seed;                 /* training sample / initialisation seed, default 0x5eed */
  uint32_t learned_root_segments; /* learned_plr root: 0 = automatic (linear; a CDF spline of <= 2048
                                     segments when the linear root routes keys to the B-tree), 1 = linear,
                                     otherwise the most spline segments (a power of two, <= 4096) */
  uint32_t learned_flags;         /* GJ_LEARNED_* bits, default 0 */
  uint32_t learned_layout;        /* GJ_LAYOUT_AUTO (default), GJ_LAYOUT_COLUMN or GJ_LAYOUT_LINES */
  uint32_t learned_line_keys;     /* lines layout: build keys per 128-byte line on average; default 10 of the
                                     15 a line holds for 8-byte keys, 20 of 30 for 4-byte keys */
  uint32_t rs_lds_bits;           /* RadixSpline: the radix table is staged in LDS when rs_radix_bits <= this
                                     (at most 14, a 64 KiB table), else read from global memory; default 13
                                     (32 KiB) */
  uint32_t placement;             /* GJ_PLACE_DEVICE (default) or GJ_PLACE_HOST: where the index's own
                                     structures live (the column is the caller's and stays where it is) */
  uint32_t btree_node_bytes;      /* GJ_BTREE: bytes per node and per leaf chunk of the column, 64, 128
                                     (default: one cache line) or 256; the learned kinds' fallback B-tree
                                     always uses 128 */
} gj_index_params;

/* placement: the reference keeps its indexes in host memory by default and reads them over
 * the interconnect (src/index_lookup_config.hpp:52-66, src/btree.cuh:440-473).  GJ_PLACE_HOST
 * builds in HBM, then moves every index structure (tree nodes, Harmonia keys and children,
 * spline knots and radix table, learned leaves and lines, the hash table) to page-locked host
 * memory mapped into the GPU's address space; lookups read it in place over the host link. */
#define GJ_PLACE_DEVICE 0u
#define GJ_PLACE_HOST 1u

/* Layouts of the learned kinds.
 *   GJ_LAYOUT_COLUMN  the model predicts a position in the borrowed column and a bounded window
 *                     search closes it (leaves too wide for the bound fall back to the B-tree)
 *   GJ_LAYOUT_LINES   the index keeps its own copy of the keys in 128-byte lines, each key placed
 *                     in the line the model assigns it (a line holds an 8-byte header {first row id,
 *                     key count} and up to 15 8-byte or 30 4-byte keys, in key order; a full line
 *                     spills its surplus into the next line).  A lookup reads the assigned line and,
 *                     only for the keys a full line pushed on, the lines after it, at most the
 *                     build's measured max_lines.  About 12.8 bytes per 8-byte key at the default
 *                     load, against the column's 8.
 *   GJ_LAYOUT_AUTO    lines for a column of 16 MiB or more (smaller ones are served from the
 *                     caches, where the column's denser lines win) when the build's mean lines per
 *                     present key stays at most 1.10, no key needs more than 4 lines and the line
 *                     copy fits; else column.  Always column with GJ_PLACE_HOST (the line copy
 *                     would put every key read on the host link) */
#define GJ_LAYOUT_AUTO 0u
#define GJ_LAYOUT_COLUMN 1u
#define GJ_LAYOUT_LINES 2u

/* learned_flags: keep the MLP root at its initialisation (the piecewise-linear CDF through
 * mlp_hidden + 1 quantiles of the column) and run no training step; the root's loss is
 * still measured and reported in gj_index_stats.train_loss. */
#define GJ_LEARNED_UNTRAINED 0x1u

void gj_index_params_default(gj_index_params* p);

/* Lookup parameters.  Zero selects the default / autotuned value. */
typedef struct gj_lookup_params {
  gj_approach approach;
  uint32_t block_size;       /* threads per block, default 256 */
  uint32_t grid_blocks;      /* 0 = enough to cover the input (plain) or 4 x CUs (persistent kinds);
                                the partitioned driver's lookup then uses one workgroup per CU
                                when the column is page-locked host memory */
  uint32_t radix_bits;       /* partitioning: log2 fan-out, 1..16 (default 10) */
  uint32_t passes;           /* partitioning: 1 or 2 (default 1); above 12 bits always 2 */
  uint32_t ignore_bits;      /* partitioning: low key bits skipped; 0xFFFFFFFF = derive from the max key
                                (partitioning_approach_dynamic_bit_range, src/index_lookup_new.cu:63-67) */
  uint64_t window;           /* partitioning: probes per window (default 2^24) */
} gj_lookup_params;

void gj_lookup_params_default(gj_lookup_params* p);

typedef struct gj_index gj_index; /* opaque; owns index-private HBM, borrows the column */

/* Build an index over a sorted column.
 * Replaces abstract_index<Key>::construct(h_column, d_column)
 * (src/indexes.cuh:36) with the build_index<Key, IndexT> dispatch
 * (src/index_lookup_common.cuh:208-215, src/index_lookup_new.cu:83-106).
 *   h_sorted_keys: optional host copy of the column (NULL: copied from the device when needed)
 *   d_sorted_keys: the device column, key_bytes 4 (uint32) or 8 (uint64), ascending;
 *                  strictly ascending for every kind except GJ_BINARY_SEARCH and GJ_NO_OP
 *   n:             keys in the column, 1 <= n < 2^32 - 1
 * GJ_HARMONIA requires n to be a multiple of 32 (src/harmonia.cuh:350-352).
 * Synchronises the stream before returning (construction is host-driven, as in the reference). */
int gj_index_build(gj_index_kind kind, const void* h_sorted_keys, const void* d_sorted_keys, uint64_t n,
                   int key_bytes, const gj_index_params* params, gj_stream_t stream, gj_index** out);

/* Batched lookup: d_rowids[i] = row id of d_probe_keys[i] or GJ_NOT_FOUND.
 * Replaces lookup_kernel<<<...>>>(index.device_index, ...) as driven by
 * plain_approach::run (src/index_lookup_new.cu:188-199), bws
 * (src/index_lookup_new.cu:227-249), partitioning_approach::run
 * (src/index_lookup_partitioning.cu:493-507) and hj_approach::run
 * (src/index_lookup_new.cu:262-301).  Probe keys have the index's key width.
 * m may be 0.  Stream-ordered.  The partitioning and hj drivers use device
 * scratch that the index owns: one buffer set per stream, grown on first use
 * and kept until gj_index_destroy, so lookups on one index from different
 * streams may run concurrently (calls on one stream are ordered anyway). */
int gj_index_lookup(const gj_index* index, const void* d_probe_keys, uint64_t m, uint32_t* d_rowids,
                    const gj_lookup_params* params, gj_stream_t stream);

/* Index-private device bytes.  Replaces abstract_index::memory_consumption()
 * (src/indexes.cuh:38). */
uint64_t gj_index_memory(const gj_index* index);

/* Descriptive statistics of a built index (no reference counterpart). */
typedef struct gj_index_stats {
  gj_index_kind kind;
  int key_bytes;
  uint64_t n;
  uint64_t memory_bytes;
  uint32_t depth;             /* tree levels (B-tree inner levels, Harmonia levels) */
  uint32_t lds_levels;        /* upper levels resident in LDS during lookups */
  uint64_t spline_points;     /* RadixSpline knots; learned_plr spline-root knots */
  uint64_t leaves;            /* learned leaves */
  uint64_t fallback_leaves;   /* learned leaves routed to the B-tree */
  uint64_t fallback_keys;     /* build keys whose leaf is routed to the B-tree */
  uint32_t max_window;        /* widest last-mile window over non-fallback leaves (learned), 2*err+2 (RadixSpline) */
  double avg_window;          /* mean last-mile window over build keys (learned) */
  float train_loss;           /* final MLP training loss (learned_mlp) */
  uint32_t root_segments;     /* learned_plr: segments of the spline root (0: linear root) */
  uint32_t split_leaves;      /* learned: leaves too wide for one line, split into 16 sub-leaves */
  uint64_t dropped_tuples;    /* probes the partitioned driver's scatter could not place (0 unless an
                                 internal fault occurred; their row ids were not written) */
  uint32_t layout;            /* learned: GJ_LAYOUT_COLUMN or GJ_LAYOUT_LINES */
  uint32_t max_lines;         /* lines layout: most lines a present key's lookup reads (1 + the largest spill) */
  uint64_t lines;             /* lines layout: 128-byte lines held */
  double avg_lines;           /* lines layout: mean lines a present key's lookup reads */
  uint32_t root_trained;      /* learned_mlp: 1 when the trained parameters serve lookups; 0 when the build
                                 ran untrained (GJ_LEARNED_UNTRAINED) or, in the lines layout, kept the
                                 quantile initialisation because it placed the keys in fewer lines */
  uint32_t placement;         /* GJ_PLACE_DEVICE or GJ_PLACE_HOST: where the index structures live */
} gj_index_stats;

/* Fills *out.  dropped_tuples is read from the device: the call waits for the
 * work already queued on the index's lookups. */
int gj_index_get_stats(const gj_index* index, gj_index_stats* out);

void gj_index_destroy(gj_index* index);

/* Name -> kind, following parse_index_type (src/indexes.cpp:5-19). */
int gj_parse_index_type(const char* name, gj_index_kind* out);
const char* gj_index_type_name(gj_index_kind kind);

/* ---- Radix partitioner ------------------------------------------------
 * Replaces the fast-interconnects prefix-sum + LA-SWWC scatter pair driven
 * by run_on_stream (src/index_lookup_partitioning.cu:300-366; declarations
 * src/gpu_radix_partition.cuh:8-45) and the partition_offsets /
 * partitioned_relation buffers (src/partitioned_relation.hpp:12-50).
 * Partition of key k = (k >> ignore_bits) & (2^radix_bits - 1).
 * Output is structure-of-arrays: partition p holds d_keys[d_offsets[p] ..
 * d_offsets[p] + d_counts[p]) and, for every tuple, the probe's original
 * index in d_payloads.  Each partition starts on a 128-byte boundary of
 * d_keys (the reference pads each partition by GPU_CACHE_LINE_SIZE,
 * src/partitioned_relation.hpp:7-10).  Tuple order inside a partition
 * follows the input order (stable). */
typedef struct gj_partition_params {
  uint32_t radix_bits;  /* 1..16 */
  uint32_t ignore_bits; /* < key bits */
  uint32_t passes;      /* 1 or 2 (2: a pass on the high ceil(radix_bits/2) bits, then the full digit).
                           Above 12 bits always two: the high ceil(radix_bits/2) bits, then each of those
                           partitions split on the low floor(radix_bits/2) bits (the reference's limit at
                           48 KiB of shared memory is 8 bits per pass, src/index_lookup_partitioning.cu:50-52) */
} gj_partition_params;

typedef struct gj_partitioned {
  void* d_keys;          /* capacity keys of the input key width */
  uint32_t* d_payloads;  /* capacity original indices */
  uint64_t* d_offsets;   /* 2^radix_bits partition starts */
  uint64_t* d_counts;    /* 2^radix_bits partition sizes */
  uint64_t capacity;     /* tuples d_keys / d_payloads can hold: >= gj_partition_capacity(m, ...) */
  uint64_t* d_dropped;   /* optional device counter: tuples whose store would fall past `capacity` are not
                            stored but added here (0 for any consistent partitioning; a non-zero value
                            reports an internal fault instead of a silently missing tuple) */
} gj_partitioned;

uint64_t gj_partition_capacity(uint64_t m, int key_bytes, uint32_t radix_bits);
/* Scratch bytes gj_partition needs (histograms, second-pass buffers). */
uint64_t gj_partition_scratch_bytes(uint64_t m, int key_bytes, const gj_partition_params* p);
int gj_partition(const void* d_keys, uint64_t m, int key_bytes, const gj_partition_params* params,
                 gj_partitioned* out, void* d_scratch, uint64_t scratch_bytes, gj_stream_t stream);

/* ---- Fact-dimension join-aggregate -------------------------------------
 * Replaces the Q14 probe kernels (ij_plain_kernel src/tpch/tpch_14_ij.cu:34-85
 * with the correct predicate of hj_probe_kernel src/tpch/tpch_14_hj.cu:38-58,
 * partitioned_ij_lookup src/tpch/tpch_14_ij_partitioning.cu:187-237) and
 * process_results (src/tpch/tpch_14_common.cu:116-132).
 * For each fact row with date_lo <= date < date_hi whose fkey is found in the
 * index over the dimension keys: s = price * (100 - discount) (int64);
 * denominator += s; numerator += s when d_flag[rowid] != 0.
 * The index must have key_bytes 4 (fkeys are read as uint32).  Results are
 * ADDED to d_result (the caller zeroes it), so several batches accumulate.
 * Workgroup partial sums go to scratch that the index owns per stream (64 KiB
 * at 256 CUs, allocated on a stream's first call), so calls from different
 * streams may overlap. */
typedef struct gj_fact_batch {
  const int32_t* d_fkey;
  const int32_t* d_date;
  const int64_t* d_price;
  const int8_t* d_discount;
  uint64_t rows;
} gj_fact_batch;

/* How the qualifying rows meet the index (the reference's Q14 approaches,
 * src/tpch/tpch_14_common.cu:748-755).  Every strategy gives the same sums.
 *   GJ_AGG_LANE_REFILL  qualifying rows queue per wave in LDS and are looked up 64 at a time,
 *                       fused with the sums (ij_np_lf_lr_plain's lane refill,
 *                       src/tpch/tpch_14_ij.cu:464-566, without materialising join entries)
 *   GJ_AGG_PLAIN        each lane looks up its own qualifying rows in place (ij_plain_kernel,
 *                       src/tpch/tpch_14_ij.cu:34-85, with the correct date predicate)
 *   GJ_AGG_PARTITIONED  compact the qualifying (fkey, summand) pairs, radix-partition them on the
 *                       fkey, then look the partitions up in key order and sum (ij_partitioning,
 *                       src/tpch/tpch_14_common.cu:651-704, src/tpch/tpch_14_ij_partitioning.cu:187-237);
 *                       synchronises the stream once, after the filter counts the qualifying rows
 *   GJ_AGG_SORTED       each workgroup buffers its qualifying rows in LDS (2048 slots), sorts them by
 *                       fkey and looks them up in key order, 64 neighbouring keys per wave step
 *                       (ij_pbws, the pipelined blockwise sort, src/tpch/tpch_14_ij.cu:93-362)
 *   GJ_AGG_LOOKUP_FIRST look up every qualifying row in place and write the matches as join entries
 *                       (fact row, dimension row), then a second kernel gathers price, discount and
 *                       flag per entry and sums (ij_np_lf_plain: ij_lookup_kernel + ij_join_kernel,
 *                       src/tpch/tpch_14_ij.cu:371-461, src/tpch/tpch_14_common.cu:302-335); at most
 *                       2^32 - 1 fact rows; synchronises the stream once, like GJ_AGG_PARTITIONED */
typedef enum gj_agg_strategy {
  GJ_AGG_LANE_REFILL = 0,
  GJ_AGG_PLAIN = 1,
  GJ_AGG_PARTITIONED = 2,
  GJ_AGG_SORTED = 3,
  GJ_AGG_LOOKUP_FIRST = 4
} gj_agg_strategy;

typedef struct gj_agg_params {
  int32_t date_lo;        /* inclusive */
  int32_t date_hi;        /* exclusive */
  const uint8_t* d_flag;  /* per row of the sorted dimension: 1 = promo */
  uint32_t strategy;      /* gj_agg_strategy, default GJ_AGG_LANE_REFILL (the only one batch plans run) */
  uint32_t radix_bits;    /* GJ_AGG_PARTITIONED: log2 fan-out over the fkey's top bits, 1..16; 0 = 11, the
                             reference's (src/tpch/tpch_14_common.cu:657) */
} gj_agg_params;

typedef struct gj_agg_result {
  int64_t numerator;   /* sum over flagged rows */
  int64_t denominator; /* sum over all matching rows */
  uint64_t matches;    /* rows that passed the filter and found their dimension row */
  uint64_t pad;
} gj_agg_result;

int gj_join_aggregate(const gj_index* index, const gj_fact_batch* batch, const gj_agg_params* params,
                      gj_agg_result* d_result, gj_stream_t stream);

/* gj_join_aggregate for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, semantics and results as
 * gj_join_aggregate (GJ_AGG_PARTITIONED and GJ_AGG_LOOKUP_FIRST, which need scratch, are rejected with
 * GJ_ERR_UNSUPPORTED), but the call allocates nothing: the workgroup partial sums
 * live in a static device array of 4096 slots (128 KiB per GPU, held until the
 * process ends) instead of index-owned scratch.  A launch uses at most 8
 * workgroups per CU, so devices with more than 512 CUs are rejected with
 * GJ_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (on different streams) race and give wrong sums. */
int gj_join_aggregate_noalloc(const gj_index* index, const gj_fact_batch* batch, const gj_agg_params* params,
                              gj_agg_result* d_result, gj_stream_t stream);

/* Final fixed-point value 100 * (num * 1000) / (den / 1000), i.e. percent x 1e6
 * (process_results, src/tpch/tpch_14_common.cu:124-127).  Returns 0 when den / 1000 == 0. */
int64_t gj_promo_revenue(int64_t numerator, int64_t denominator);

/* ---- Batched execution (uneven batches against one resident index) -------
 * No reference counterpart: the reference launches one kernel per run.
 * A plan holds a fixed list of batches; gj_batch_plan_run executes all of
 * them, by one of three modes:
 *   GJ_BATCH_LAUNCH      one kernel launch per batch, spread over the caller's stream and
 *                        3 side streams the plan owns (forked from and joined back into
 *                        the caller's stream on every run)
 *   GJ_BATCH_GRAPH       the same launches captured once into a hipGraph, replayed per run
 *   GJ_BATCH_PERSISTENT  one persistent kernel; workgroups pull (batch, rows) work items
 *                        from a device counter until every batch is done
 * A run is ordered on the caller's stream as a whole; batches may overlap within it.
 * Lookup plans write rowids per batch; aggregate plans set one gj_agg_result per batch
 * (zeroed at the start of each run). */
typedef enum gj_batch_mode { GJ_BATCH_LAUNCH = 0, GJ_BATCH_GRAPH = 1, GJ_BATCH_PERSISTENT = 2 } gj_batch_mode;

typedef struct gj_probe_batch {
  const void* d_probe_keys;
  uint32_t* d_rowids;
  uint64_t m;
} gj_probe_batch;

typedef struct gj_batch_plan gj_batch_plan;

int gj_lookup_plan_create(const gj_index* index, const gj_probe_batch* h_batches, uint32_t num_batches,
                          gj_batch_mode mode, gj_stream_t stream, gj_batch_plan** out);
int gj_aggregate_plan_create(const gj_index* index, const gj_fact_batch* h_batches, uint32_t num_batches,
                             const gj_agg_params* params, gj_agg_result* d_results, gj_batch_mode mode,
                             gj_stream_t stream, gj_batch_plan** out);
int gj_batch_plan_run(gj_batch_plan* plan, gj_stream_t stream);
void gj_batch_plan_destroy(gj_batch_plan* plan);

/* ---- Autotuner (no reference counterpart) ------------------------------
 * Times candidate lookup configurations (approach, block size, partition
 * fan-out, window) on the given probe sample and writes the fastest into
 * *best.  Synchronises the stream. */
int gj_autotune_lookup(const gj_index* index, const void* d_probe_sample, uint64_t m, uint32_t* d_rowids_scratch,
                       gj_lookup_params* best, double* best_lookups_per_s, gj_stream_t stream);

/* Index autotuner (no reference counterpart; the reference fixes these at
 * compile time).  Builds `kind` over the column once per candidate of its own
 * parameters, times the plain lookup of the probe sample on each, and writes
 * the fastest parameter set into *best.  Candidates, starting from *base (or
 * the defaults when base is NULL):
 *   harmonia      harmonia_group_lanes 8 / 16 (/ 32 for 4-byte keys): threads per lookup
 *   btree         btree_lds_bytes 4096 / 16384 / 32768: the LDS-resident upper levels
 *   radix_spline  rs_radix_bits 12 / 13 (table in LDS) / 16 / 18 (global) x rs_max_error 16 / 32 / 64
 *   learned_plr   learned_leaf_bits 0 (the default: compact 8-byte leaves where the build
 *                 admits them), then 16-byte leaves at default - 1 / default / default + 1
 *                 bits x learned_max_error 64 / 128 / 256
 *   learned_mlp   learned_leaf_bits 0, then default - 1 / default / default + 1
 *   other kinds   *base only
 * Candidates the column does not admit are skipped.  The candidate indexes
 * are destroyed; build the chosen one with gj_index_build.  Synchronises the stream. */
int gj_autotune_index(gj_index_kind kind, const void* h_sorted_keys, const void* d_sorted_keys, uint64_t n,
                      int key_bytes, const gj_index_params* base, const void* d_probe_sample, uint64_t m,
                      uint32_t* d_rowids_scratch, gj_index_params* best, double* best_lookups_per_s,
                      gj_stream_t stream);

/* ---- Learned-model kernels exposed for gradient checks (no reference counterpart)
 * MLP root f(x) = b2 + sum_k w2_k * ramp(w1_k * x + b1_k), ramp(a) = min(max(a, 0), 1),
 * hidden width h (8, 16, 32 or 64), params laid out [w1(h), b1(h), w2(h), b2(1)] in
 * FP32.  Loss = mean((f(x) - y)^2).
 * Writes the loss and d loss / d params (3h + 1 floats). */
int gj_mlp_loss_grad(const float* d_params, uint32_t hidden, const float* d_x, const float* d_y, uint64_t n,
                     float* d_grad, float* d_loss, gj_stream_t stream);
/* Forward pass only: d_out[i] = f(d_x[i]). */
int gj_mlp_forward(const float* d_params, uint32_t hidden, const float* d_x, uint64_t n, float* d_out,
                   gj_stream_t stream);

/* ---- Microbenchmarks: the random-gather roofline probe ------------------
 * Successor of random_access<512> (src/random_access.cu:53-66): for each of
 * m slots, read d_probes[i] (8 B, streamed), gather
 * d_array[murmur3_64(d_probes[i]) & (n_pow2 - 1)] (8 B) and write its low
 * 32 bits to d_out[i] (4 B).  n_pow2 must be a power of two. */
int gj_gather(const uint64_t* d_array, uint64_t n_pow2, const uint64_t* d_probes, uint64_t m, uint32_t* d_out,
              uint32_t block_size, uint32_t items_per_thread, gj_stream_t stream);
/* Dependent chain: cur = d_next[cur] for steps steps from `chains` independent
 * starts; writes the final positions to d_out (chains entries).  Chain c starts
 * at d_start[c] when d_start is given (it may be d_out itself, so that each
 * launch continues where the previous one stopped and never re-walks lines a
 * cache may still hold), else at (c * 0x9E3779B97F4A7C15) mod n.  Every d_next
 * entry and start must be < n. */
int gj_pointer_chase(const uint32_t* d_next, uint64_t n, uint32_t chains, uint32_t steps, const uint32_t* d_start,
                     uint32_t* d_out, gj_stream_t stream);
/* Streaming copy of bytes (multiple of 16) from d_src to d_dst. */
int gj_stream_copy(const void* d_src, void* d_dst, uint64_t bytes, gj_stream_t stream);
/* Streaming read of bytes (multiple of 16, 16-byte aligned) from d_src: XORs every
 * 8-byte word into *d_xor (which the caller sets first), the bandwidth bound of a
 * read-only kernel such as the join-aggregate. */
int gj_stream_read(const void* d_src, uint64_t bytes, uint64_t* d_xor, gj_stream_t stream);

/* ---- Diagnostics --------------------------------------------------------- */
const char* gj_last_error(void);
const char* gj_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GJ_H */
