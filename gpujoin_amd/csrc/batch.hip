// batch.hip - fact-dimension join-aggregate and batched execution plans.
//
// gj_join_aggregate replaces the Q14 probe kernels: the filter / lookup /
// aggregate of hj_probe_kernel (src/tpch/tpch_14_hj.cu:21-71, the correct
// shipdate predicate lo <= d < hi, :39-42) and ij_plain_kernel
// (src/tpch/tpch_14_ij.cu:34-85), with the lane-refill idea of
// ij_np_lane_refill_scan_lookup (src/tpch/tpch_14_ij.cu:464-566): qualifying
// rows are queued per wave in LDS and looked up 64 at a time, so the
// cooperative index lookup always runs on a full wave although only ~1 % of
// the fact rows pass the filter.  Only the date column is read for every row;
// fkey, price and discount are read for qualifying rows only.  Sums are exact
// int64: wave shuffle and workgroup reductions write one partial per
// workgroup, and a one-block kernel adds the partials to the result (64-bit
// atomics on one address from every wave would serialise).
//
// Plans run a fixed list of uneven batches: one launch per batch, the same
// launches captured once into a hipGraph, or one persistent kernel whose
// workgroups pull (batch, rows) work items from a device counter until all are
// done.
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "lookup_dispatch.hpp"

namespace gj {

constexpr int kAggBlock = 256;
// 64-row slices per wave step (the next step's dates load behind this step's columns):
// 4 for both kernels (one batch 547 G rows/s against 459 with 2 and 540 with 8; the
// persistent kernel 454 against 370 with 8)
constexpr int kAggUnrollBatch = 4;
constexpr int kAggUnrollTile = 4;
// queued qualifying rows per wave: a full step plus fewer than 64 left over
constexpr int agg_queue_rows(int unroll) { return kWave * (unroll + 1); }
constexpr uint64_t kAggWaves = kAggBlock / kWave;  // partial sums per persistent work item
constexpr uint64_t kAggDirectBlocks = 128;         // grids up to this size skip the partial-sum kernel
constexpr int kGraphLanes = 4;  // streams a graph plan's captured launches are spread over
constexpr uint64_t kAggTileRows = 65536;    // fact rows per persistent work item (one workgroup)
// probes per persistent work item (one workgroup): small items keep the last
// round of items short (65536 left about 2.7 items per workgroup and lost 6 %)
constexpr uint64_t kLookupTileRows = 4096;

struct FactBatchDev {
  const int32_t* fkey;
  const int32_t* date;
  const int64_t* price;
  const int8_t* disc;
  uint64_t rows;
  uint64_t tile_start;  // first work item of this batch (persistent plans)
  gj_agg_result* result;
};

struct ProbeBatchDev {
  const void* probes;
  uint32_t* out;
  uint64_t m;
  uint64_t tile_start;
};

struct AggAcc {
  int64_t num = 0, den = 0;
  uint64_t matches = 0;
};

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Per-wave LDS queue of qualifying rows (dimension key, summand).
struct AggQueue {
  uint32_t* key;
  int64_t* val;
  int count;  // wave-uniform
};

// Looks up the 64 (or fewer, at a flush) newest queued rows on the full wave.
template <class Dev>
__device__ __forceinline__ void agg_drain(const Dev& d, const uint8_t* lds, AggQueue& q, int take,
                                          const uint8_t* __restrict__ flag, AggAcc& acc) {
  const int ln = lane();
  const bool active = ln < take;
  const int slot = q.count - take + ln;
  const uint32_t k = active ? q.key[slot] : 0u;
  const int64_t s = active ? q.val[slot] : 0;
  const uint32_t r = d.lookup(k, active, lds);
  if (active && r != kNotFound) {
    acc.den += s;
    if (flag[r]) acc.num += s;
    acc.matches += 1;
  }
  q.count -= take;
}

// Rows of one fact batch: the wave takes U x 64 rows per step (row
// base + k * 64 + lane), starting at r0 and advancing by `step`, below r1.  All
// date loads of a step are issued together, then the fkey / price / discount
// loads of every qualifying row, and right behind those the next step's date
// loads, so the two memory latencies of consecutive steps overlap; qualifying rows
// queue in LDS and are looked up 64 at a time.
template <int U, class Dev>
__device__ __forceinline__ void agg_rows(const Dev& d, const uint8_t* lds, AggQueue& q, const FactBatchDev& b,
                                         uint64_t r0, uint64_t r1, uint64_t step, int32_t lo, int32_t hi,
                                         const uint8_t* __restrict__ flag, AggAcc& acc) {
  const int ln = lane();
  const uint64_t lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  int32_t dt[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t i = r0 + static_cast<uint64_t>(k * kWave + ln);
    dt[k] = i < r1 ? load_stream(b.date + i) : 0;
  }
  for (uint64_t base = r0; base < r1; base += step) {
    uint32_t qm = 0;  // bit k: row of slice k qualifies
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
      qm |= ((i < r1 && dt[k] >= lo && dt[k] < hi) ? 1u : 0u) << k;
    }
    const bool any = ballot(qm != 0) != 0;  // wave-uniform
    uint32_t fk[U];
    int64_t pr[U];
    int32_t dc[U];
    if (any) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        fk[k] = 0;
        pr[k] = 0;
        dc[k] = 0;
        if ((qm >> k) & 1u) {
          const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
          fk[k] = static_cast<uint32_t>(b.fkey[i]);
          pr[k] = b.price[i];
          dc[k] = b.disc[i];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {  // the next step's dates, behind this step's column loads
      const uint64_t i = base + step + static_cast<uint64_t>(k * kWave + ln);
      dt[k] = i < r1 ? load_stream(b.date + i) : 0;
    }
    if (!any) continue;
    int64_t sv[U];
#pragma unroll
    for (int k = 0; k < U; ++k) sv[k] = pr[k] * (100 - static_cast<int64_t>(dc[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool qual = (qm >> k) & 1u;
      const uint64_t mask = ballot(qual);
      if (qual) {
        const int pos = q.count + __popcll(mask & lt);
        q.key[pos] = fk[k];
        q.val[pos] = sv[k];
      }
      q.count += __popcll(mask);
    }
    __builtin_amdgcn_wave_barrier();
    while (q.count >= kWave) agg_drain(d, lds, q, kWave, flag, acc);
  }
}

// The plain strategy (ij_plain_kernel, src/tpch/tpch_14_ij.cu:34-85): each lane
// takes its own rows and looks up the qualifying ones in place; the wave's lookup
// runs with only the qualifying lanes active (about 1 % of them at the Q14 window).
template <int U, class Dev>
__device__ __forceinline__ void agg_rows_plain(const Dev& d, const uint8_t* lds, const FactBatchDev& b, uint64_t r0,
                                               uint64_t r1, uint64_t step, int32_t lo, int32_t hi,
                                               const uint8_t* __restrict__ flag, AggAcc& acc) {
  // (loading the next step's dates ahead here, as agg_rows does, cost 10 %: the per-slice
  // lookups wait on them)
  const int ln = lane();
  for (uint64_t base = r0; base < r1; base += step) {
    int32_t dt[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
      dt[k] = i < r1 ? load_stream(b.date + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
      const bool qual = i < r1 && dt[k] >= lo && dt[k] < hi;
      if (ballot(qual) == 0) continue;  // wave-uniform
      uint32_t fk = 0;
      int64_t sv = 0;
      if (qual) {
        fk = static_cast<uint32_t>(b.fkey[i]);
        sv = b.price[i] * (100 - static_cast<int64_t>(b.disc[i]));
      }
      const uint32_t r = d.lookup(fk, qual, lds);
      if (qual && r != kNotFound) {
        acc.den += sv;
        if (flag[r]) acc.num += sv;
        acc.matches += 1;
      }
    }
  }
}

// Per-workgroup sums of the one-batch kernel and per-wave sums of the persistent
// one (no result atomics; see agg_sum_kernel).
struct AggPartial {
  int64_t num, den;
  uint64_t matches, pad;
};

template <class Dev>
__device__ __forceinline__ void agg_flush(const Dev& d, const uint8_t* lds, AggQueue& q,
                                          const uint8_t* __restrict__ flag, AggAcc& acc, gj_agg_result* out) {
  while (q.count > 0) agg_drain(d, lds, q, q.count < kWave ? q.count : kWave, flag, acc);
  const int64_t num = wave_sum_i64(acc.num);
  const int64_t den = wave_sum_i64(acc.den);
  const int64_t mt = wave_sum_i64(static_cast<int64_t>(acc.matches));
  if (lane() == 0 && (num | den | mt) != 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->numerator), static_cast<unsigned long long>(num));
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->denominator), static_cast<unsigned long long>(den));
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->matches), static_cast<unsigned long long>(mt));
  }
  acc = AggAcc{};
}

template <int U>
__device__ __forceinline__ AggQueue wave_queue(uint8_t* lds_tail) {
  constexpr int rows = agg_queue_rows(U);
  const int w = threadIdx.x / kWave;
  AggQueue q;
  q.key = reinterpret_cast<uint32_t*>(lds_tail) + w * rows;
  q.val = reinterpret_cast<int64_t*>(lds_tail + (kAggBlock / kWave) * rows * 4) + w * rows;
  q.count = 0;
  return q;
}

inline size_t agg_queue_bytes(int unroll) { return (kAggBlock / kWave) * agg_queue_rows(unroll) * (4 + 8); }

// One batch, grid-stride over 64-row steps.  Each workgroup writes its sums to
// partial[blockIdx.x]; agg_sum_kernel adds them to the result.  (One 64-bit atomic
// per wave onto the three result words serialises at their L2 channel: at ~4000
// waves that took about half of the kernel's time.  Letting the last workgroup
// add the partials instead needs a device-scope fence in every workgroup, whose
// L2 write-back cost more than the second launch.)
template <class Dev, bool Plain>
__global__ __launch_bounds__(kAggBlock) void join_agg_kernel(Dev d, FactBatchDev b, int32_t lo, int32_t hi,
                                                             const uint8_t* __restrict__ flag,
                                                             AggPartial* __restrict__ partial, bool direct) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ int64_t red[kAggBlock / kWave][3];
  d.stage(lds);
  const size_t ilds = (d.lds_bytes() + 15) / 16 * 16;
  AggQueue q = wave_queue<kAggUnrollBatch>(lds + ilds);
  __syncthreads();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const int w = threadIdx.x / kWave;
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + w;
  AggAcc acc;
  const uint64_t span = static_cast<uint64_t>(kWave) * kAggUnrollBatch;
  if constexpr (Plain) {
    agg_rows_plain<kAggUnrollBatch>(d, lds, b, wave * span, b.rows, waves * span, lo, hi, flag, acc);
  } else {
    agg_rows<kAggUnrollBatch>(d, lds, q, b, wave * span, b.rows, waves * span, lo, hi, flag, acc);
    while (q.count > 0) agg_drain(d, lds, q, q.count < kWave ? q.count : kWave, flag, acc);
  }
  const int64_t num = wave_sum_i64(acc.num);
  const int64_t den = wave_sum_i64(acc.den);
  const int64_t mt = wave_sum_i64(static_cast<int64_t>(acc.matches));
  if (lane() == 0) {
    red[w][0] = num;
    red[w][1] = den;
    red[w][2] = mt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    AggPartial p{0, 0, 0, 0};
    for (int v = 0; v < kAggBlock / kWave; ++v) {
      p.num += red[v][0];
      p.den += red[v][1];
      p.matches += static_cast<uint64_t>(red[v][2]);
    }
    if (!direct) {
      partial[blockIdx.x] = p;
    } else if ((p.num | p.den | static_cast<int64_t>(p.matches)) != 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&b.result->numerator), static_cast<unsigned long long>(p.num));
      atomicAdd(reinterpret_cast<unsigned long long*>(&b.result->denominator), static_cast<unsigned long long>(p.den));
      atomicAdd(reinterpret_cast<unsigned long long*>(&b.result->matches), static_cast<unsigned long long>(p.matches));
    }
  }
}

// The blockwise-sorted strategy (ij_pbws, src/tpch/tpch_14_ij.cu:93-362): a
// workgroup collects its qualifying rows (dimension key, summand) in an LDS buffer
// of kAggSortTile slots, sorts the buffered keys (bitonic network over
// key << 32 | slot), and looks them up in key order, 64 neighbouring keys per wave
// step, so a wave's lanes walk neighbouring index paths and share lines; then it
// refills.  Rows are taken as by join_agg_kernel (64 x kAggSortUnroll per wave
// step, grid stride).  Slots are claimed with one LDS atomic per wave and 64-row
// slice.  A wave starts a step only while at most kAggSortTile - (waves per
// workgroup) x (rows per step) slots are taken: at most one step per wave is then
// in flight past that point, so the buffer never overflows.
constexpr int kAggSortTile = 2048;  // 4096: 278 against 396 G rows/s (half the workgroups per CU)
constexpr int kAggSortUnroll = 4;

template <class Dev>
__global__ __launch_bounds__(kAggBlock) void join_agg_sorted_kernel(Dev d, FactBatchDev b, int32_t lo, int32_t hi,
                                                                    const uint8_t* __restrict__ flag,
                                                                    AggPartial* __restrict__ partial, bool direct) {
  constexpr int U = kAggSortUnroll;
  constexpr int kWaves = kAggBlock / kWave;
  constexpr int kSpan = kWave * U;
  static_assert(kWaves * kSpan <= kAggSortTile / 2, "a fill round must fit the buffer's second half");
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ int64_t red[kWaves][3];
  __shared__ int taken;
  d.stage(lds);
  const size_t ilds = (d.lds_bytes() + 15) / 16 * 16;
  uint64_t* ent = reinterpret_cast<uint64_t*>(lds + ilds);                      // key << 32 | slot
  int64_t* val = reinterpret_cast<int64_t*>(lds + ilds + kAggSortTile * 8);     // summand by slot
  if (threadIdx.x == 0) taken = 0;
  __syncthreads();
  const int ln = lane();
  const uint64_t lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  const int w = threadIdx.x / kWave;
  const uint64_t step = static_cast<uint64_t>(gridDim.x) * kWaves * kSpan;
  uint64_t base = (static_cast<uint64_t>(blockIdx.x) * kWaves + w) * kSpan;
  AggAcc acc;
  int32_t dt[U];  // the dates of the step at `base`, loaded one step ahead
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
    dt[k] = i < b.rows ? load_stream(b.date + i) : 0;
  }
  for (;;) {
    while (base < b.rows &&
           __hip_atomic_load(&taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) <= kAggSortTile - kWaves * kSpan) {
      uint32_t qm = 0;
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
        qm |= ((i < b.rows && dt[k] >= lo && dt[k] < hi) ? 1u : 0u) << k;
      }
      const bool any = ballot(qm != 0) != 0;  // wave-uniform
      uint32_t fk[U];
      int64_t pr[U];
      int32_t dc[U];
      if (any) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
          fk[k] = 0;
          pr[k] = 0;
          dc[k] = 0;
          if ((qm >> k) & 1u) {
            const uint64_t i = base + static_cast<uint64_t>(k * kWave + ln);
            fk[k] = static_cast<uint32_t>(b.fkey[i]);
            pr[k] = b.price[i];
            dc[k] = b.disc[i];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {  // the next step's dates, behind this step's column loads
        const uint64_t i = base + step + static_cast<uint64_t>(k * kWave + ln);
        dt[k] = i < b.rows ? load_stream(b.date + i) : 0;
      }
      if (any) {
        int64_t sv[U];
#pragma unroll
        for (int k = 0; k < U; ++k) sv[k] = pr[k] * (100 - static_cast<int64_t>(dc[k]));
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const bool qual = (qm >> k) & 1u;
          const uint64_t mask = ballot(qual);
          if (mask == 0) continue;  // wave-uniform
          int first = 0;
          if (ln == 0) first = atomicAdd(&taken, __popcll(mask));
          first = __shfl(first, 0, kWave);
          if (qual) {
            const int slot = first + __popcll(mask & lt);
            ent[slot] = (static_cast<uint64_t>(fk[k]) << 32) | static_cast<uint32_t>(slot);
            val[slot] = sv[k];
          }
        }
      }
      base += step;
    }
    __syncthreads();  // every wave has stopped filling: `taken` is final for this round
    const int total = taken;
    const bool more = __syncthreads_or(base < b.rows) != 0;
    if (total > 0) {  // workgroup-uniform
      int n2 = kWave;
      while (n2 < total) n2 <<= 1;
      for (int t = total + threadIdx.x; t < n2; t += kAggBlock) ent[t] = ~0ull;
      __syncthreads();
      for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int t = threadIdx.x; t < n2 / 2; t += kAggBlock) {
            const int a = 2 * t - (t & (j - 1));  // lower element of the pair (a, a + j)
            const uint64_t ea = ent[a], eb = ent[a + j];
            if ((ea > eb) == ((a & k) == 0)) {
              ent[a] = eb;
              ent[a + j] = ea;
            }
          }
          __syncthreads();
        }
      }
      for (int t0 = w * kWave; t0 < total; t0 += kAggBlock) {  // wave-uniform bound
        const int t = t0 + ln;
        const bool active = t < total;
        const uint64_t e = active ? ent[t] : 0ull;
        const uint32_t key = static_cast<uint32_t>(e >> 32);
        const int64_t sv = active ? val[static_cast<uint32_t>(e)] : 0;
        const uint32_t r = d.lookup(key, active, lds);
        if (active && r != kNotFound) {
          acc.den += sv;
          if (flag[r]) acc.num += sv;
          acc.matches += 1;
        }
      }
      __syncthreads();  // the buffer is read out before it refills
      if (threadIdx.x == 0) taken = 0;
      __syncthreads();
    }
    if (!more) break;
  }
  const int64_t num = wave_sum_i64(acc.num);
  const int64_t den = wave_sum_i64(acc.den);
  const int64_t mt = wave_sum_i64(static_cast<int64_t>(acc.matches));
  if (ln == 0) {
    red[w][0] = num;
    red[w][1] = den;
    red[w][2] = mt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    AggPartial p{0, 0, 0, 0};
    for (int v = 0; v < kWaves; ++v) {
      p.num += red[v][0];
      p.den += red[v][1];
      p.matches += static_cast<uint64_t>(red[v][2]);
    }
    if (!direct) {
      partial[blockIdx.x] = p;
    } else if ((p.num | p.den | static_cast<int64_t>(p.matches)) != 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&b.result->numerator), static_cast<unsigned long long>(p.num));
      atomicAdd(reinterpret_cast<unsigned long long*>(&b.result->denominator), static_cast<unsigned long long>(p.den));
      atomicAdd(reinterpret_cast<unsigned long long*>(&b.result->matches), static_cast<unsigned long long>(p.matches));
    }
  }
}

// Adds the workgroups' partial sums to the result: one workgroup, exact int64.
__global__ __launch_bounds__(1024) void agg_sum_kernel(const AggPartial* __restrict__ partial, uint32_t blocks,
                                                       gj_agg_result* __restrict__ out) {
  __shared__ int64_t red[1024 / kWave][3];
  int64_t num = 0, den = 0, mt = 0;
  for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x) {
    num += partial[i].num;
    den += partial[i].den;
    mt += static_cast<int64_t>(partial[i].matches);
  }
  num = wave_sum_i64(num);
  den = wave_sum_i64(den);
  mt = wave_sum_i64(mt);
  const int w = threadIdx.x / kWave;
  if (lane() == 0) {
    red[w][0] = num;
    red[w][1] = den;
    red[w][2] = mt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 1; v < static_cast<int>(blockDim.x / kWave); ++v) {
      num += red[v][0];
      den += red[v][1];
      mt += red[v][2];
    }
    // results are added (the caller zeroes them); a single writer per call
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->numerator), static_cast<unsigned long long>(num));
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->denominator), static_cast<unsigned long long>(den));
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->matches), static_cast<unsigned long long>(mt));
  }
}

// Persistent join-aggregate: workgroups pull work items until the counter passes
// total_tiles.  Each wave's sums over an item go to its own partial slot (one
// owner, so a plain store); agg_batch_sum_kernel then adds each batch's slots to
// its result, so the waves issue no result atomics.
template <class Dev>
__global__ __launch_bounds__(kAggBlock) void join_agg_persistent_kernel(Dev d, const FactBatchDev* __restrict__ batches,
                                                                        const uint32_t* __restrict__ tile_batch,
                                                                        uint64_t total_tiles, int32_t lo,
                                                                        int32_t hi, const uint8_t* __restrict__ flag,
                                                                        unsigned long long* counter,
                                                                        AggPartial* __restrict__ partial) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  const size_t ilds = (d.lds_bytes() + 15) / 16 * 16;
  AggQueue q = wave_queue<kAggUnrollTile>(lds + ilds);
  __syncthreads();
  // A workgroup claims an item of kAggTileRows rows when it is free and its
  // waves interleave over it in 64 x kAggUnrollTile-row steps (claiming ahead would
  // leave the last items with busy workgroups; one stream per wave over its own
  // rows reads HBM worse than adjacent waves on adjacent rows).  tile_batch maps
  // an item to its batch.  Fewer than 64 queued rows stay queued into the next
  // item when it is of the same batch (their sums count in that item's
  // partials), so the part-empty lookup round that ends a queue is paid once
  // per batch change.  Wave w of item t writes partial[t * waves + w].
  constexpr int kWaves = kAggBlock / kWave;
  const int w = threadIdx.x / kWave;
  const uint64_t span = static_cast<uint64_t>(kWave) * kAggUnrollTile;
  __shared__ uint64_t claimed;
  auto claim = [&]() -> uint64_t {
    __syncthreads();  // every wave has read the previous item
    if (threadIdx.x == 0) claimed = atomicAdd(counter, 1ull);
    __syncthreads();
    return claimed;
  };
  uint64_t t = claim();
  while (t < total_tiles) {  // workgroup-uniform
    const uint32_t bi = tile_batch[t];
    const FactBatchDev b = batches[bi];
    const uint64_t r0 = (t - b.tile_start) * kAggTileRows;
    const uint64_t r1 = std::min<uint64_t>(b.rows, r0 + kAggTileRows);
    AggAcc acc;
    agg_rows<kAggUnrollTile>(d, lds, q, b, r0 + static_cast<uint64_t>(w) * span, r1, span * kWaves, lo, hi, flag, acc);
    const uint64_t tn = claim();
    if (tn >= total_tiles || tile_batch[tn] != bi)
      while (q.count > 0) agg_drain(d, lds, q, q.count < kWave ? q.count : kWave, flag, acc);
    const int64_t num = wave_sum_i64(acc.num);
    const int64_t den = wave_sum_i64(acc.den);
    const int64_t mt = wave_sum_i64(static_cast<int64_t>(acc.matches));
    if (lane() == 0) partial[t * kWaves + w] = AggPartial{num, den, static_cast<uint64_t>(mt), 0};
    t = tn;
  }
}

// One wave per batch: the sum of its items' per-wave partials, added to its result.
__global__ __launch_bounds__(256) void agg_batch_sum_kernel(const FactBatchDev* __restrict__ batches,
                                                            const uint64_t* __restrict__ tile_start, uint32_t nb,
                                                            uint64_t total_tiles,
                                                            const AggPartial* __restrict__ partial) {
  const uint64_t bw = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (bw >= nb) return;  // wave-uniform
  const uint64_t t0 = tile_start[bw];
  const uint64_t t1 = bw + 1 < nb ? tile_start[bw + 1] : total_tiles;
  int64_t num = 0, den = 0, mt = 0;
  for (uint64_t t = t0 * kAggWaves + static_cast<uint64_t>(lane()); t < t1 * kAggWaves; t += kWave) {
    num += partial[t].num;
    den += partial[t].den;
    mt += static_cast<int64_t>(partial[t].matches);
  }
  num = wave_sum_i64(num);
  den = wave_sum_i64(den);
  mt = wave_sum_i64(mt);
  if (lane() == 0 && (num | den | mt) != 0) {
    gj_agg_result* out = batches[bw].result;
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->numerator), static_cast<unsigned long long>(num));
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->denominator), static_cast<unsigned long long>(den));
    atomicAdd(reinterpret_cast<unsigned long long*>(&out->matches), static_cast<unsigned long long>(mt));
  }
}

// Persistent lookups over many probe batches.
template <class Dev, class K>
__global__ __launch_bounds__(256) void lookup_persistent_kernel(Dev d, const ProbeBatchDev* __restrict__ batches,
                                                                const uint32_t* __restrict__ tile_batch,
                                                                uint64_t total_tiles, unsigned long long* counter) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  d.stage(lds);
  __syncthreads();
  // as in the join-aggregate: a free workgroup claims an item and its waves
  // interleave over it 64 probes at a time
  const uint64_t w = threadIdx.x / kWave;
  const uint64_t step = blockDim.x;
  __shared__ uint64_t claimed;
  auto claim = [&]() -> uint64_t {
    __syncthreads();  // every wave has read the previous item
    if (threadIdx.x == 0) claimed = atomicAdd(counter, 1ull);
    __syncthreads();
    return claimed;
  };
  for (uint64_t t = claim(); t < total_tiles; t = claim()) {  // workgroup-uniform
    const ProbeBatchDev b = batches[tile_batch[t]];
    const K* probes = static_cast<const K*>(b.probes);
    const uint64_t r0 = (t - b.tile_start) * kLookupTileRows;
    const uint64_t r1 = std::min<uint64_t>(b.m, r0 + kLookupTileRows);
    for (uint64_t base = r0 + w * kWave; base < r1; base += step) {
      const uint64_t i = base + static_cast<uint64_t>(lane());
      const bool active = i < r1;
      const K x = active ? load_stream(probes + i) : K(0);
      const uint32_t r = d.lookup(x, active, lds);
      if (active) store_stream(b.out + i, r);
    }
  }
}

// ---- partitioned strategy (ij_partitioning, src/tpch/tpch_14_common.cu:651-704,
// src/tpch/tpch_14_ij_partitioning.cu:187-237): the qualifying rows are compacted to
// (fkey, summand) pairs, radix-partitioned on the fkey, and the partitions are looked
// up in key order, so neighbouring lanes and waves probe neighbouring index paths.

constexpr uint64_t kFilterChunk = 16384;  // fact rows per filter workgroup
// date loads in flight per thread and step: one at a time held the count at 3.3 TB/s
constexpr int kFilterIn = 4;

// counts[c] = qualifying rows of chunk c.
__global__ __launch_bounds__(kAggBlock) void agg_filter_count_kernel(const int32_t* __restrict__ date, uint64_t rows,
                                                                     int32_t lo, int32_t hi,
                                                                     uint32_t* __restrict__ counts) {
  __shared__ uint32_t wc[kAggBlock / kWave];
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * kFilterChunk;
  const uint64_t c1 = c0 + kFilterChunk < rows ? c0 + kFilterChunk : rows;
  uint32_t cnt = 0;
  for (uint64_t i0 = c0 + threadIdx.x; i0 < c1; i0 += kFilterIn * blockDim.x) {
    int32_t dt[kFilterIn];
#pragma unroll
    for (int u = 0; u < kFilterIn; ++u) {
      const uint64_t i = i0 + static_cast<uint64_t>(u) * blockDim.x;
      dt[u] = i < c1 ? load_stream(date + i) : 0;
    }
#pragma unroll
    for (int u = 0; u < kFilterIn; ++u) {
      const uint64_t i = i0 + static_cast<uint64_t>(u) * blockDim.x;
      cnt += (i < c1 && dt[u] >= lo && dt[u] < hi) ? 1u : 0u;
    }
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, kWave);
  if (lane() == 0) wc[threadIdx.x / kWave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kAggBlock / kWave; ++w) t += wc[w];
    counts[blockIdx.x] = t;
  }
}

// In place: counts[c] <- counts[0] + ... + counts[c - 1]; *total <- the sum.  One block.
__global__ __launch_bounds__(1024) void agg_scan_counts_kernel(uint32_t* __restrict__ counts, uint64_t n,
                                                               uint64_t* __restrict__ total) {
  __shared__ uint32_t wsum[1024 / kWave];
  const int w = threadIdx.x / kWave;
  const int ln = lane();
  uint64_t carry = 0;
  for (uint64_t t0 = 0; t0 < n; t0 += blockDim.x) {
    const uint64_t i = t0 + threadIdx.x;
    const uint32_t v = i < n ? counts[i] : 0u;
    uint32_t x = v;
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, kWave);
      if (ln >= off) x += y;
    }
    if (ln == kWave - 1) wsum[w] = x;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (int u = 0; u < static_cast<int>(blockDim.x / kWave); ++u) {
      before += u < w ? wsum[u] : 0u;
      all += wsum[u];
    }
    if (i < n) counts[i] = static_cast<uint32_t>(carry + before + x - v);
    carry += all;
    __syncthreads();  // wsum is rewritten by the next tile
  }
  if (threadIdx.x == 0) *total = carry;
}

// Writes chunk c's qualifying rows into positions [offsets[c], offsets[c + 1]):
// key = fkey, val = price * (100 - discount).  A step covers kFilterIn x 64 rows per wave
// (slice u = rows t0 + u * 64 + lane), with all its date loads in flight together.  The
// order inside a chunk is the order in which its waves reserve their slots (one LDS
// atomic per wave and step), which the partitioned strategy does not need: an ordered
// write, with a block-wide prefix and two barriers per step, ran at 3.2 TB/s.
__global__ __launch_bounds__(kAggBlock) void agg_filter_write_kernel(FactBatchDev b, int32_t lo, int32_t hi,
                                                                     const uint32_t* __restrict__ offsets,
                                                                     uint32_t* __restrict__ out_key,
                                                                     int64_t* __restrict__ out_val) {
  __shared__ uint32_t used;  // slots of the chunk taken so far
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * kFilterChunk;
  const uint64_t c1 = c0 + kFilterChunk < b.rows ? c0 + kFilterChunk : b.rows;
  const int ln = lane();
  const uint64_t lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  const uint32_t first = offsets[blockIdx.x];
  if (threadIdx.x == 0) used = 0;
  __syncthreads();
  const uint64_t waves = blockDim.x / kWave;
  const uint64_t wbase = c0 + (threadIdx.x / kWave) * (kFilterIn * kWave);
  for (uint64_t t0 = wbase; t0 < c1; t0 += waves * kFilterIn * kWave) {
    int32_t dt[kFilterIn];
#pragma unroll
    for (int u = 0; u < kFilterIn; ++u) {
      const uint64_t i = t0 + static_cast<uint64_t>(u * kWave + ln);
      dt[u] = i < c1 ? load_stream(b.date + i) : 0;
    }
    uint64_t mask[kFilterIn];
    uint32_t cnt = 0;
#pragma unroll
    for (int u = 0; u < kFilterIn; ++u) {
      const uint64_t i = t0 + static_cast<uint64_t>(u * kWave + ln);
      mask[u] = ballot(i < c1 && dt[u] >= lo && dt[u] < hi);
      cnt += static_cast<uint32_t>(__popcll(mask[u]));
    }
    if (cnt == 0) continue;  // wave-uniform
    uint32_t at0 = 0;
    if (ln == 0) at0 = atomicAdd(&used, cnt);
    at0 = first + shfl32(at0, 0);
#pragma unroll
    for (int u = 0; u < kFilterIn; ++u) {
      if ((mask[u] >> ln) & 1ull) {
        const uint64_t i = t0 + static_cast<uint64_t>(u * kWave + ln);
        const uint32_t at = at0 + static_cast<uint32_t>(__popcll(mask[u] & lt));
        out_key[at] = static_cast<uint32_t>(b.fkey[i]);
        out_val[at] = b.price[i] * (100 - static_cast<int64_t>(b.disc[i]));
      }
      at0 += static_cast<uint32_t>(__popcll(mask[u]));
    }
  }
}

// The lookup-first strategy, first pass (ij_lookup_kernel of ij_np_lf_plain,
// src/tpch/tpch_14_ij.cu:371-407, with the correct date predicate): chunk c's
// qualifying rows are looked up in place, one per lane, and every match is
// written as a join entry (fact row, dimension row) from position offsets[c] on;
// matched[c] receives the chunk's match count.  (The reference reserves entries with
// one global atomic counter; here the filter count has sized every chunk's region.)
template <class Dev>
__global__ __launch_bounds__(kAggBlock) void agg_lookup_entries_kernel(Dev d, FactBatchDev b, int32_t lo, int32_t hi,
                                                                       const uint32_t* __restrict__ offsets,
                                                                       uint32_t* __restrict__ matched,
                                                                       uint32_t* __restrict__ ent_row,
                                                                       uint32_t* __restrict__ ent_dim) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ uint32_t used;  // entries of the chunk written so far
  d.stage(lds);
  if (threadIdx.x == 0) used = 0;
  __syncthreads();
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * kFilterChunk;
  const uint64_t c1 = c0 + kFilterChunk < b.rows ? c0 + kFilterChunk : b.rows;
  const int ln = lane();
  const uint64_t lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  const uint32_t first = offsets[blockIdx.x];
  // each wave walks its own 64-row slices of the chunk; each row's date is loaded one
  // slice ahead, so it arrives during the slice before's lookups; a wave reserves its
  // entries with one LDS atomic (the join pass sums them in any order), so the waves of
  // a workgroup never wait for one another
  const uint64_t step = blockDim.x;
  uint64_t i = c0 + threadIdx.x;
  int32_t next = i < c1 ? load_stream(b.date + i) : 0;
  for (uint64_t t0 = c0 + (threadIdx.x / kWave) * kWave; t0 < c1; t0 += step, i += step) {
    const int32_t dt = next;
    next = i + step < c1 ? load_stream(b.date + i + step) : 0;
    const bool qual = i < c1 && dt >= lo && dt < hi;
    bool match = false;
    uint32_t r = kNotFound;
    if (ballot(qual) != 0) {  // wave-uniform
      const uint32_t k = qual ? static_cast<uint32_t>(b.fkey[i]) : 0u;
      r = d.lookup(k, qual, lds);
      match = qual && r != kNotFound;
    }
    const uint64_t mask = ballot(match);
    if (mask == 0) continue;  // wave-uniform
    uint32_t at0 = 0;
    if (ln == 0) at0 = atomicAdd(&used, static_cast<uint32_t>(__popcll(mask)));
    at0 = first + shfl32(at0, 0);
    if (match) {
      const uint32_t at = at0 + static_cast<uint32_t>(__popcll(mask & lt));
      ent_row[at] = static_cast<uint32_t>(i);
      ent_dim[at] = r;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) matched[blockIdx.x] = used;
}

// The lookup-first strategy, second pass (ij_join_kernel, src/tpch/tpch_14_ij.cu:416-461):
// chunk c's join entries gather their price and discount by fact row and their
// flag by dimension row; the workgroup's sums go to partial[c].
__global__ __launch_bounds__(kAggBlock) void agg_join_entries_kernel(FactBatchDev b,
                                                                     const uint32_t* __restrict__ offsets,
                                                                     const uint32_t* __restrict__ matched,
                                                                     const uint32_t* __restrict__ ent_row,
                                                                     const uint32_t* __restrict__ ent_dim,
                                                                     const uint8_t* __restrict__ flag,
                                                                     AggPartial* __restrict__ partial) {
  __shared__ int64_t red[kAggBlock / kWave][3];
  const uint64_t e0 = offsets[blockIdx.x], e1 = e0 + matched[blockIdx.x];
  AggAcc acc;
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const uint32_t row = load_stream(ent_row + e);
    const uint32_t dim = load_stream(ent_dim + e);
    const int64_t sv = b.price[row] * (100 - static_cast<int64_t>(b.disc[row]));
    acc.den += sv;
    if (flag[dim]) acc.num += sv;
    acc.matches += 1;
  }
  const int w = threadIdx.x / kWave;
  const int64_t num = wave_sum_i64(acc.num);
  const int64_t den = wave_sum_i64(acc.den);
  const int64_t mt = wave_sum_i64(static_cast<int64_t>(acc.matches));
  if (lane() == 0) {
    red[w][0] = num;
    red[w][1] = den;
    red[w][2] = mt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    AggPartial p{0, 0, 0, 0};
    for (int v = 0; v < kAggBlock / kWave; ++v) {
      p.num += red[v][0];
      p.den += red[v][1];
      p.matches += static_cast<uint64_t>(red[v][2]);
    }
    partial[blockIdx.x] = p;
  }
}

// Looks up partitioned (fkey, position) tuples and sums the matches' summands; the
// sweep over the tuples is lookup_tuples_kernel's (each XCD one contiguous eighth,
// its workgroups in step).  Padding slots carry position 0xFFFFFFFF.
template <class Dev>
__global__ __launch_bounds__(kAggBlock) void agg_tuples_kernel(Dev d, const uint32_t* __restrict__ keys,
                                                               const uint32_t* __restrict__ pos,
                                                               const uint64_t* __restrict__ d_total,
                                                               uint64_t total_max, const int64_t* __restrict__ vals,
                                                               const uint8_t* __restrict__ flag,
                                                               AggPartial* __restrict__ partial) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ int64_t red[kAggBlock / kWave][3];
  d.stage(lds);
  __syncthreads();
  uint64_t total = total_max;
  if (d_total) total = *d_total < total_max ? *d_total : total_max;
  const uint32_t groups = (gridDim.x % 8 == 0) ? 8u : 1u;  // XCDs
  const uint32_t per = gridDim.x / groups;
  const uint64_t chunk = ((total + groups - 1) / groups + kWave - 1) / kWave * kWave;
  const uint64_t lo = static_cast<uint64_t>(blockIdx.x % groups) * chunk;
  const uint64_t hi = (lo + chunk < total) ? lo + chunk : total;
  const uint64_t w = threadIdx.x / kWave;
  const uint64_t step = static_cast<uint64_t>(per) * blockDim.x;
  AggAcc acc;
  for (uint64_t base = lo + static_cast<uint64_t>(blockIdx.x / groups) * blockDim.x + w * kWave; base < hi;
       base += step) {
    const uint64_t i = base + static_cast<uint64_t>(lane());
    uint32_t p = 0xFFFFFFFFu, k = 0;
    if (i < hi) {
      p = load_stream(pos + i);
      k = load_stream(keys + i);
    }
    const bool active = p != 0xFFFFFFFFu;
    const uint32_t r = d.lookup(k, active, lds);
    if (active && r != kNotFound) {
      const int64_t sv = vals[p];
      acc.den += sv;
      if (flag[r]) acc.num += sv;
      acc.matches += 1;
    }
  }
  const int64_t num = wave_sum_i64(acc.num);
  const int64_t den = wave_sum_i64(acc.den);
  const int64_t mt = wave_sum_i64(static_cast<int64_t>(acc.matches));
  if (lane() == 0) {
    red[w][0] = num;
    red[w][1] = den;
    red[w][2] = mt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    AggPartial p{0, 0, 0, 0};
    for (int v = 0; v < kAggBlock / kWave; ++v) {
      p.num += red[v][0];
      p.den += red[v][1];
      p.matches += static_cast<uint64_t>(red[v][2]);
    }
    partial[blockIdx.x] = p;
  }
}

namespace {

// Workgroup partial sums: at most 8 workgroups of 256 threads fit one CU.
inline uint64_t agg_partial_cap(int cu) { return 8ull * static_cast<uint64_t>(cu); }

// partial holds partial_cap slots (0: agg_partial_cap(cu), what the callers' buffers are sized for).
// Strategy: GJ_AGG_LANE_REFILL, GJ_AGG_PLAIN or GJ_AGG_SORTED (one kernel each, plus the sum kernel).
template <class Dev, int Strategy = GJ_AGG_LANE_REFILL>
void launch_join_agg(const Dev& d, const FactBatchDev& b, int32_t lo, int32_t hi, const uint8_t* flag, int cu,
                     AggPartial* partial, hipStream_t s, uint64_t partial_cap = 0) {
  static_assert(Strategy == GJ_AGG_LANE_REFILL || Strategy == GJ_AGG_PLAIN || Strategy == GJ_AGG_SORTED,
                "one-kernel join-aggregate strategies only");
  if (b.rows == 0) return;
  constexpr bool kSorted = Strategy == GJ_AGG_SORTED;
  constexpr bool kPlain = Strategy == GJ_AGG_PLAIN;
  const size_t ilds = (d.lds_bytes() + 15) / 16 * 16;
  const size_t lds = ilds + (kSorted ? static_cast<size_t>(kAggSortTile) * 16
                                     : (kPlain ? 0 : agg_queue_bytes(kAggUnrollBatch)));
  const void* fn = kSorted ? reinterpret_cast<const void*>(&join_agg_sorted_kernel<Dev>)
                           : reinterpret_cast<const void*>(&join_agg_kernel<Dev, kPlain>);
  const uint64_t rows_per_step = static_cast<uint64_t>(kAggBlock) * (kSorted ? kAggSortUnroll : kAggUnrollBatch);
  uint64_t blocks = static_cast<uint64_t>(blocks_per_cu(fn, kAggBlock, lds)) * cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(b.rows, rows_per_step));
  blocks = std::max<uint64_t>(blocks, 1);
  if (partial_cap == 0) {
    GJ_REQUIRE(blocks <= agg_partial_cap(cu), GJ_ERR_INTERNAL, "join-aggregate grid exceeds the partial buffer");
  } else {
    GJ_REQUIRE(blocks <= kAggDirectBlocks || blocks <= partial_cap, GJ_ERR_UNSUPPORTED,
               "join-aggregate grid needs more workgroup partial slots than the static array holds");
  }
  // a small grid adds its workgroup sums to the result itself (a few dozen atomics
  // cost less than the second launch)
  const bool direct = blocks <= kAggDirectBlocks;
  if constexpr (kSorted) {
    hipLaunchKernelGGL((join_agg_sorted_kernel<Dev>), dim3(static_cast<unsigned>(blocks)), dim3(kAggBlock), lds, s, d,
                       b, lo, hi, flag, partial, direct);
  } else {
    hipLaunchKernelGGL((join_agg_kernel<Dev, kPlain>), dim3(static_cast<unsigned>(blocks)), dim3(kAggBlock), lds, s,
                       d, b, lo, hi, flag, partial, direct);
  }
  GJ_HIP_CHECK(hipGetLastError());
  if (direct) return;
  hipLaunchKernelGGL(agg_sum_kernel, dim3(1), dim3(1024), 0, s, partial, static_cast<uint32_t>(blocks), b.result);
  GJ_HIP_CHECK(hipGetLastError());
}

// The one-kernel strategies by their run-time value.
template <class Dev>
void launch_join_agg_strategy(uint32_t strategy, const Dev& d, const FactBatchDev& b, int32_t lo, int32_t hi,
                              const uint8_t* flag, int cu, AggPartial* partial, hipStream_t s,
                              uint64_t partial_cap = 0) {
  switch (strategy) {
    case GJ_AGG_PLAIN: launch_join_agg<Dev, GJ_AGG_PLAIN>(d, b, lo, hi, flag, cu, partial, s, partial_cap); return;
    case GJ_AGG_SORTED: launch_join_agg<Dev, GJ_AGG_SORTED>(d, b, lo, hi, flag, cu, partial, s, partial_cap); return;
    default: launch_join_agg<Dev, GJ_AGG_LANE_REFILL>(d, b, lo, hi, flag, cu, partial, s, partial_cap); return;
  }
}

// Workgroup partial sums of gj_join_aggregate_noalloc, which must not allocate: a
// file-scope device array (each GPU holds its own copy until the process ends)
// instead of index-owned scratch.  Host code never passes `scratch` itself (on
// the host that name is only a stand-in): noalloc_partials() resolves each
// device's address once, with the device current, and every use goes through it.
// Each workgroup writes its own slot before agg_sum_kernel reads the slots of
// that launch, so nothing carried over from an earlier call is ever read.
constexpr uint64_t kNoAllocPartials = 4096;  // 128 KiB: 8 workgroups per CU on up to 512 CUs
static __device__ __align__(256) AggPartial scratch[kNoAllocPartials];

AggPartial* noalloc_partials() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static AggPartial* ptr[kMaxDevices] = {};
  int dev = 0;
  GJ_HIP_CHECK(hipGetDevice(&dev));
  GJ_REQUIRE(dev >= 0 && dev < kMaxDevices, GJ_ERR_UNSUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> g(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    GJ_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = static_cast<AggPartial*>(p);
  }
  return ptr[dev];
}

void check_agg_index(const gj_index& ix) {
  GJ_REQUIRE(ix.key_bytes == 4, GJ_ERR_INVALID, "join-aggregate needs an index over 4-byte keys");
  GJ_REQUIRE(ix.kind != GJ_NO_OP, GJ_ERR_INVALID, "join-aggregate needs a real index");
}

uint32_t key_bits(uint64_t max_key) { return max_key ? 64u - static_cast<uint32_t>(__builtin_clzll(max_key)) : 1u; }

// The partitioned strategy.  The count of qualifying rows sizes the buffers, so the
// stream is synchronised once between the filter and the rest (the reference splits
// its run the same way, phase1 / phase2, src/tpch/tpch_14_common.cu:720-731).
template <class Dev>
void join_aggregate_partitioned(const gj_index& ix, const Dev& d, const FactBatchDev& b, const gj_agg_params& p,
                                hipStream_t s) {
  if (b.rows == 0) return;
  const uint32_t bits = p.radix_bits ? p.radix_bits : 11u;
  GJ_REQUIRE(bits >= 1 && bits <= 16, GJ_ERR_INVALID, "join-aggregate radix_bits must be in [1, 16]");
  const uint64_t chunks = ceil_div(b.rows, kFilterChunk);
  GJ_REQUIRE(chunks < (1ull << 31), GJ_ERR_INVALID, "fact batch too large");
  auto up = [](uint64_t v) { return (v + 255) / 256 * 256; };
  // the chunk offsets (and, before them, the qualifying-row total) stay at the front of
  // the stream's scratch; the rest is sized by the total
  dbuf& sc = ix.scratch_for(s).a;
  sc.reserve(256 + up(chunks * 4));
  auto* total = sc.as<uint64_t>();
  auto* offs = reinterpret_cast<uint32_t*>(sc.as<uint8_t>() + 256);
  hipLaunchKernelGGL(agg_filter_count_kernel, dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), 0, s, b.date,
                     b.rows, p.date_lo, p.date_hi, offs);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(agg_scan_counts_kernel, dim3(1), dim3(1024), 0, s, offs, chunks, total);
  GJ_HIP_CHECK(hipGetLastError());
  uint64_t m = 0;
  GJ_HIP_CHECK(hipMemcpyAsync(&m, total, sizeof(m), hipMemcpyDeviceToHost, s));
  GJ_HIP_CHECK(hipStreamSynchronize(s));
  if (m == 0) return;
  GJ_REQUIRE(m < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "more than 2^32 - 2 qualifying rows");
  const uint32_t ignore = key_bits(ix.max_key) > bits ? key_bits(ix.max_key) - bits : 0u;
  const uint64_t fan = 1ull << bits;
  const uint64_t cap = partition_capacity(m, 4, bits);
  const uint64_t nparts = agg_partial_cap(ix.cu);
  // total | offsets | compact keys | compact summands | tuple keys | tuple positions | start
  // | count | partitioner scratch | workgroup partials (reserve keeps the front's contents
  // only when it does not reallocate, so a larger buffer is set up before the count)
  const uint64_t o_off = 256, o_key = o_off + up(chunks * 4), o_val = o_key + up(m * 4), o_tk = o_val + up(m * 8),
                 o_tp = o_tk + up(cap * 4), o_st = o_tp + up(cap * 4), o_ct = o_st + up((fan + 1) * 8),
                 o_ps = o_ct + up(fan * 8), o_pa = o_ps + up(partition_scratch(m, 4, bits, 1)),
                 bytes = o_pa + up(nparts * sizeof(AggPartial));
  if (bytes > sc.bytes) {
    // grow and redo the count into the new buffer (dbuf::reserve does not keep contents)
    sc.reserve(bytes);
    total = sc.as<uint64_t>();
    offs = reinterpret_cast<uint32_t*>(sc.as<uint8_t>() + 256);
    hipLaunchKernelGGL(agg_filter_count_kernel, dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), 0, s, b.date,
                       b.rows, p.date_lo, p.date_hi, offs);
    GJ_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(agg_scan_counts_kernel, dim3(1), dim3(1024), 0, s, offs, chunks, total);
    GJ_HIP_CHECK(hipGetLastError());
  }
  uint8_t* base = sc.as<uint8_t>();
  auto* ckey = reinterpret_cast<uint32_t*>(base + o_key);
  auto* cval = reinterpret_cast<int64_t*>(base + o_val);
  auto* tkey = reinterpret_cast<uint32_t*>(base + o_tk);
  auto* tpos = reinterpret_cast<uint32_t*>(base + o_tp);
  auto* start = reinterpret_cast<uint64_t*>(base + o_st);
  auto* count = reinterpret_cast<uint64_t*>(base + o_ct);
  auto* partial = reinterpret_cast<AggPartial*>(base + o_pa);
  hipLaunchKernelGGL(agg_filter_write_kernel, dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), 0, s, b,
                     p.date_lo, p.date_hi, offs, ckey, cval);
  GJ_HIP_CHECK(hipGetLastError());
  partition_keys<uint32_t>(ckey, m, 0, bits, ignore, 1, false, tkey, tpos, start, count, base + o_ps,
                           ix.dropped.as<unsigned long long>(), ix.cu, s);
  const size_t lds = (d.lds_bytes() + 15) / 16 * 16;
  const void* fn = reinterpret_cast<const void*>(&agg_tuples_kernel<Dev>);
  uint64_t blocks = static_cast<uint64_t>(blocks_per_cu(fn, kAggBlock, lds)) * ix.cu;
  blocks = std::min<uint64_t>(blocks, ceil_div(cap, kAggBlock));
  blocks = std::min<uint64_t>(blocks, nparts);
  if (blocks >= 8) blocks = blocks / 8 * 8;  // whole XCD rounds for the XCD-major sweep
  blocks = std::max<uint64_t>(blocks, 1);
  hipLaunchKernelGGL((agg_tuples_kernel<Dev>), dim3(static_cast<unsigned>(blocks)), dim3(kAggBlock), lds, s, d, tkey,
                     tpos, start + fan, cap, cval, p.d_flag, partial);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(agg_sum_kernel, dim3(1), dim3(1024), 0, s, partial, static_cast<uint32_t>(blocks), b.result);
  GJ_HIP_CHECK(hipGetLastError());
}

// The lookup-first strategy: filter count, scan, then the two passes above and the
// sum of the chunks' partials.  Like the partitioned strategy it synchronises the
// stream once, to size the join entries by the qualifying rows.
template <class Dev>
void join_aggregate_lookup_first(const gj_index& ix, const Dev& d, const FactBatchDev& b, const gj_agg_params& p,
                                 hipStream_t s) {
  if (b.rows == 0) return;
  GJ_REQUIRE(b.rows < (1ull << 32), GJ_ERR_INVALID, "the lookup-first strategy takes at most 2^32 - 1 fact rows");
  const uint64_t chunks = ceil_div(b.rows, kFilterChunk);
  auto up = [](uint64_t v) { return (v + 255) / 256 * 256; };
  dbuf& sc = ix.scratch_for(s).a;
  sc.reserve(256 + up(chunks * 4));
  auto* total = sc.as<uint64_t>();
  auto* offs = reinterpret_cast<uint32_t*>(sc.as<uint8_t>() + 256);
  hipLaunchKernelGGL(agg_filter_count_kernel, dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), 0, s, b.date,
                     b.rows, p.date_lo, p.date_hi, offs);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(agg_scan_counts_kernel, dim3(1), dim3(1024), 0, s, offs, chunks, total);
  GJ_HIP_CHECK(hipGetLastError());
  uint64_t m = 0;
  GJ_HIP_CHECK(hipMemcpyAsync(&m, total, sizeof(m), hipMemcpyDeviceToHost, s));
  GJ_HIP_CHECK(hipStreamSynchronize(s));
  if (m == 0) return;
  // total | offsets | match counts | entry rows | entry dimension rows | chunk partials
  const uint64_t o_off = 256, o_mt = o_off + up(chunks * 4), o_er = o_mt + up(chunks * 4), o_ed = o_er + up(m * 4),
                 o_pa = o_ed + up(m * 4), bytes = o_pa + up(chunks * sizeof(AggPartial));
  if (bytes > sc.bytes) {
    // grow and redo the count into the new buffer (dbuf::reserve does not keep contents)
    sc.reserve(bytes);
    total = sc.as<uint64_t>();
    offs = reinterpret_cast<uint32_t*>(sc.as<uint8_t>() + 256);
    hipLaunchKernelGGL(agg_filter_count_kernel, dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), 0, s, b.date,
                       b.rows, p.date_lo, p.date_hi, offs);
    GJ_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(agg_scan_counts_kernel, dim3(1), dim3(1024), 0, s, offs, chunks, total);
    GJ_HIP_CHECK(hipGetLastError());
  }
  uint8_t* base = sc.as<uint8_t>();
  auto* matched = reinterpret_cast<uint32_t*>(base + o_mt);
  auto* erow = reinterpret_cast<uint32_t*>(base + o_er);
  auto* edim = reinterpret_cast<uint32_t*>(base + o_ed);
  auto* partial = reinterpret_cast<AggPartial*>(base + o_pa);
  const size_t lds = (d.lds_bytes() + 15) / 16 * 16;
  hipLaunchKernelGGL((agg_lookup_entries_kernel<Dev>), dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), lds, s,
                     d, b, p.date_lo, p.date_hi, offs, matched, erow, edim);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(agg_join_entries_kernel, dim3(static_cast<unsigned>(chunks)), dim3(kAggBlock), 0, s, b, offs,
                     matched, erow, edim, p.d_flag, partial);
  GJ_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(agg_sum_kernel, dim3(1), dim3(1024), 0, s, partial, static_cast<uint32_t>(chunks), b.result);
  GJ_HIP_CHECK(hipGetLastError());
}

FactBatchDev to_dev(const gj_fact_batch& b, gj_agg_result* res) {
  GJ_REQUIRE(b.rows == 0 || (b.d_fkey && b.d_date && b.d_price && b.d_discount), GJ_ERR_INVALID, "null fact column");
  return FactBatchDev{b.d_fkey, b.d_date, b.d_price, b.d_discount, b.rows, 0, res};
}

}  // namespace

}  // namespace gj

using namespace gj;

namespace gj {

// Side streams and events over which a plan spreads its per-batch launches
// (graph capture, or the launch mode at run time): lanes[0] is the caller's (or
// capturing) stream, lanes[1..] wait on the fork event, which is recorded after
// everything the batches depend on (the results' fill), and lanes[0] waits on
// each side stream's join event after the last launch.
struct GraphLanes {
  GraphLanes() = default;
  GraphLanes(const GraphLanes&) = delete;  // owns the side streams and events
  GraphLanes& operator=(const GraphLanes&) = delete;
  std::vector<hipStream_t> lanes;
  std::vector<hipStream_t> own;
  std::vector<hipEvent_t> events;  // [0] fork, [1..] one join per side stream
  hipError_t create(hipStream_t cs) {
    lanes.push_back(cs);
    for (int i = 0; i < kGraphLanes; ++i) {
      hipEvent_t ev = nullptr;
      hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) return e;
      events.push_back(ev);
      if (i == 0) continue;
      hipStream_t st = nullptr;
      e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      if (e != hipSuccess) return e;
      own.push_back(st);
      lanes.push_back(st);
    }
    return hipSuccess;
  }
  hipError_t fork(hipStream_t cs) {
    hipError_t e = hipEventRecord(events[0], cs);
    for (size_t i = 1; i < lanes.size() && e == hipSuccess; ++i) e = hipStreamWaitEvent(lanes[i], events[0], 0);
    return e;
  }
  hipError_t join(hipStream_t cs) {
    hipError_t e = hipSuccess;
    for (size_t i = 1; i < lanes.size() && e == hipSuccess; ++i) {
      e = hipEventRecord(events[i], lanes[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(cs, events[i], 0);
    }
    return e;
  }
  void destroy() {
    for (hipStream_t st : own) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
    own.clear();
    events.clear();
    lanes.clear();
  }
  ~GraphLanes() { destroy(); }
};

}  // namespace gj

struct gj_batch_plan {
  const gj_index* ix = nullptr;
  bool aggregate = false;
  gj_batch_mode mode = GJ_BATCH_LAUNCH;
  // host copies
  std::vector<FactBatchDev> fact;
  std::vector<ProbeBatchDev> probe;
  std::vector<uint64_t> tile_start;
  uint64_t total_tiles = 0;
  int32_t lo = 0, hi = 0;
  const uint8_t* flag = nullptr;
  gj_agg_result* results = nullptr;
  // device copies (persistent)
  dbuf d_batches, d_tile_start, d_counter;
  dbuf d_tile_batch;  // batch of each persistent work item
  dbuf d_partial;  // workgroup partial sums of the per-batch join-aggregate launches
  // side streams of the launch mode
  gj::GraphLanes lanes;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~gj_batch_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace gj {
namespace {


// Enqueues one execution of the plan (no graph replay) on s.  With `gl` (graph
// capture), batch i goes to lane i % lanes instead, so the captured launches may
// overlap: lookup batches share no scratch, and the aggregate launches of lane j
// use their own slice j of the partial buffer.
void enqueue_plan(gj_batch_plan& pl, hipStream_t s, GraphLanes* gl = nullptr) {
  const std::vector<hipStream_t>* lanes = gl ? &gl->lanes : nullptr;
  const gj_index& ix = *pl.ix;
  if (pl.aggregate) {
    fill_bytes(pl.results, 0, pl.fact.size() * sizeof(gj_agg_result), s);
    with_index_dev<uint32_t>(ix, [&](auto d) {
      using Dev = decltype(d);
      if (pl.mode == GJ_BATCH_PERSISTENT) {
        fill_bytes(pl.d_counter.p, 0, 16, s);
        if (pl.total_tiles == 0) return;
        const size_t lds = (d.lds_bytes() + 15) / 16 * 16 + agg_queue_bytes(kAggUnrollTile);
        const void* fn = reinterpret_cast<const void*>(&join_agg_persistent_kernel<Dev>);
        uint64_t blocks = static_cast<uint64_t>(blocks_per_cu(fn, kAggBlock, lds)) * ix.cu;
        blocks = std::min<uint64_t>(blocks, pl.total_tiles);
        blocks = std::max<uint64_t>(blocks, 1);
        hipLaunchKernelGGL((join_agg_persistent_kernel<Dev>), dim3(static_cast<unsigned>(blocks)), dim3(kAggBlock),
                           lds, s, d, pl.d_batches.as<const FactBatchDev>(), pl.d_tile_batch.as<const uint32_t>(),
                           pl.total_tiles, pl.lo, pl.hi, pl.flag,
                           pl.d_counter.as<unsigned long long>(), pl.d_partial.as<AggPartial>());
        GJ_HIP_CHECK(hipGetLastError());
        const uint32_t nbat = static_cast<uint32_t>(pl.fact.size());
        hipLaunchKernelGGL(agg_batch_sum_kernel, dim3(static_cast<unsigned>(ceil_div(nbat, 256 / kWave))), dim3(256), 0,
                           s, pl.d_batches.as<const FactBatchDev>(), pl.d_tile_start.as<const uint64_t>(), nbat,
                           pl.total_tiles, pl.d_partial.as<const AggPartial>());
        GJ_HIP_CHECK(hipGetLastError());
      } else {
        const uint64_t cap = agg_partial_cap(ix.cu);
        if (gl) GJ_HIP_CHECK(gl->fork(s));  // after the results' fill
        for (size_t i = 0; i < pl.fact.size(); ++i) {
          const size_t lane = lanes ? i % lanes->size() : 0;
          launch_join_agg(d, pl.fact[i], pl.lo, pl.hi, pl.flag, ix.cu, pl.d_partial.as<AggPartial>() + lane * cap,
                          lanes ? (*lanes)[lane] : s);
        }
        if (gl) GJ_HIP_CHECK(gl->join(s));
      }
    });
    return;
  }
  dispatch_key(ix.key_bytes, [&]<class K>() {
    with_index_dev<K>(ix, [&](auto d) {
      using Dev = decltype(d);
      if (pl.mode == GJ_BATCH_PERSISTENT) {
        fill_bytes(pl.d_counter.p, 0, 16, s);
        if (pl.total_tiles == 0) return;
        const size_t lds = d.lds_bytes();
        const void* fn = reinterpret_cast<const void*>(&lookup_persistent_kernel<Dev, K>);
        uint64_t blocks = static_cast<uint64_t>(blocks_per_cu(fn, 256, lds)) * ix.cu;
        blocks = std::min<uint64_t>(blocks, pl.total_tiles);
        blocks = std::max<uint64_t>(blocks, 1);
        hipLaunchKernelGGL((lookup_persistent_kernel<Dev, K>), dim3(static_cast<unsigned>(blocks)), dim3(256), lds, s,
                           d, pl.d_batches.as<const ProbeBatchDev>(), pl.d_tile_batch.as<const uint32_t>(),
                           pl.total_tiles,
                           pl.d_counter.as<unsigned long long>());
        GJ_HIP_CHECK(hipGetLastError());
      } else {
        gj_lookup_params lp;
        gj_lookup_params_default(&lp);
        if (gl) GJ_HIP_CHECK(gl->fork(s));
        for (size_t i = 0; i < pl.probe.size(); ++i) {
          const ProbeBatchDev& b = pl.probe[i];
          const hipStream_t ls = lanes ? (*lanes)[i % lanes->size()] : s;
          launch_plain(d, static_cast<const K*>(b.probes), b.m, b.out, lp, ls, ix.cu);
        }
        if (gl) GJ_HIP_CHECK(gl->join(s));
      }
    });
  });
}


void finish_plan(gj_batch_plan& pl, hipStream_t s) {
  if (pl.mode == GJ_BATCH_PERSISTENT) {
    const size_t nb = pl.aggregate ? pl.fact.size() : pl.probe.size();
    const size_t esz = pl.aggregate ? sizeof(FactBatchDev) : sizeof(ProbeBatchDev);
    pl.d_batches.alloc(std::max<size_t>(nb * esz, 16));
    pl.d_tile_start.alloc(std::max<size_t>(nb * 8, 16));
    pl.d_counter.alloc(16);
    std::vector<uint32_t> tile_batch(pl.total_tiles);
    for (size_t b = 0; b < nb; ++b) {
      const uint64_t end = b + 1 < nb ? pl.tile_start[b + 1] : pl.total_tiles;
      for (uint64_t t = pl.tile_start[b]; t < end; ++t) tile_batch[t] = static_cast<uint32_t>(b);
    }
    pl.d_tile_batch.alloc(std::max<size_t>(tile_batch.size() * 4, 16));
    const void* src = pl.aggregate ? static_cast<const void*>(pl.fact.data()) : static_cast<const void*>(pl.probe.data());
    if (nb) {
      GJ_HIP_CHECK(hipMemcpyAsync(pl.d_batches.p, src, nb * esz, hipMemcpyHostToDevice, s));
      GJ_HIP_CHECK(hipMemcpyAsync(pl.d_tile_start.p, pl.tile_start.data(), nb * 8, hipMemcpyHostToDevice, s));
    }
    if (!tile_batch.empty())
      GJ_HIP_CHECK(hipMemcpyAsync(pl.d_tile_batch.p, tile_batch.data(), tile_batch.size() * 4, hipMemcpyHostToDevice, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
  } else if (pl.mode == GJ_BATCH_GRAPH) {
    // capture needs a non-null stream
    hipStream_t cs = s;
    hipStream_t own = nullptr;
    if (!cs) {
      GJ_HIP_CHECK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
      cs = own;
    }
    // The batches fork over kGraphLanes captured streams (joined again before
    // the graph ends), so small batches' launches overlap on replay.
    GraphLanes gl;
    hipError_t e = hipSuccess;
    if ((pl.aggregate ? pl.fact.size() : pl.probe.size()) > 1) e = gl.create(cs);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      try {
        enqueue_plan(pl, cs, gl.lanes.empty() ? nullptr : &gl);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        gl.destroy();
        if (own) (void)hipStreamDestroy(own);
        throw;
      }
      e = hipStreamEndCapture(cs, &pl.graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0);
    gl.destroy();
    if (own) (void)hipStreamDestroy(own);
    GJ_HIP_CHECK(e);
  } else if ((pl.aggregate ? pl.fact.size() : pl.probe.size()) > 1) {
    // launch mode: the side streams live with the plan; lanes[0] is set per run
    GJ_HIP_CHECK(pl.lanes.create(nullptr));
  }
}

}  // namespace
}  // namespace gj

extern "C" {

int gj_join_aggregate(const gj_index* index, const gj_fact_batch* batch, const gj_agg_params* params,
                      gj_agg_result* d_result, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(index && batch && params && d_result, GJ_ERR_INVALID, "null argument");
    check_agg_index(*index);
    GJ_REQUIRE(params->d_flag, GJ_ERR_INVALID, "null flag column");
    const FactBatchDev b = to_dev(*batch, d_result);
    GJ_REQUIRE(params->strategy <= GJ_AGG_LOOKUP_FIRST, GJ_ERR_INVALID, "unknown join-aggregate strategy");
    const hipStream_t s = S(stream);
    with_index_dev<uint32_t>(*index, [&](auto d) {
      if (params->strategy == GJ_AGG_PARTITIONED) {
        join_aggregate_partitioned(*index, d, b, *params, s);
        return;
      }
      if (params->strategy == GJ_AGG_LOOKUP_FIRST) {
        join_aggregate_lookup_first(*index, d, b, *params, s);
        return;
      }
      dbuf& partial = index->scratch_for(s).a;
      partial.reserve(agg_partial_cap(index->cu) * sizeof(AggPartial));
      launch_join_agg_strategy(params->strategy, d, b, params->date_lo, params->date_hi, params->d_flag, index->cu,
                               partial.as<AggPartial>(), s);
    });
  });
}

int gj_join_aggregate_noalloc(const gj_index* index, const gj_fact_batch* batch, const gj_agg_params* params,
                              gj_agg_result* d_result, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(index && batch && params && d_result, GJ_ERR_INVALID, "null argument");
    check_agg_index(*index);
    GJ_REQUIRE(params->d_flag, GJ_ERR_INVALID, "null flag column");
    GJ_REQUIRE(params->strategy <= GJ_AGG_LOOKUP_FIRST, GJ_ERR_INVALID, "unknown join-aggregate strategy");
    GJ_REQUIRE(params->strategy != GJ_AGG_PARTITIONED && params->strategy != GJ_AGG_LOOKUP_FIRST, GJ_ERR_UNSUPPORTED,
               "the partitioned and lookup-first strategies need scratch memory; use gj_join_aggregate");
    const FactBatchDev b = to_dev(*batch, d_result);
    AggPartial* partial = noalloc_partials();
    with_index_dev<uint32_t>(*index, [&](auto d) {
      launch_join_agg_strategy(params->strategy, d, b, params->date_lo, params->date_hi, params->d_flag, index->cu,
                               partial, S(stream), kNoAllocPartials);
    });
  });
}

int64_t gj_promo_revenue(int64_t numerator, int64_t denominator) {
  // process_results, src/tpch/tpch_14_common.cu:124-127.  Its int64 products
  // overflow for large sums; they wrap here in two's complement (defined, through
  // uint64), as the reference's build does in practice.
  const int64_t n = static_cast<int64_t>(static_cast<uint64_t>(numerator) * 1000u);
  denominator /= 1000;
  if (denominator == 0) return 0;
  return static_cast<int64_t>(static_cast<uint64_t>(n) * 100u) / denominator;
}

int gj_lookup_plan_create(const gj_index* index, const gj_probe_batch* h_batches, uint32_t num_batches,
                          gj_batch_mode mode, gj_stream_t stream, gj_batch_plan** out) {
  return guarded([&] {
    GJ_REQUIRE(index && out && (num_batches == 0 || h_batches), GJ_ERR_INVALID, "null argument");
    GJ_REQUIRE(mode >= GJ_BATCH_LAUNCH && mode <= GJ_BATCH_PERSISTENT, GJ_ERR_INVALID, "unknown batch mode");
    *out = nullptr;
    auto pl = std::make_unique<gj_batch_plan>();
    pl->ix = index;
    pl->mode = mode;
    uint64_t tiles = 0;
    for (uint32_t i = 0; i < num_batches; ++i) {
      const gj_probe_batch& b = h_batches[i];
      GJ_REQUIRE(b.m == 0 || (b.d_probe_keys && b.d_rowids), GJ_ERR_INVALID, "null probe batch buffer");
      pl->probe.push_back(ProbeBatchDev{b.d_probe_keys, b.d_rowids, b.m, tiles});
      pl->tile_start.push_back(tiles);
      tiles += ceil_div(b.m, kLookupTileRows);
    }
    pl->total_tiles = tiles;
    finish_plan(*pl, S(stream));
    *out = pl.release();
  });
}

int gj_aggregate_plan_create(const gj_index* index, const gj_fact_batch* h_batches, uint32_t num_batches,
                             const gj_agg_params* params, gj_agg_result* d_results, gj_batch_mode mode,
                             gj_stream_t stream, gj_batch_plan** out) {
  return guarded([&] {
    GJ_REQUIRE(index && out && params && d_results && (num_batches == 0 || h_batches), GJ_ERR_INVALID,
               "null argument");
    GJ_REQUIRE(mode >= GJ_BATCH_LAUNCH && mode <= GJ_BATCH_PERSISTENT, GJ_ERR_INVALID, "unknown batch mode");
    check_agg_index(*index);
    GJ_REQUIRE(params->d_flag, GJ_ERR_INVALID, "null flag column");
    GJ_REQUIRE(params->strategy == GJ_AGG_LANE_REFILL, GJ_ERR_UNSUPPORTED,
               "aggregate plans run the lane-refill strategy only");
    *out = nullptr;
    auto pl = std::make_unique<gj_batch_plan>();
    pl->ix = index;
    pl->aggregate = true;
    pl->mode = mode;
    pl->lo = params->date_lo;
    pl->hi = params->date_hi;
    pl->flag = params->d_flag;
    pl->results = d_results;
    uint64_t tiles = 0;
    for (uint32_t i = 0; i < num_batches; ++i) {
      FactBatchDev b = to_dev(h_batches[i], d_results + i);
      b.tile_start = tiles;
      pl->fact.push_back(b);
      pl->tile_start.push_back(tiles);
      tiles += ceil_div(b.rows, kAggTileRows);
    }
    pl->total_tiles = tiles;
    // per-workgroup sums of the per-batch launches, or per-tile sums of the persistent kernel
    // (one slice per graph lane for the per-batch launches)
    pl->d_partial.alloc(std::max<uint64_t>(agg_partial_cap(index->cu) * kGraphLanes, tiles * kAggWaves) *
                        sizeof(AggPartial));
    finish_plan(*pl, S(stream));
    *out = pl.release();
  });
}

int gj_batch_plan_run(gj_batch_plan* plan, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(plan, GJ_ERR_INVALID, "null plan");
    if (plan->mode == GJ_BATCH_GRAPH) {
      GJ_HIP_CHECK(hipGraphLaunch(plan->exec, S(stream)));
    } else if (plan->mode == GJ_BATCH_LAUNCH && !plan->lanes.lanes.empty()) {
      plan->lanes.lanes[0] = S(stream);
      enqueue_plan(*plan, S(stream), &plan->lanes);
    } else {
      enqueue_plan(*plan, S(stream));
    }
  });
}

void gj_batch_plan_destroy(gj_batch_plan* plan) { delete plan; }

}  // extern "C"
