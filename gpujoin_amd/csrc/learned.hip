// learned.hip - GPU training and calibration of the learned two-stage position
// models (no reference counterpart; they close over the same exact last-mile
// search as the reference's RadixSpline, src/indexes.cuh:205-215).
//
// Pipeline, all on the GPU over the (key, position) pairs of the sorted column:
//   1. root: linear least squares of cdf = i / n on x = unit(key)  (reduction kernel),
//            an error-bounded greedy spline of the CDF (learned_plr on skewed keys), or
//            a 1xH MLP f(x) = b2 + sum_k w2_k clamp(w1_k x + b1_k, 0, 1) whose ramps start
//            at the column's quantiles (the piecewise-linear CDF through h + 1 quantiles) and
//            are then trained with minibatch Adam by hand-written forward/backward kernels;
//            w1 and w2 are projected to >= 0 after every step, so the MLP is monotone and
//            every leaf owns a contiguous key range
//   2. routing: leaf = floor(root(x) * L); leaf start positions by a boundary kernel
//   3. leaves: per-leaf linear least squares of local position on FP32 key offsets,
//      packed into the 16-byte lookup model (intercept rounded into the start)
//   4. calibration: the error of every build key under the exact FP32 lookup code
//      gives each leaf its error range; leaves wider than the bound are routed to the
//      B-tree, and the extremes over the other leaves form the lookup window
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "lookup_dispatch.hpp"
#include "rs_builder.hpp"

namespace gj {

namespace {

unsigned grid_for(uint64_t work, unsigned block, unsigned cap = 65536) {
  uint64_t g = ceil_div(work, block);
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct RootFitSums {
  double sx, sxx, st, sxt;
};

// Per-block partial sums of x, x^2, t, x*t with x = unit(key_i), t = i / n.
template <class K>
__global__ void root_sums_kernel(const K* __restrict__ col, uint64_t n, K min_key, float inv_range,
                                 RootFitSums* __restrict__ partial) {
  double sx = 0, sxx = 0, st = 0, sxt = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double x = key_unit(col[i], min_key, inv_range);
    const double t = static_cast<double>(i) / static_cast<double>(n);
    sx += x;
    sxx += x * x;
    st += t;
    sxt += x * t;
  }
  sx = wave_sum(sx);
  sxx = wave_sum(sxx);
  st = wave_sum(st);
  sxt = wave_sum(sxt);
  __shared__ RootFitSums s[4];
  const int w = threadIdx.x / 64;
  if (lane() == 0) s[w] = {sx, sxx, st, sxt};
  __syncthreads();
  if (threadIdx.x == 0) {
    RootFitSums r{0, 0, 0, 0};
    for (int i = 0; i < static_cast<int>(blockDim.x / 64); ++i) {
      r.sx += s[i].sx;
      r.sxx += s[i].sxx;
      r.st += s[i].st;
      r.sxt += s[i].sxt;
    }
    partial[blockIdx.x] = r;
  }
}

// ---- MLP root: loss and gradient -------------------------------------------
// Samples come either from explicit arrays (x, y) or from the sorted column at
// seeded random positions (x = unit(key), y = position / n).
template <class K>
struct SampleSource {
  const float* x;
  const float* y;
  const K* col;
  uint64_t n;
  K min_key;
  float inv_range;
  uint64_t seed;
  __device__ __forceinline__ void get(uint64_t j, float& xs, float& ys) const {
    if (x) {
      xs = x[j];
      ys = y[j];
    } else {
      const uint64_t r = mix64(seed + (j + 1) * 0x9E3779B97F4A7C15ull);
      const uint64_t i = static_cast<uint64_t>((static_cast<unsigned __int128>(r) * n) >> 64);
      xs = key_unit(col[i], min_key, inv_range);
      ys = static_cast<float>(static_cast<double>(i) / static_cast<double>(n));
    }
  }
};

// grad[3H+1] += d/dp mean((f(x) - y)^2) over `count` samples; loss += mean squared error.
template <int H, class K>
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float* __restrict__ params, SampleSource<K> src,
                                                       uint64_t count, float* __restrict__ grad,
                                                       float* __restrict__ loss) {
  __shared__ float p[3 * H + 1];
  __shared__ float red[3 * H + 2];
  for (int i = threadIdx.x; i < 3 * H + 1; i += blockDim.x) p[i] = params[i];
  for (int i = threadIdx.x; i < 3 * H + 2; i += blockDim.x) red[i] = 0.0f;
  __syncthreads();
  float gw1[H], gb1[H], gw2[H];
#pragma unroll
  for (int k = 0; k < H; ++k) gw1[k] = gb1[k] = gw2[k] = 0.0f;
  float gb2 = 0.0f, l = 0.0f;
  const float scale = 2.0f / static_cast<float>(count);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < count; j += stride) {
    float x, y;
    src.get(j, x, y);
    float a[H];
    float f = p[3 * H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      a[k] = __builtin_fmaf(p[k], x, p[H + k]);
      f = __builtin_fmaf(p[2 * H + k], fminf(fmaxf(a[k], 0.0f), 1.0f), f);
    }
    const float r = f - y;
    l = __builtin_fmaf(r, r, l);
    const float g = scale * r;
    gb2 += g;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const float hk = fminf(fmaxf(a[k], 0.0f), 1.0f);
      gw2[k] = __builtin_fmaf(g, hk, gw2[k]);
      if (a[k] > 0.0f && a[k] < 1.0f) {  // the ramp passes gradient only on its slope
        const float gw = g * p[2 * H + k];
        gb1[k] += gw;
        gw1[k] = __builtin_fmaf(gw, x, gw1[k]);
      }
    }
  }
  // wave sums, then block sums in LDS, then one global atomic per parameter per block
#pragma unroll
  for (int k = 0; k < H; ++k) {
    gw1[k] = wave_sum(gw1[k]);
    gb1[k] = wave_sum(gb1[k]);
    gw2[k] = wave_sum(gw2[k]);
  }
  gb2 = wave_sum(gb2);
  l = wave_sum(l);
  if (lane() == 0) {
#pragma unroll
    for (int k = 0; k < H; ++k) {
      atomicAdd(&red[k], gw1[k]);
      atomicAdd(&red[H + k], gb1[k]);
      atomicAdd(&red[2 * H + k], gw2[k]);
    }
    atomicAdd(&red[3 * H], gb2);
    atomicAdd(&red[3 * H + 1], l);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * H + 1; i += blockDim.x) atomicAdd(&grad[i], red[i]);
  if (threadIdx.x == 0) atomicAdd(loss, red[3 * H + 1] / static_cast<float>(count));
}

// One Adam step; w1 and w2 are projected onto >= 0 (monotone root); clears grad.
__global__ void adam_kernel(float* __restrict__ params, float* __restrict__ grad, float* __restrict__ m1,
                            float* __restrict__ m2, uint32_t h, float lr, int step) {
  const uint32_t i = threadIdx.x;
  if (i >= 3 * h + 1) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = grad[i];
  grad[i] = 0.0f;
  m1[i] = b1 * m1[i] + (1.0f - b1) * g;
  m2[i] = b2 * m2[i] + (1.0f - b2) * g * g;
  const float mh = m1[i] / (1.0f - powf(b1, static_cast<float>(step)));
  const float vh = m2[i] / (1.0f - powf(b2, static_cast<float>(step)));
  float p = params[i] - lr * mh / (sqrtf(vh) + eps);
  const bool nonneg = (i < h) || (i >= 2 * h && i < 3 * h);
  if (nonneg && p < 0.0f) p = 0.0f;
  params[i] = p;
}

template <class K>
void launch_mlp_grad(uint32_t h, const float* params, const SampleSource<K>& src, uint64_t count, float* grad,
                     float* loss, hipStream_t s) {
  const unsigned blocks = grid_for(count, 256, 1024);
  switch (h) {
    case 8: hipLaunchKernelGGL((mlp_grad_kernel<8, K>), dim3(blocks), dim3(256), 0, s, params, src, count, grad, loss); break;
    case 16: hipLaunchKernelGGL((mlp_grad_kernel<16, K>), dim3(blocks), dim3(256), 0, s, params, src, count, grad, loss); break;
    case 32: hipLaunchKernelGGL((mlp_grad_kernel<32, K>), dim3(blocks), dim3(256), 0, s, params, src, count, grad, loss); break;
    case 64: hipLaunchKernelGGL((mlp_grad_kernel<64, K>), dim3(blocks), dim3(256), 0, s, params, src, count, grad, loss); break;
    default: throw error(GJ_ERR_INVALID, "mlp hidden width must be 8, 16, 32 or 64");
  }
  GJ_HIP_CHECK(hipGetLastError());
}

__global__ void mlp_forward_kernel(const float* __restrict__ params, uint32_t h, const float* __restrict__ x, uint64_t n,
                                   float* __restrict__ out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = mlp_eval(params, h, x[i]);
}

// ---- routing, leaf fit, calibration ----------------------------------------
template <class K>
struct Router {
  K min_key;
  RootModel root;
  uint32_t leaves;
  __device__ __forceinline__ uint32_t operator()(K key) const {
    float cdf;
    if (root.segments) {
      cdf = spline_cdf(static_cast<uint64_t>(key), reinterpret_cast<const uint8_t*>(root.knots), root);
    } else {
      const float u = key_unit(key, min_key, root.inv_range);
      cdf = root.hidden ? mlp_eval(root.mlp, root.hidden, u) : __builtin_fmaf(root.a, u, root.b);
    }
    return leaf_of(cdf, leaves);
  }
};

// start[l] = first position routed to a leaf >= l (start[] pre-filled with n); counts
// routing inversions (which must not occur: the root is monotone).
template <class K>
__global__ void leaf_bounds_kernel(const K* __restrict__ col, uint64_t n, Router<K> route, uint32_t* __restrict__ start,
                                   unsigned long long* inversions) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t li = route(col[i]);
    const int64_t lp = (i == 0) ? -1 : static_cast<int64_t>(route(col[i - 1]));
    if (static_cast<int64_t>(li) < lp) {
      atomicAdd(inversions, 1ull);
      continue;
    }
    for (int64_t l = lp + 1; l <= static_cast<int64_t>(li); ++l) start[l] = static_cast<uint32_t>(i);
  }
}

// Build-time facts of one leaf (not read by lookups).
struct LeafFit {
  uint32_t start, count;
  int32_t err_lo, err_hi;  // extremes of position - pred over the leaf's keys
};

// One block per leaf: least squares of local position on FP32 key offsets.
// Leaf l covers positions [lo[l], hi[l]).
template <class K>
__global__ __launch_bounds__(256) void leaf_fit_kernel(const K* __restrict__ col, uint64_t n,
                                                       const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                                                       uint32_t leaves, LeafModel* __restrict__ out,
                                                       LeafFit* __restrict__ fit) {
  const uint32_t l = blockIdx.x;
  if (l >= leaves) return;
  const uint64_t s = lo[l];
  const uint64_t e = hi[l];
  const uint64_t cnt = e > s ? e - s : 0;
  const uint64_t base = cnt ? static_cast<uint64_t>(col[s]) : 0;
  double sd = 0, sdd = 0, sj = 0, sdj = 0;
  for (uint64_t j = threadIdx.x; j < cnt; j += blockDim.x) {
    const double d = leaf_dkey(col[s + j], base);
    const double jj = static_cast<double>(j);
    sd += d;
    sdd += d * d;
    sj += jj;
    sdj += d * jj;
  }
  sd = wave_sum(sd);
  sdd = wave_sum(sdd);
  sj = wave_sum(sj);
  sdj = wave_sum(sdj);
  __shared__ double red[4][4];
  const int w = threadIdx.x / 64;
  if (lane() == 0) {
    red[w][0] = sd;
    red[w][1] = sdd;
    red[w][2] = sj;
    red[w][3] = sdj;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < static_cast<int>(blockDim.x / 64); ++i) {
      sd += red[i][0];
      sdd += red[i][1];
      sj += red[i][2];
      sdj += red[i][3];
    }
    double slope = 0, icpt = 0;
    if (cnt > 1) {
      const double c = static_cast<double>(cnt);
      const double var = sdd - sd * sd / c;
      if (var > 0) slope = (sdj - sd * sj / c) / var;
      icpt = (sj - slope * sd) / c;
    }
    double st = static_cast<double>(s) + std::floor(icpt + 0.5);
    st = st < 0 ? 0 : (st > static_cast<double>(n) ? static_cast<double>(n) : st);
    LeafModel m;
    m.base_key = base;
    m.start = static_cast<uint32_t>(st);
    m.slope = static_cast<float>(slope);
    out[l] = m;
    fit[l] = LeafFit{static_cast<uint32_t>(s), static_cast<uint32_t>(cnt), INT_MAX, INT_MIN};
  }
}

// err_lo/err_hi of the leaf each key is routed to, under the exact lookup arithmetic.
template <class K>
__global__ void leaf_error_kernel(const K* __restrict__ col, uint64_t n, Router<K> route,
                                  const LeafModel* __restrict__ lm, LeafFit* __restrict__ fit) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const K key = col[i];
    const uint32_t l = route(key);
    const LeafModel m = lm[l];
    const int64_t pred = leaf_pred(m, leaf_dkey(key, m.base_key));
    int64_t e = static_cast<int64_t>(i) - pred;
    e = e < INT_MIN / 2 ? INT_MIN / 2 : (e > INT_MAX / 2 ? INT_MAX / 2 : e);
    atomicMin(&fit[l].err_lo, static_cast<int>(e));
    atomicMax(&fit[l].err_hi, static_cast<int>(e));
  }
}

// Empty leaves predict their start (no key is routed to them); leaves wider than
// max_width go to the B-tree (slope = NaN, error range emptied: err_hi < err_lo).
__global__ void leaf_finalize_kernel(LeafModel* __restrict__ lm, LeafFit* __restrict__ fit, uint32_t leaves,
                                     int64_t max_width) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= leaves) return;
  LeafFit f = fit[l];
  if (f.count == 0) {
    f.err_lo = 0;
    f.err_hi = 0;
    lm[l].slope = 0.0f;
  } else if (static_cast<int64_t>(f.err_hi) - f.err_lo + 1 > max_width) {
    f.err_lo = 1;
    f.err_hi = 0;
    lm[l].slope = __builtin_nanf("");
    lm[l].start = kNoSplit;
  }
  fit[l] = f;
}

// Separator line of split j: sub-leaf t covers [sub_lo, sub_hi); its separator is the
// key at sub_hi - 1 (at the leaf's first position for an empty leading sub-leaf), the
// last separator is all ones, so the count of separators below a key of the split
// leaf is the index of the sub-leaf that holds it.
template <class K>
__global__ void split_sep_kernel(const K* __restrict__ col, const uint32_t* __restrict__ sub_lo,
                                 const uint32_t* __restrict__ sub_hi, const uint32_t* __restrict__ leaf_lo,
                                 uint32_t splits, uint64_t* __restrict__ sep) {
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= static_cast<uint64_t>(splits) * kSplitWays) return;
  const uint32_t t = static_cast<uint32_t>(g % kSplitWays);
  const uint32_t j = static_cast<uint32_t>(g / kSplitWays);
  if (t == kSplitWays - 1) {
    sep[g] = ~0ull;
    return;
  }
  const uint32_t h = sub_hi[g];
  const uint32_t at = h > leaf_lo[j] ? h - 1 : leaf_lo[j];
  sep[g] = static_cast<uint64_t>(col[at]);
}

// err_lo/err_hi of the sub-leaves: one block per split leaf, over the leaf's keys,
// with the separator count and arithmetic of the lookup.
template <class K>
__global__ void split_error_kernel(const K* __restrict__ col, const uint32_t* __restrict__ leaf_lo,
                                   const uint32_t* __restrict__ leaf_hi, const uint64_t* __restrict__ sep,
                                   const LeafModel* __restrict__ subs, LeafFit* __restrict__ fit) {
  const uint32_t j = blockIdx.x;
  const uint64_t* sp = sep + static_cast<uint64_t>(j) * kSplitWays;
  for (uint64_t i = leaf_lo[j] + threadIdx.x; i < leaf_hi[j]; i += blockDim.x) {
    const K key = col[i];
    uint32_t t = 0;
    for (uint32_t u = 0; u < kSplitWays; ++u) t += sp[u] < static_cast<uint64_t>(key) ? 1u : 0u;
    t = t < kSplitWays ? t : kSplitWays - 1;
    const uint64_t g = static_cast<uint64_t>(j) * kSplitWays + t;
    const LeafModel m = subs[g];
    int64_t e = static_cast<int64_t>(i) - leaf_pred(m, leaf_dkey(key, m.base_key));
    e = e < INT_MIN / 2 ? INT_MIN / 2 : (e > INT_MAX / 2 ? INT_MAX / 2 : e);
    atomicMin(&fit[g].err_lo, static_cast<int>(e));
    atomicMax(&fit[g].err_hi, static_cast<int>(e));
  }
}

// The split leaves' own models: slope = NaN, start = split index.
__global__ void split_mark_kernel(LeafModel* __restrict__ lm, const uint32_t* __restrict__ leaf_id, uint32_t splits) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= splits) return;
  LeafModel m;
  m.base_key = 0;
  m.start = j;
  m.slope = __builtin_nanf("");
  lm[leaf_id[j]] = m;
}

// Group g of the compact leaves starts at the first position routed to leaf
// g << gshift or a later leaf (start[] of leaf_bounds_kernel).
template <class K>
__global__ void compact_groups_kernel(const K* __restrict__ col, uint64_t n, const uint32_t* __restrict__ start,
                                      uint32_t groups, uint32_t gshift, LeafGroup* __restrict__ out) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= groups) return;
  const uint32_t st = start[static_cast<uint64_t>(g) << gshift];
  LeafGroup r;
  r.start = st;
  r.base_key = st < n ? static_cast<uint64_t>(col[st]) : static_cast<uint64_t>(col[n - 1]);
  r.pad = 0;
  out[g] = r;
}

// 16-byte leaf -> compact leaf over its group: start + slope * (key - base) equals
// group.start + slope * (key - group.base) + icpt for icpt = start - group.start
// + slope * (group.base - base).  Resets the leaf's error range for re-measuring.
__global__ void compact_convert_kernel(const LeafModel* __restrict__ lm, const LeafGroup* __restrict__ groups,
                                       uint32_t leaves, uint32_t gshift, LeafCompact* __restrict__ out,
                                       LeafFit* __restrict__ fit) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= leaves) return;
  const LeafModel m = lm[l];
  const LeafGroup g = groups[l >> gshift];
  double icpt = static_cast<double>(static_cast<int64_t>(m.start) - static_cast<int64_t>(g.start));
  if (m.slope != 0.0f)
    icpt += static_cast<double>(m.slope) * static_cast<double>(static_cast<int64_t>(g.base_key - m.base_key));
  out[l] = LeafCompact{m.slope, static_cast<float>(icpt)};
  fit[l].err_lo = INT_MAX;
  fit[l].err_hi = INT_MIN;
}

// err_lo/err_hi of every compact leaf under the exact lookup arithmetic.
template <class K>
__global__ void compact_error_kernel(const K* __restrict__ col, uint64_t n, Router<K> route,
                                     const LeafCompact* __restrict__ lc, const LeafGroup* __restrict__ groups,
                                     uint32_t gshift, LeafFit* __restrict__ fit) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const K key = col[i];
    const uint32_t l = route(key);
    const LeafGroup g = groups[l >> gshift];
    const int64_t pred = compact_pred(lc[l], g, leaf_dkey(key, g.base_key));
    int64_t e = static_cast<int64_t>(i) - pred;
    e = e < INT_MIN / 2 ? INT_MIN / 2 : (e > INT_MAX / 2 ? INT_MAX / 2 : e);
    atomicMin(&fit[l].err_lo, static_cast<int>(e));
    atomicMax(&fit[l].err_hi, static_cast<int>(e));
  }
}

// ---- lines layout ------------------------------------------------------------

// first[l] = first position whose target line is >= l (first[] pre-filled with n).
// Counts keys whose target falls below their predecessor's or at or past
// `targets` (neither may occur: the target is monotone and bounded).
template <class K>
__global__ void line_first_kernel(const K* __restrict__ col, uint64_t n, LinesDev<K> d, const uint8_t* rp,
                                  uint64_t targets, uint32_t* __restrict__ first, unsigned long long* bad) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t t = d.target(col[i], rp);
    const int64_t tp = (i == 0) ? -1 : static_cast<int64_t>(d.target(col[i - 1], rp));
    if (t >= targets || static_cast<int64_t>(t) < tp) {
      atomicAdd(bad, 1ull);
      continue;
    }
    for (int64_t l = tp + 1; l <= static_cast<int64_t>(t); ++l) first[l] = static_cast<uint32_t>(i);
  }
}

// Line l holds the keys at positions [base[l], base[l + 1]), after its header.
template <class K>
__global__ void line_fill_kernel(const K* __restrict__ col, const uint32_t* __restrict__ base, uint64_t lines,
                                 K* __restrict__ out) {
  using D = LinesDev<K>;
  const uint64_t slots = lines * D::SLOTS;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < slots; g += stride) {
    const uint64_t l = g / D::SLOTS;
    const uint32_t sl = static_cast<uint32_t>(g % D::SLOTS);
    const uint32_t b = base[l];
    const uint32_t cnt = base[l + 1] - b;
    K v;
    if (sl < D::HDR) {
      const uint64_t hdr = (static_cast<uint64_t>(cnt) << 32) | b;
      v = static_cast<K>(sizeof(K) == 8 ? hdr : (sl == 0 ? (hdr & 0xFFFFFFFFull) : (hdr >> 32)));
    } else {
      const uint32_t k = sl - D::HDR;
      v = k < cnt ? col[static_cast<uint64_t>(b) + k] : key_max<K>();
    }
    out[g] = v;
  }
}

// Two-stage leaves: leaf l maps its keys [start[l], start[l + 1]) linearly onto its
// lines [line0[l], line0[l + 1]).
template <class K>
__global__ void line_leaf_kernel(const K* __restrict__ col, const uint32_t* __restrict__ start,
                                 const uint32_t* __restrict__ line0, uint32_t leaves, LineLeaf* __restrict__ out) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= leaves) return;
  const uint32_t a = start[l], b = start[l + 1];
  LineLeaf f;
  f.line0 = line0[l];
  if (b > a) {
    const uint64_t k0 = static_cast<uint64_t>(col[a]);
    const uint64_t span = static_cast<uint64_t>(col[b - 1]) - k0;
    const uint32_t nl = line0[l + 1] - line0[l];
    f.key0 = k0;
    f.scale = static_cast<float>(static_cast<double>(nl) / (static_cast<double>(span) + 1.0));
  } else {
    f.key0 = 0;
    f.scale = 0.0f;
  }
  out[l] = f;
}

// A lines-layout build under one target model; the candidate holds its buffers
// until it is adopted by the index or dropped.
struct LinesBuild {
  dbuf lines, leaves;
  uint64_t num_lines = 0, mult = 0;
  uint32_t num_leaves = 0, max_lines = 0;
  double avg_lines = 0.0;
  RootModel root{};
  bool ok = false;
};

// Places the keys under target model d (d.leaves / d.mult already set; d.lines unused)
// over `targets` target lines plus whatever tail lines the last full lines push keys into.
// src_line (optional) receives, per line, the smallest target line among its keys.
template <class K>
void place_lines(const gj_index& ix, LinesDev<K> d, const uint8_t* rp, uint64_t targets, LinesBuild& out,
                 hipStream_t s, std::vector<uint32_t>* src_line = nullptr) {
  using D = LinesDev<K>;
  const K* col = static_cast<const K*>(ix.d_col);
  const uint64_t n = ix.n;
  dbuf first((targets + 1) * sizeof(uint32_t));
  {
    std::vector<uint32_t> init(targets + 1, static_cast<uint32_t>(n));
    GJ_HIP_CHECK(hipMemcpyAsync(first.p, init.data(), init.size() * 4, hipMemcpyHostToDevice, s));
    dbuf bad(sizeof(unsigned long long));
    fill_bytes(bad.p, 0, bad.bytes, s);
    hipLaunchKernelGGL(line_first_kernel<K>, dim3(grid_for(n, 256)), dim3(256), 0, s, col, n, d, rp, targets,
                       first.as<uint32_t>(), bad.as<unsigned long long>());
    GJ_HIP_CHECK(hipGetLastError());
    unsigned long long hbad = 0;
    GJ_HIP_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(hbad), hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipMemcpyAsync(init.data(), first.p, init.size() * 4, hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    GJ_REQUIRE(hbad == 0, GJ_ERR_INTERNAL, "learned lines: target model is not monotone");
    // Greedy placement in key order: line l takes its carried-over keys and its own,
    // up to CAP; the rest moves on.  base[l] = row id of line l's first key.
    std::vector<uint32_t> base;
    base.reserve(targets + 16);
    uint64_t carry = 0, pos = 0;
    double sum_line = 0.0, sum_target = 0.0;
    for (uint64_t l = 0; l < targets; ++l) {
      const uint64_t c = static_cast<uint64_t>(init[l + 1]) - init[l];
      const uint64_t avail = carry + c;
      const uint64_t placed = avail < D::CAP ? avail : D::CAP;
      carry = avail - placed;
      base.push_back(static_cast<uint32_t>(pos));
      pos += placed;
      sum_line += static_cast<double>(l) * static_cast<double>(placed);
      sum_target += static_cast<double>(l) * static_cast<double>(c);
    }
    for (uint64_t l = targets; carry > 0; ++l) {  // tail lines for what the last full lines push on
      const uint64_t placed = carry < D::CAP ? carry : D::CAP;
      base.push_back(static_cast<uint32_t>(pos));
      pos += placed;
      carry -= placed;
      sum_line += static_cast<double>(l) * static_cast<double>(placed);
    }
    GJ_REQUIRE(pos == n, GJ_ERR_INTERNAL, "learned lines: placement lost keys");
    const uint64_t total = base.size();
    base.push_back(static_cast<uint32_t>(n));
    // largest spill: the first key of a line has the smallest target in it
    uint64_t t = 0, max_spill = 0;
    if (src_line) src_line->assign(total, 0);
    for (uint64_t l = 0; l < total; ++l) {
      if (base[l + 1] == base[l]) {
        if (src_line) (*src_line)[l] = static_cast<uint32_t>(l);
        continue;
      }
      while (t + 1 < targets && init[t + 1] <= base[l]) ++t;
      max_spill = std::max<uint64_t>(max_spill, l - t);
      if (src_line) (*src_line)[l] = static_cast<uint32_t>(t);
    }
    out.num_lines = total;
    out.avg_lines = 1.0 + (sum_line - sum_target) / static_cast<double>(n);
    out.max_lines = static_cast<uint32_t>(std::min<uint64_t>(max_spill + 1, 0xFFFFFFFFull));
    first.release();
    dbuf d_base(base.size() * sizeof(uint32_t));
    GJ_HIP_CHECK(hipMemcpyAsync(d_base.p, base.data(), base.size() * 4, hipMemcpyHostToDevice, s));
    out.lines.alloc(total * 128);
    hipLaunchKernelGGL(line_fill_kernel<K>, dim3(grid_for(total * D::SLOTS, 256)), dim3(256), 0, s, col,
                       d_base.as<const uint32_t>(), total, out.lines.as<K>());
    GJ_HIP_CHECK(hipGetLastError());
    GJ_HIP_CHECK(hipStreamSynchronize(s));
  }
  out.ok = true;
}

// Root-only lines: the key range spread linearly over ceil(n / line_keys) lines.
template <class K>
void lines_root_only(const gj_index& ix, uint32_t line_keys, LinesBuild& out, hipStream_t s) {
  const uint64_t targets = std::max<uint64_t>(1, ceil_div(ix.n, line_keys));
  const unsigned __int128 den = static_cast<unsigned __int128>(ix.max_key - ix.min_key) + 1;
  // mult = floor(targets * 2^64 / (range + 1)), so umulhi(range, mult) < targets; capped at
  // 2^64 - 1 for a range narrower than the line count (then target <= key - min_key)
  const unsigned __int128 q = (static_cast<unsigned __int128>(targets) << 64) / den;
  out.mult = q > static_cast<unsigned __int128>(~0ull) ? ~0ull : static_cast<uint64_t>(q);
  LinesDev<K> d{};
  d.min_key = static_cast<K>(ix.min_key);
  d.max_key = static_cast<K>(ix.max_key);
  d.mult = out.mult;
  out.num_leaves = 0;
  out.root = RootModel{};
  place_lines<K>(ix, d, nullptr, targets, out, s);
}

// Two-stage lines: `root` routes keys to 2^bits leaves; each leaf gets lines for its
// keys at line_keys per line and maps its key span linearly onto them.
template <class K>
void lines_two_stage(const gj_index& ix, const RootModel& root, uint32_t bits, uint32_t line_keys, LinesBuild& out,
                     hipStream_t s) {
  const K* col = static_cast<const K*>(ix.d_col);
  const uint64_t n = ix.n;
  const uint32_t L = 1u << bits;
  const K kmin = static_cast<K>(ix.min_key);
  Router<K> route{kmin, root, L};
  dbuf start((static_cast<uint64_t>(L) + 1) * sizeof(uint32_t));
  std::vector<uint32_t> hs(L + 1, static_cast<uint32_t>(n));
  {
    GJ_HIP_CHECK(hipMemcpyAsync(start.p, hs.data(), hs.size() * 4, hipMemcpyHostToDevice, s));
    dbuf inv(sizeof(unsigned long long));
    fill_bytes(inv.p, 0, inv.bytes, s);
    hipLaunchKernelGGL(leaf_bounds_kernel<K>, dim3(grid_for(n, 256)), dim3(256), 0, s, col, n, route,
                       start.as<uint32_t>(), inv.as<unsigned long long>());
    GJ_HIP_CHECK(hipGetLastError());
    unsigned long long hinv = 0;
    GJ_HIP_CHECK(hipMemcpyAsync(&hinv, inv.p, sizeof(hinv), hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipMemcpyAsync(hs.data(), start.p, hs.size() * 4, hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    GJ_REQUIRE(hinv == 0, GJ_ERR_INTERNAL, "learned lines: root model is not monotone");
  }
  // A leaf whose keys are not evenly spread over its key span overfills some of its lines,
  // and the surplus runs on through the following lines; a lookup wave walks as far as its
  // farthest probe.  So the leaves that source a run of two or more lines get twice their
  // lines (half the load where their keys crowd), and the placement is redone, up to
  // kInflateRounds times.
  constexpr int kInflateRounds = 4;
  constexpr uint32_t kMaxInflate = 16;
  std::vector<uint32_t> inflate(L, 1), line0(L + 1, 0), src;
  dbuf d_line0((static_cast<uint64_t>(L) + 1) * sizeof(uint32_t));
  out.leaves.alloc(sizeof(LineLeaf) * L);
  out.num_leaves = L;
  out.root = root;
  for (int round = 0;; ++round) {
    uint64_t lines = 0;
    for (uint32_t l = 0; l < L; ++l) {
      line0[l] = static_cast<uint32_t>(lines);
      const uint64_t c = static_cast<uint64_t>(hs[l + 1]) - hs[l];
      lines += ceil_div(c * inflate[l], line_keys);
    }
    GJ_REQUIRE(lines < 0xFFFFFFFFull, GJ_ERR_OVERFLOW, "learned lines: too many lines");
    line0[L] = static_cast<uint32_t>(lines);
    GJ_HIP_CHECK(hipMemcpyAsync(d_line0.p, line0.data(), line0.size() * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(line_leaf_kernel<K>, dim3(grid_for(L, 256)), dim3(256), 0, s, col,
                       start.as<const uint32_t>(), d_line0.as<const uint32_t>(), L, out.leaves.as<LineLeaf>());
    GJ_HIP_CHECK(hipGetLastError());
    LinesDev<K> d{};
    d.min_key = kmin;
    d.max_key = static_cast<K>(ix.max_key);
    d.root = root;
    d.leaves = out.leaves.as<const LineLeaf>();
    d.num_leaves = L;
    // FP32 rounding can carry a leaf's last keys to floor(p) = its line count, the first
    // line of the next leaf (still monotone); for the last leaf that is target line `lines`
    place_lines<K>(ix, d, reinterpret_cast<const uint8_t*>(root.params()), lines + 1, out, s, &src);
    if (out.max_lines <= 2 || round + 1 >= kInflateRounds) break;
    std::vector<uint8_t> mark(L, 0);
    for (uint64_t l = 0; l < src.size(); ++l) {
      if (l - src[l] < 2) continue;
      // the leaf whose lines hold the source line
      const uint32_t j = static_cast<uint32_t>(
          std::upper_bound(line0.begin(), line0.end(), src[l]) - line0.begin() - 1);
      if (j < L) mark[j] = 1;
    }
    bool grew = false;
    for (uint32_t j = 0; j < L; ++j) {
      if (mark[j] && inflate[j] < kMaxInflate) {
        inflate[j] *= 2;
        grew = true;
      }
    }
    if (!grew) break;
  }
}

// Adopts a lines build into the index: the column layout's leaves and the B-tree
// fallback are not needed.
void adopt_lines(gj_index& ix, LinesBuild& b) {
  ix.lm_layout = GJ_LAYOUT_LINES;
  ix.lm_lines = std::move(b.lines);
  ix.lm_line_leaves = std::move(b.leaves);
  ix.lm_num_lines = b.num_lines;
  ix.lm_mult = b.mult;
  ix.lm_num_line_leaves = b.num_leaves;
  ix.lm_root = b.root;
  ix.lm_leaves.release();
  ix.lm_groups.release();
  ix.lm_seps.release();
  ix.lm_subs.release();
  ix.lm_compact = false;
  ix.lm_splits = 0;
  ix.lm_fallback_any = 0;
  ix.btree_nodes.release();
  ix.bt = BTreeLayout{};
  ix.stats.layout = GJ_LAYOUT_LINES;
  ix.stats.lines = b.num_lines;
  ix.stats.avg_lines = b.avg_lines;
  ix.stats.max_lines = b.max_lines;
  ix.stats.leaves = b.num_leaves;
  ix.stats.root_segments = b.root.segments;
  ix.stats.fallback_leaves = 0;
  ix.stats.fallback_keys = 0;
  ix.stats.split_leaves = 0;
  ix.stats.max_window = 0;
  ix.stats.avg_window = 0.0;
  ix.stats.depth = b.num_leaves ? 2 : 1;
  ix.stats.lds_levels = 0;
}

}  // namespace

template <class K>
void build_learned(gj_index& ix, hipStream_t s) {
  const K* col = static_cast<const K*>(ix.d_col);
  const uint64_t n = ix.n;
  // default: 2^16 16-byte leaves (1 MiB, L2-resident) up to 2^26 keys, then about 1024
  // keys per leaf; twice as many compact leaves in the same bytes where the build admits them
  uint32_t bits = ix.params.learned_leaf_bits;
  if (bits == 0) bits = static_cast<uint32_t>(std::clamp(static_cast<int>(ilog2_ceil(n)) - 10, 16, 24));
  GJ_REQUIRE(bits >= 1 && bits <= 24, GJ_ERR_INVALID, "learned_leaf_bits must be in [1, 24]");
  uint32_t L = 1u << bits;  // leaves of the build in progress
  const K kmin = static_cast<K>(ix.min_key);
  const uint64_t range = ix.max_key - ix.min_key;
  RootModel root{};
  root.inv_range = range ? 1.0f / static_cast<float>(range) : 0.0f;
  std::vector<float> mlp_init;  // the MLP root's parameters before training
  ix.stats.root_trained = (ix.kind == GJ_LEARNED_MLP && !(ix.params.learned_flags & GJ_LEARNED_UNTRAINED)) ? 1u : 0u;

  if (ix.kind == GJ_LEARNED_PLR) {
    const unsigned blocks = grid_for(n, 256, 2048);
    dbuf partial(sizeof(RootFitSums) * blocks);
    hipLaunchKernelGGL(root_sums_kernel<K>, dim3(blocks), dim3(256), 0, s, col, n, kmin, root.inv_range,
                       partial.as<RootFitSums>());
    GJ_HIP_CHECK(hipGetLastError());
    std::vector<RootFitSums> h(blocks);
    GJ_HIP_CHECK(hipMemcpyAsync(h.data(), partial.p, partial.bytes, hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    RootFitSums t{0, 0, 0, 0};
    for (const auto& r : h) {
      t.sx += r.sx;
      t.sxx += r.sxx;
      t.st += r.st;
      t.sxt += r.sxt;
    }
    const double c = static_cast<double>(n);
    const double var = t.sxx - t.sx * t.sx / c;
    double a = var > 0 ? (t.sxt - t.sx * t.st / c) / var : 0.0;
    if (a < 0) a = 0;
    root.a = static_cast<float>(a);
    root.b = static_cast<float>((t.st - a * t.sx) / c);
    root.hidden = 0;
  } else {
    const uint32_t h = ix.params.mlp_hidden;
    GJ_REQUIRE(h == 8 || h == 16 || h == 32 || h == 64, GJ_ERR_INVALID, "mlp_hidden must be 8, 16, 32 or 64");
    const uint32_t np = 3 * h + 1;
    // ramp k rises from 0 at quantile k/h of the keys to 1 at quantile (k+1)/h, weight 1/h:
    // the initial root is the piecewise-linear CDF through h + 1 quantiles
    std::vector<float> q(h + 1);
    for (uint32_t k = 0; k <= h; ++k) {
      const uint64_t i = std::min<uint64_t>(n - 1, (static_cast<uint64_t>(k) * n) / h);
      K key;
      GJ_HIP_CHECK(hipMemcpyAsync(&key, col + i, sizeof(K), hipMemcpyDeviceToHost, s));
      GJ_HIP_CHECK(hipStreamSynchronize(s));
      q[k] = static_cast<float>(static_cast<uint64_t>(key - kmin)) * root.inv_range;
    }
    std::vector<float> p0(np, 0.0f);
    for (uint32_t k = 0; k < h; ++k) {
      const float w = q[k + 1] - q[k];
      p0[k] = w > 0.0f ? 1.0f / w : 0.0f;
      p0[h + k] = w > 0.0f ? -q[k] / w : (q[k] <= 0.0f ? 1.0f : 0.0f);
      p0[2 * h + k] = 1.0f / static_cast<float>(h);
    }
    p0[3 * h] = 0.0f;
    mlp_init = p0;
    ix.lm_mlp.alloc(((np * 4 + 15) / 16) * 16);
    fill_bytes(ix.lm_mlp.p, 0, ix.lm_mlp.bytes, s);
    GJ_HIP_CHECK(hipMemcpyAsync(ix.lm_mlp.p, p0.data(), np * 4, hipMemcpyHostToDevice, s));
    dbuf state(4 * 64 * 4 * 4);  // grad, m1, m2, loss
    fill_bytes(state.p, 0, state.bytes, s);
    float* grad = state.as<float>();
    float* m1 = grad + 256;
    float* m2 = grad + 512;
    float* loss = grad + 768;
    SampleSource<K> src{nullptr, nullptr, col, n, kmin, root.inv_range, ix.params.seed};
    const uint64_t batch = std::min<uint64_t>(ix.params.mlp_batch, std::max<uint64_t>(n, 1));
    const uint32_t steps = (ix.params.learned_flags & GJ_LEARNED_UNTRAINED) ? 0u : ix.params.mlp_steps;
    for (uint32_t step = 1; step <= steps; ++step) {
      src.seed = ix.params.seed + static_cast<uint64_t>(step) * 0x632BE59BD9B4E019ull;
      if (step == steps) fill_bytes(loss, 0, 16, s);
      launch_mlp_grad<K>(h, ix.lm_mlp.as<float>(), src, batch, grad, loss, s);
      hipLaunchKernelGGL(adam_kernel, dim3(1), dim3(256), 0, s, ix.lm_mlp.as<float>(), grad, m1, m2, h,
                         ix.params.mlp_lr, static_cast<int>(step));
      GJ_HIP_CHECK(hipGetLastError());
    }
    if (steps == 0) {
      // untrained root: its loss on one minibatch of the same sampler, parameters untouched
      src.seed = ix.params.seed + 0x632BE59BD9B4E019ull;
      launch_mlp_grad<K>(h, ix.lm_mlp.as<float>(), src, batch, grad, loss, s);
    }
    float hl = 0;
    GJ_HIP_CHECK(hipMemcpyAsync(&hl, loss, sizeof(float), hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    ix.stats.train_loss = hl;
    root.hidden = h;
    root.mlp = ix.lm_mlp.as<const float>();
  }

  // Spline root: GreedySplineCorridor knots over (key, position) with the error bound
  // doubled until at most `max_segs` segments remain; stored as (key, cdf = pos / n,
  // slope) with the segment count padded to a power of two by repeating the last knot.
  auto spline_root = [&](uint32_t max_segs) {
    std::vector<uint8_t> hcopy;
    const void* hk = ix.h_col;
    if (!hk) {
      hcopy.resize(n * sizeof(K));
      GJ_HIP_CHECK(hipMemcpyAsync(hcopy.data(), col, hcopy.size(), hipMemcpyDeviceToHost, s));
      GJ_HIP_CHECK(hipStreamSynchronize(s));
      hk = hcopy.data();
    }
    uint64_t err = std::max<uint64_t>(16, n / (8ull * max_segs));
    RadixSplineHost sp;
    for (;;) {
      sp = build_radix_spline(hk, n, sizeof(K), 1, static_cast<uint32_t>(std::min<uint64_t>(err, 0x7FFFFFFFull)));
      if (sp.knot_x.size() <= static_cast<size_t>(max_segs) + 1 || err >= n) break;
      err *= 2;
    }
    const uint32_t knots = static_cast<uint32_t>(sp.knot_x.size());
    uint32_t segs = 1;
    while (segs + 1 < knots) segs *= 2;
    RootModel r{};
    r.inv_range = root.inv_range;
    r.segments = segs;
    std::vector<uint8_t> h(r.bytes(), 0);
    auto* key = reinterpret_cast<uint64_t*>(h.data());
    auto* cdf = reinterpret_cast<float*>(h.data() + r.key_bytes());
    auto* slope = reinterpret_cast<float*>(h.data() + r.key_bytes() + r.cdf_bytes());
    const double inv_n = 1.0 / static_cast<double>(n);
    for (uint32_t i = 0; i <= segs; ++i) {
      const uint32_t j = std::min(i, knots - 1);
      key[i] = sp.knot_x[j];
      cdf[i] = static_cast<float>(sp.knot_y[j] * inv_n);
    }
    for (uint32_t i = 0; i < segs; ++i) {
      const uint64_t dx = key[i + 1] - key[i];
      slope[i] = dx ? static_cast<float>((static_cast<double>(cdf[i + 1]) - cdf[i]) / static_cast<double>(dx)) : 0.0f;
    }
    ix.lm_knots.alloc(h.size());
    GJ_HIP_CHECK(hipMemcpyAsync(ix.lm_knots.p, h.data(), h.size(), hipMemcpyHostToDevice, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    r.knots = ix.lm_knots.as<const uint64_t>();
    ix.stats.spline_points = knots;
    return r;
  };
  if (ix.kind == GJ_LEARNED_PLR && ix.params.learned_root_segments >= 2) {
    const uint32_t want = ix.params.learned_root_segments;
    GJ_REQUIRE((want & (want - 1)) == 0 && want <= 4096, GJ_ERR_INVALID,
               "learned_root_segments must be 0, 1 or a power of two <= 4096");
    root = spline_root(want);
  }

  // Lines layout: the model places the keys itself, so a present key is read from the
  // line the model names (or, behind a run of full lines, the next ones).  Candidates,
  // cheapest first: the linear root alone (no table to read), then the root routing to
  // leaves that each spread their keys over their own lines, then (learned_plr) a
  // spline root of the column's CDF routing to leaves.  The one with the fewest lines
  // per present key wins; the column layout takes over in auto mode when even that
  // reads more than 1.10 lines per key or up to more than 4 lines, or when the line
  // copy does not fit in memory.
  // A column below 16 MiB is served from L2 and the MALL, where the column's denser lines
  // (16 or 32 keys against 10 or 20) win; auto keeps it.  So it does for structures placed
  // in host memory: the line copy would move every key lookup onto the host link, where
  // the column layout's model (a few MiB, cached in L2) leaves the key reads in the column.
  const uint32_t layout = ix.params.learned_layout;
  GJ_REQUIRE(layout <= GJ_LAYOUT_LINES, GJ_ERR_INVALID, "learned_layout must be 0, 1 or 2");
  constexpr uint64_t kLinesMinColumnBytes = 16ull << 20;
  if (layout == GJ_LAYOUT_LINES || (layout == GJ_LAYOUT_AUTO && n * sizeof(K) >= kLinesMinColumnBytes &&
                                    ix.params.placement != GJ_PLACE_HOST)) {
    using D = LinesDev<K>;
    const uint32_t lk = ix.params.learned_line_keys ? ix.params.learned_line_keys : (sizeof(K) == 8 ? 10u : 20u);
    GJ_REQUIRE(lk >= 1 && lk <= D::CAP, GJ_ERR_INVALID, "learned_line_keys must be in [1, keys per line]");
    const uint32_t lbits = ix.params.learned_leaf_bits
                               ? ix.params.learned_leaf_bits
                               : static_cast<uint32_t>(std::clamp(static_cast<int>(ilog2_ceil(n)) - 10, 8, 20));
    LinesBuild best;
    auto consider = [&](auto&& make) {
      LinesBuild c;
      try {
        make(c);
      } catch (const error& e) {
        if (e.status != GJ_ERR_NOMEM || layout == GJ_LAYOUT_LINES) throw;
        (void)hipGetLastError();
        return;  // auto: the line copy does not fit; try the next candidate or the column layout
      }
      // a candidate within the auto bounds beats one outside them; then fewer lines per key
      auto good = [](const LinesBuild& b) { return b.avg_lines <= 1.10 && b.max_lines <= 4; };
      if (!best.ok || (good(c) && !good(best)) || (good(c) == good(best) && c.avg_lines < best.avg_lines))
        best = std::move(c);
    };
    constexpr double kGoodLines = 1.05;  // below this, a costlier candidate cannot gain much
    const bool linear_root = root.hidden == 0 && root.segments == 0;
    if (linear_root && ix.params.learned_leaf_bits == 0)
      consider([&](LinesBuild& c) { lines_root_only<K>(ix, lk, c, s); });
    if (!best.ok || best.avg_lines > kGoodLines)
      consider([&](LinesBuild& c) { lines_two_stage<K>(ix, root, lbits, lk, c, s); });
    if (ix.kind == GJ_LEARNED_PLR && linear_root && ix.params.learned_root_segments == 0 &&
        (!best.ok || best.avg_lines > kGoodLines) && n >= 4)
      consider([&](LinesBuild& c) { lines_two_stage<K>(ix, spline_root(2048), lbits, lk, c, s); });
    // Training lowers the MLP root's mean squared error, but that can leave a flat stretch
    // that routes a crowd of keys to a few leaves (lognormal keys at 2^26: 700 lines per
    // key); the quantile initialisation routes equal key counts per ramp.  So the build
    // also places the keys under the initial root, and keeps the trained one unless the
    // initial one does better.
    bool init_root = false;
    if (root.hidden && !mlp_init.empty() && !(ix.params.learned_flags & GJ_LEARNED_UNTRAINED) &&
        (!best.ok || best.avg_lines > kGoodLines || best.max_lines > 4)) {
      dbuf init_params(ix.lm_mlp.bytes);
      fill_bytes(init_params.p, 0, init_params.bytes, s);
      GJ_HIP_CHECK(hipMemcpyAsync(init_params.p, mlp_init.data(), mlp_init.size() * 4, hipMemcpyHostToDevice, s));
      RootModel r0 = root;
      r0.mlp = init_params.as<const float>();
      consider([&](LinesBuild& c) { lines_two_stage<K>(ix, r0, lbits, lk, c, s); });
      init_root = best.ok && best.root.mlp == r0.mlp;
      if (init_root) {
        // the index serves lookups with the initial parameters
        GJ_HIP_CHECK(hipMemcpyAsync(ix.lm_mlp.p, init_params.p, init_params.bytes, hipMemcpyDeviceToDevice, s));
        GJ_HIP_CHECK(hipStreamSynchronize(s));
        best.root.mlp = ix.lm_mlp.as<const float>();
      }
    }
    if (init_root) ix.stats.root_trained = 0;
    // a lookup wave walks as many lines as its farthest probe needs, so auto also bounds
    // the longest walk, not only the mean
    if (best.ok && (layout == GJ_LAYOUT_LINES || (best.avg_lines <= 1.10 && best.max_lines <= 4))) {
      if (best.root.segments == 0) {
        ix.lm_knots.release();
        ix.stats.spline_points = 0;
      }
      adopt_lines(ix, best);
      return;
    }
    GJ_REQUIRE(layout != GJ_LAYOUT_LINES, GJ_ERR_NOMEM, "learned lines: the line copy does not fit");
    ix.lm_knots.release();
    ix.stats.spline_points = 0;
  }
  ix.lm_layout = GJ_LAYOUT_COLUMN;
  ix.stats.layout = GJ_LAYOUT_COLUMN;

  dbuf last_start, last_fit;  // leaf starts and fits of the last build_leaves
  // routing, leaf fits and calibration under a root; returns the build keys routed to the B-tree
  auto build_leaves = [&](const RootModel& root) -> uint64_t {
    ix.lm_compact = false;
    ix.lm_groups.release();
    Router<K> route{kmin, root, L};
    dbuf start((static_cast<uint64_t>(L) + 1 + 3) * sizeof(uint32_t));
    {
      // fill start[] with n
      std::vector<uint32_t> init(L + 1, static_cast<uint32_t>(n));
      GJ_HIP_CHECK(hipMemcpyAsync(start.p, init.data(), init.size() * 4, hipMemcpyHostToDevice, s));
      dbuf inv(sizeof(unsigned long long));
      fill_bytes(inv.p, 0, inv.bytes, s);
      hipLaunchKernelGGL(leaf_bounds_kernel<K>, dim3(grid_for(n, 256)), dim3(256), 0, s, col, n, route,
                         start.as<uint32_t>(), inv.as<unsigned long long>());
      GJ_HIP_CHECK(hipGetLastError());
      unsigned long long hinv = 0;
      GJ_HIP_CHECK(hipMemcpyAsync(&hinv, inv.p, sizeof(hinv), hipMemcpyDeviceToHost, s));
      GJ_HIP_CHECK(hipStreamSynchronize(s));
      GJ_REQUIRE(hinv == 0, GJ_ERR_INTERNAL, "learned index: root model is not monotone");
    }
    ix.lm_leaves.alloc(sizeof(LeafModel) * L);
    dbuf fit(sizeof(LeafFit) * L);
    hipLaunchKernelGGL(leaf_fit_kernel<K>, dim3(L), dim3(256), 0, s, col, n, start.as<const uint32_t>(),
                       start.as<const uint32_t>() + 1, L, ix.lm_leaves.as<LeafModel>(), fit.as<LeafFit>());
    GJ_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(leaf_error_kernel<K>, dim3(grid_for(n, 256)), dim3(256), 0, s, col, n, route,
                       ix.lm_leaves.as<const LeafModel>(), fit.as<LeafFit>());
    GJ_HIP_CHECK(hipGetLastError());
    const int64_t max_width = 2 * static_cast<int64_t>(ix.params.learned_max_error) + 2;
    hipLaunchKernelGGL(leaf_finalize_kernel, dim3(grid_for(L, 256)), dim3(256), 0, s, ix.lm_leaves.as<LeafModel>(),
                       fit.as<LeafFit>(), L, max_width);
    GJ_HIP_CHECK(hipGetLastError());

    std::vector<LeafFit> hf(L);
    GJ_HIP_CHECK(hipMemcpyAsync(hf.data(), fit.p, fit.bytes, hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));

    // split a leaf that is too wide into kSplitWays equal-count sub-leaves, unless its
    // sub-leaves would still hold over 4 average leaves each (then no line fits them
    // either, and the extra separator read would only delay the B-tree)
    const uint64_t split_max = static_cast<uint64_t>(kSplitWays) * 4 * std::max<uint64_t>(1, ceil_div(n, L));
    std::vector<uint32_t> split_id, split_lo, split_hi, sub_lo, sub_hi;
    for (uint32_t l = 0; l < L; ++l) {
      const LeafFit& f = hf[l];
      if (f.count == 0 || f.err_hi >= f.err_lo || f.count > split_max) continue;
      split_id.push_back(l);
      split_lo.push_back(f.start);
      split_hi.push_back(f.start + f.count);
      for (uint32_t t = 0; t < kSplitWays; ++t) {
        sub_lo.push_back(f.start + static_cast<uint32_t>((static_cast<uint64_t>(f.count) * t) / kSplitWays));
        sub_hi.push_back(f.start + static_cast<uint32_t>((static_cast<uint64_t>(f.count) * (t + 1)) / kSplitWays));
      }
    }
    const uint32_t J = static_cast<uint32_t>(split_id.size());
    std::vector<LeafFit> hs;
    ix.lm_splits = J;
    ix.lm_seps.release();
    ix.lm_subs.release();
    if (J) {
      const uint64_t S = static_cast<uint64_t>(J) * kSplitWays;
      dbuf d_id(J * 4), d_lo(J * 4), d_hi(J * 4), d_slo(S * 4), d_shi(S * 4), sfit(S * sizeof(LeafFit));
      GJ_HIP_CHECK(hipMemcpyAsync(d_id.p, split_id.data(), J * 4, hipMemcpyHostToDevice, s));
      GJ_HIP_CHECK(hipMemcpyAsync(d_lo.p, split_lo.data(), J * 4, hipMemcpyHostToDevice, s));
      GJ_HIP_CHECK(hipMemcpyAsync(d_hi.p, split_hi.data(), J * 4, hipMemcpyHostToDevice, s));
      GJ_HIP_CHECK(hipMemcpyAsync(d_slo.p, sub_lo.data(), S * 4, hipMemcpyHostToDevice, s));
      GJ_HIP_CHECK(hipMemcpyAsync(d_shi.p, sub_hi.data(), S * 4, hipMemcpyHostToDevice, s));
      ix.lm_seps.alloc(S * 8);
      ix.lm_subs.alloc(S * sizeof(LeafModel));
      hipLaunchKernelGGL(split_sep_kernel<K>, dim3(grid_for(S, 256)), dim3(256), 0, s, col, d_slo.as<const uint32_t>(),
                         d_shi.as<const uint32_t>(), d_lo.as<const uint32_t>(), J, ix.lm_seps.as<uint64_t>());
      GJ_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(leaf_fit_kernel<K>, dim3(static_cast<unsigned>(S)), dim3(256), 0, s, col, n,
                         d_slo.as<const uint32_t>(), d_shi.as<const uint32_t>(), static_cast<uint32_t>(S),
                         ix.lm_subs.as<LeafModel>(), sfit.as<LeafFit>());
      GJ_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(split_error_kernel<K>, dim3(J), dim3(256), 0, s, col, d_lo.as<const uint32_t>(),
                         d_hi.as<const uint32_t>(), ix.lm_seps.as<const uint64_t>(), ix.lm_subs.as<const LeafModel>(),
                         sfit.as<LeafFit>());
      GJ_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(leaf_finalize_kernel, dim3(grid_for(S, 256)), dim3(256), 0, s, ix.lm_subs.as<LeafModel>(),
                         sfit.as<LeafFit>(), static_cast<uint32_t>(S), max_width);
      GJ_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(split_mark_kernel, dim3(grid_for(J, 256)), dim3(256), 0, s, ix.lm_leaves.as<LeafModel>(),
                         d_id.as<const uint32_t>(), J);
      GJ_HIP_CHECK(hipGetLastError());
      hs.resize(S);
      GJ_HIP_CHECK(hipMemcpyAsync(hs.data(), sfit.p, sfit.bytes, hipMemcpyDeviceToHost, s));
      GJ_HIP_CHECK(hipStreamSynchronize(s));
    }

    // window bounds over the (sub-)leaves that are not routed to the B-tree, and statistics
    uint64_t fb_leaves = 0, fb_keys = 0, wsum = 0, wkeys = 0;
    uint32_t wmax = 0;
    int32_t elo = 0, ehi = 0;
    std::vector<uint8_t> is_split(L, 0);
    for (const uint32_t l : split_id) is_split[l] = 1;
    std::vector<const LeafFit*> live_fits;
    for (uint32_t l = 0; l < L; ++l)
      if (!is_split[l]) live_fits.push_back(&hf[l]);  // split leaves count through their sub-leaves
    for (const auto& f : hs) live_fits.push_back(&f);
    for (const LeafFit* fp : live_fits) {
      const LeafFit& f = *fp;
      if (f.count == 0) continue;
      if (f.err_hi < f.err_lo) {
        ++fb_leaves;
        fb_keys += f.count;
      } else {
        const uint32_t w = static_cast<uint32_t>(f.err_hi - f.err_lo + 1);
        wmax = std::max(wmax, w);
        wsum += static_cast<uint64_t>(w) * f.count;
        wkeys += f.count;
        elo = std::min(elo, f.err_lo);
        ehi = std::max(ehi, f.err_hi);
      }
    }
    ix.lm_err_lo = elo;
    ix.lm_err_hi = ehi;
    ix.lm_root = root;
    ix.stats.root_segments = root.segments;
    ix.lm_num_leaves = L;
    ix.lm_fallback_any = fb_leaves ? 1u : 0u;
    ix.stats.leaves = L;
    ix.stats.fallback_leaves = fb_leaves;
    ix.stats.fallback_keys = fb_keys;
    ix.stats.split_leaves = J;
    ix.stats.max_window = wmax;
    ix.stats.avg_window = wkeys ? static_cast<double>(wsum) / static_cast<double>(wkeys) : 0.0;
    ix.stats.depth = 2;
    ix.stats.lds_levels = ix.bt.lds_levels;
    last_start = std::move(start);
    last_fit = std::move(fit);
    return fb_keys;
  };

  // compact leaves from the last build's 16-byte leaves; false (index unchanged) when a
  // leaf's window under the compact arithmetic exceeds the bound
  auto make_compact = [&](const RootModel& root) -> bool {
    const uint32_t lbits = ilog2_ceil(L);
    const uint32_t gshift = lbits > 9 ? lbits - 9 : 0;  // at most 512 groups (8 KiB of LDS)
    const uint32_t G = L >> gshift;
    dbuf groups(sizeof(LeafGroup) * G), cl(sizeof(LeafCompact) * L);
    hipLaunchKernelGGL(compact_groups_kernel<K>, dim3(grid_for(G, 256)), dim3(256), 0, s, col, n,
                       last_start.as<const uint32_t>(), G, gshift, groups.as<LeafGroup>());
    GJ_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(compact_convert_kernel, dim3(grid_for(L, 256)), dim3(256), 0, s,
                       ix.lm_leaves.as<const LeafModel>(), groups.as<const LeafGroup>(), L, gshift,
                       cl.as<LeafCompact>(), last_fit.as<LeafFit>());
    GJ_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(compact_error_kernel<K>, dim3(grid_for(n, 256)), dim3(256), 0, s, col, n,
                       Router<K>{kmin, root, L}, cl.as<const LeafCompact>(), groups.as<const LeafGroup>(), gshift,
                       last_fit.as<LeafFit>());
    GJ_HIP_CHECK(hipGetLastError());
    std::vector<LeafFit> hf(L);
    GJ_HIP_CHECK(hipMemcpyAsync(hf.data(), last_fit.p, last_fit.bytes, hipMemcpyDeviceToHost, s));
    GJ_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t max_width = 2 * static_cast<int64_t>(ix.params.learned_max_error) + 2;
    uint64_t wsum = 0, wkeys = 0;
    uint32_t wmax = 0;
    int32_t elo = 0, ehi = 0;
    for (const LeafFit& f : hf) {
      if (f.count == 0) continue;
      const int64_t w = static_cast<int64_t>(f.err_hi) - f.err_lo + 1;
      if (w > max_width) return false;
      wmax = std::max(wmax, static_cast<uint32_t>(w));
      wsum += static_cast<uint64_t>(w) * f.count;
      wkeys += f.count;
      elo = std::min(elo, f.err_lo);
      ehi = std::max(ehi, f.err_hi);
    }
    ix.lm_leaves = std::move(cl);
    ix.lm_groups = std::move(groups);
    ix.lm_gshift = gshift;
    ix.lm_compact = true;
    ix.lm_err_lo = elo;
    ix.lm_err_hi = ehi;
    ix.stats.max_window = wmax;
    ix.stats.avg_window = wkeys ? static_cast<double>(wsum) / static_cast<double>(wkeys) : 0.0;
    return true;
  };

  // With the default leaf count, a root that spreads the keys evenly gets twice the
  // leaves in compact form (the same table bytes); otherwise 16-byte leaves.  Below
  // 2^22 keys the lookups are bound by L2 hits rather than misses and the compact
  // form's LDS group table only costs occupancy, so small indexes keep 16-byte leaves.
  bool compact = false;
  if (ix.params.learned_leaf_bits == 0 && bits < 24 && root.segments == 0 && n >= (1ull << 22)) {
    L = 2u << bits;
    const uint64_t fb2 = build_leaves(root);
    compact = fb2 == 0 && ix.lm_splits == 0 && make_compact(root);
    if (!compact) L = 1u << bits;
  }
  if (!compact) {
    const uint64_t fb = build_leaves(root);
    if (ix.kind == GJ_LEARNED_PLR && ix.params.learned_root_segments == 0 && fb > 0 && n >= 4) {
      // the linear root cannot spread these keys evenly over the leaves: route by a
      // spline of the column's own CDF instead
      build_leaves(spline_root(2048));
    }
  }
}
template void build_learned<uint32_t>(gj_index&, hipStream_t);
template void build_learned<uint64_t>(gj_index&, hipStream_t);

}  // namespace gj

using namespace gj;

extern "C" {

int gj_mlp_loss_grad(const float* d_params, uint32_t hidden, const float* d_x, const float* d_y, uint64_t n,
                     float* d_grad, float* d_loss, gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_params && d_x && d_y && d_grad && d_loss && n > 0, GJ_ERR_INVALID, "bad argument");
    const hipStream_t s = S(stream);
    fill_bytes(d_grad, 0, (3 * hidden + 1) * sizeof(float), s);
    fill_bytes(d_loss, 0, sizeof(float), s);
    SampleSource<uint64_t> src{d_x, d_y, nullptr, n, 0, 0.0f, 0};
    launch_mlp_grad<uint64_t>(hidden, d_params, src, n, d_grad, d_loss, s);
  });
}

int gj_mlp_forward(const float* d_params, uint32_t hidden, const float* d_x, uint64_t n, float* d_out,
                   gj_stream_t stream) {
  return guarded([&] {
    GJ_REQUIRE(d_params && d_x && d_out, GJ_ERR_INVALID, "bad argument");
    GJ_REQUIRE(hidden >= 1 && hidden <= static_cast<uint32_t>(kMaxHidden), GJ_ERR_INVALID, "bad hidden width");
    if (n == 0) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), d_params, hidden, d_x, n,
                       d_out);
    GJ_HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
