// kernels.hip — gfx950 device code of the GBDT engine.
//
// Reference behaviour restated here (paths relative to crates/boosters/src/):
//   value_to_bin          data/binned/bin_mapper.rs:210-258
//   objectives            training/objectives/{regression.rs:20-42, classification.rs:20-187}
//   histograms            training/gbdt/histograms/ops.rs:79-283 (+ subtraction :345-351)
//   split scan            training/gbdt/split/find.rs:286-573, gain.rs:16-52,123-137
//   partition             training/gbdt/partition.rs:73-145,296-382
//   prediction            inference/gbdt/predictor.rs:197-302, repr/gbdt/tree_view.rs:119-160
//   transforms            model/transform.rs:83-115
//
// Numerics: compiled with -ffp-contract=off so every f32/f64 operation rounds
// as written (like rustc).  Gradients and hessians are quantised per round to
// 64-bit fixed point with a power-of-two scale 2^e chosen so that any sum over
// the rows fits in 62 bits; histogram sums are then exact integer sums,
// independent of atomic order, and are converted to f64 only for the gain.
#include "engine.h"

#include <float.h>
#include <math.h>

#include <stdexcept>
#include <type_traits>
#include <utility>

namespace bst {

constexpr int WAVE = 64;

__device__ __forceinline__ float exp32(float x) { return (float)exp((double)x); }

__device__ __forceinline__ long long shfl_up_ll(long long v, int d) {
  int lo = (int)(v & 0xFFFFFFFFLL), hi = (int)(v >> 32);
  lo = __shfl_up(lo, d, WAVE);
  hi = __shfl_up(hi, d, WAVE);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ unsigned long long shfl_xor_ull(unsigned long long v, int m) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo = __shfl_xor(lo, m, WAVE);
  hi = __shfl_xor(hi, m, WAVE);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ long long shfl_xor_ll(long long v, int m) {
  return (long long)shfl_xor_ull((unsigned long long)v, m);
}
// Inclusive scan over the 64 lanes of a wave.
__device__ __forceinline__ long long wave_incl_scan(long long v) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    long long o = shfl_up_ll(v, d);
    if (lane >= d) v += o;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_ull(unsigned long long v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    unsigned long long o = shfl_xor_ull(v, m);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_ll(v, m);
  return v;
}
// Order-preserving map of f32 to u32 (larger float -> larger integer).
__device__ __forceinline__ uint32_t ord_f32(float x) {
  uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// ===========================================================================
// Quantisation: value_to_bin
// ===========================================================================
// Numeric: first index i < n_search with v <= bound[i] (f64), else n_search.
// thr[] holds, per bound, the largest f32 not above it, so the test is done
// exactly in f32.  NaN -> n_bins-1 if the feature has a NaN bin, else the
// default bin.  Categorical: (i32)trunc(v) looked up in the sorted category
// list, unknown -> 0, NaN -> default bin.
__global__ __launch_bounds__(256) void k_value_to_bin(const float* __restrict__ X, int64_t n,
                                                      const BinMeta* __restrict__ meta,
                                                      const float* __restrict__ thr,
                                                      const int32_t* __restrict__ cats,
                                                      uint8_t* __restrict__ bins_cm,
                                                      uint32_t* __restrict__ hi_cm,
                                                      int64_t hi_words) {
  __shared__ float s_thr[kMaxBinsDev + 1];
  __shared__ int32_t s_cat[kMaxBinsDev + 1];
  const int e = blockIdx.y;
  const BinMeta m = meta[e];
  const int tabn = m.is_cat ? (int)m.n_bins : (int)m.n_search;
  for (int i = threadIdx.x; i < tabn; i += blockDim.x) {
    if (m.is_cat)
      s_cat[i] = cats[m.table_off + i];
    else
      s_thr[i] = thr[m.table_off + i];
  }
  __syncthreads();
  const float* col = X + (int64_t)m.orig * n;
  const int lane = threadIdx.x & (WAVE - 1);
  // Each wave owns aligned groups of 64 rows so the hi-bit ballot maps to two words.
  for (int64_t base = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)));
       base < n; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = base + lane;
    uint32_t b = 0;
    if (r < n) {
      const float v = col[r];
      if (isnan(v)) {
        b = m.missing_nan ? m.n_bins - 1 : m.default_bin;
      } else if (m.is_cat) {
        const double dv = (double)v;
        int32_t c = dv >= 2147483647.0 ? 2147483647
                                       : (dv <= -2147483648.0 ? (int32_t)0x80000000 : (int32_t)dv);
        int lo = 0, hi = (int)m.n_bins;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (s_cat[mid] < c)
            lo = mid + 1;
          else
            hi = mid;
        }
        b = (lo < (int)m.n_bins && s_cat[lo] == c) ? (uint32_t)lo : 0u;
      } else {
        int lo = 0, hi = (int)m.n_search;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (v <= s_thr[mid])
            hi = mid;
          else
            lo = mid + 1;
        }
        b = (uint32_t)lo;
      }
      bins_cm[(int64_t)e * n + r] = (uint8_t)(b & 0xFF);
    }
    const unsigned long long hb = __ballot(b >= 256u);
    if (hi_cm && lane == 0) {
      const int64_t w = base >> 5;
      if (w < hi_words) hi_cm[(int64_t)e * hi_words + w] = (uint32_t)hb;
      if (w + 1 < hi_words) hi_cm[(int64_t)e * hi_words + w + 1] = (uint32_t)(hb >> 32);
    }
  }
}

void launch_value_to_bin(hipStream_t s, const float* X, int64_t n, int E, const BinMeta* meta,
                         const float* thr, const int32_t* cats, uint8_t* bins_cm, uint32_t* hi_cm,
                         int64_t hi_words) {
  int64_t blocks_x = (n + 255) / 256;
  if (blocks_x > 2048) blocks_x = 2048;
  dim3 grid((unsigned)blocks_x, (unsigned)E);
  hipLaunchKernelGGL(k_value_to_bin, grid, dim3(256), 0, s, X, n, meta, thr, cats, bins_cm, hi_cm,
                     hi_words);
}

// Column-major -> row-major: 64 rows per block staged through LDS.
__global__ __launch_bounds__(256) void k_to_rowmajor(const uint8_t* __restrict__ bins_cm,
                                                     const uint32_t* __restrict__ hi_cm,
                                                     int64_t n, int E, int S, int hi_off,
                                                     int64_t hi_words, uint8_t* __restrict__ bins_rm) {
  extern __shared__ uint32_t s_tile32[];  // [64][S/4]
  uint8_t* s_tile = (uint8_t*)s_tile32;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  for (int i = threadIdx.x; i < 64 * S / 4; i += blockDim.x) s_tile32[i] = 0;
  __syncthreads();
  // Bins: thread -> (feature, row) with rows fastest (coalesced column reads).
  for (int i = threadIdx.x; i < 64 * E; i += blockDim.x) {
    const int e = i >> 6, r = i & 63;
    if (r0 + r < n) s_tile[r * S + e] = bins_cm[(int64_t)e * n + r0 + r];
  }
  if (hi_off >= 0) {
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * E; i += blockDim.x) {
      const int e = i >> 6, r = i & 63;
      const int64_t row = r0 + r;
      if (row < n) {
        uint32_t w = hi_cm[(int64_t)e * hi_words + (row >> 5)];
        if ((w >> (row & 31)) & 1u) atomicOr(&s_tile32[(r * S + hi_off + (e >> 3)) >> 2],
                                             1u << ((((r * S + hi_off + (e >> 3)) & 3) << 3) + (e & 7)));
      }
    }
  }
  __syncthreads();
  const int words = S / 4;
  uint32_t* out = (uint32_t*)(bins_rm + r0 * S);
  for (int i = threadIdx.x; i < 64 * words; i += blockDim.x) {
    const int r = i / words;
    if (r0 + r < n) out[i] = s_tile32[i];
  }
}

void launch_to_rowmajor(hipStream_t s, const DatasetDev& d) {
  const int64_t blocks = (d.n + 63) / 64;
  hipLaunchKernelGGL(k_to_rowmajor, dim3((unsigned)blocks), dim3(256), 64 * d.S, s, d.bins_cm,
                     d.hi_cm, d.n, d.E, d.S, d.hi_off, d.hi_words, d.bins_rm);
}

// ===========================================================================
// Objectives
// ===========================================================================
constexpr int kGradMaxK = 16;   // outputs handled by one gradient launch
constexpr int kGradThreads = 1024;
constexpr int kRowBatch = 4;  // rows per thread per pass in the streaming row kernels

__device__ __forceinline__ uint32_t abs_bits(float v) { return __float_as_uint(fabsf(v)); }

// grad/hess [K][n]; absmax[2k] = max|g|, absmax[2k+1] = max|h| (as f32 bits).
// Grid-stride over rows with per-thread running maxima; one atomicMax per
// block and value, so the maxima cost no global contention.
// KM: outputs the launch handles at most (1: one output and not softmax,
// so the per-output maxima and their reduction hold one entry).
// One output's (g, h) for the single-output objectives other than softmax:
// squared regression.rs:20-42, logistic classification.rs:24-52, absolute
// regression.rs:80-104, hinge classification.rs:87-116, pinball
// regression.rs:129-156 (alpha of this output), pseudo-Huber
// regression.rs:185-216, Poisson regression.rs:244-270.  f32 arithmetic in
// the reference's order.
__device__ __forceinline__ void objective_pair(int objective, const ObjParams& op, int k, float p0,
                                               float yi, float wi, float& gv, float& hv) {
  const float HESS_MIN = 1e-6f;
  if (objective == BST_OBJ_LOGISTIC) {
    const float p = 1.0f / (1.0f + exp32(-p0));
    gv = wi * (p - yi);
    hv = fmaxf(wi * p * (1.0f - p), HESS_MIN);
  } else if (objective == BST_OBJ_ABSOLUTE) {
    const float diff = p0 - yi;
    gv = wi * (isnan(diff) ? diff : copysignf(1.0f, diff));  // f32::signum
    hv = wi;
  } else if (objective == BST_OBJ_HINGE) {
    const float yy = yi > 0.5f ? 1.0f : -1.0f;
    const float margin = yy * p0;
    gv = margin < 1.0f ? -wi * yy : 0.0f;
    hv = wi;
  } else if (objective == BST_OBJ_PINBALL) {
    const float alpha = op.alphas[k];
    const float diff = p0 - yi;
    gv = wi * (diff < 0.0f ? -alpha : 1.0f - alpha);
    hv = wi;
  } else if (objective == BST_OBJ_PSEUDO_HUBER) {
    const float inv_delta_sq = 1.0f / (op.huber_delta * op.huber_delta);
    const float residual = p0 - yi;
    const float factor = 1.0f + residual * residual * inv_delta_sq;
    const float sqrt_factor = sqrtf(factor);
    gv = wi * residual / sqrt_factor;
    hv = wi / (factor * sqrt_factor);
  } else if (objective == BST_OBJ_POISSON) {
    float pc = p0;  // f32::clamp keeps NaN
    pc = pc < -30.0f ? -30.0f : pc;
    pc = pc > 30.0f ? 30.0f : pc;
    const float e = exp32(pc);
    gv = wi * (e - yi);
    hv = fmaxf(wi * e, HESS_MIN);
  } else {  // squared
    gv = wi * (p0 - yi);
    hv = wi;
  }
}

// Threads per block of k_gradients<KM>: the many-output instance keeps KM
// running maxima and the class loop's values per row, which fit in registers
// only at 256 threads per block (1024 would spill them to scratch).
template <int KM>
constexpr int grad_threads() { return KM == 1 ? kGradThreads : 256; }

template <int KM>
__global__ __launch_bounds__(grad_threads<KM>()) void k_gradients(
    int objective, ObjParams op, const float* __restrict__ preds, const float* __restrict__ y,
    const float* __restrict__ w, int64_t n, int K, float* __restrict__ g, float* __restrict__ h,
    uint32_t* __restrict__ absmax) {
  const float HESS_MIN = 1e-6f;
  constexpr int NT = grad_threads<KM>();
  __shared__ uint32_t s_max[2 * KM][NT / WAVE];
  uint32_t mx_g[KM], mx_h[KM];
#pragma unroll
  for (int k = 0; k < KM; k++) {
    mx_g[k] = 0u;
    mx_h[k] = 0u;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i_start = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (objective != 2 && K == 1) {
    // One output: kRowBatch rows per thread per pass, loads issued together.
    for (; i_start < n; i_start += kRowBatch * stride) {
      float p0[kRowBatch], yv[kRowBatch], wv_[kRowBatch];
#pragma unroll
      for (int k = 0; k < kRowBatch; k++) {
        const int64_t i = i_start + k * stride;
        const bool in = i < n;
        p0[k] = in ? preds[i] : 0.0f;
        yv[k] = in ? y[i] : 0.0f;
        wv_[k] = in && w ? w[i] : 1.0f;
      }
#pragma unroll
      for (int k = 0; k < kRowBatch; k++) {
        const int64_t i = i_start + k * stride;
        if (i < n) {
          float gv, hv;
          objective_pair(objective, op, 0, p0[k], yv[k], wv_[k], gv, hv);
          g[i] = gv;
          h[i] = hv;
          mx_g[0] = max(mx_g[0], abs_bits(gv));
          mx_h[0] = max(mx_h[0], abs_bits(hv));
        }
      }
    }
  }
  for (int64_t i = i_start; i < n; i += stride) {
    const float wi = w ? w[i] : 1.0f;
    const float yi = y[i];
    if (objective == 2) {  // softmax, classification.rs:141-187
      float mx = -INFINITY;
#pragma unroll
      for (int k = 0; k < KM; k++)
        if (k < K) mx = fmaxf(mx, preds[(int64_t)k * n + i]);
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < KM; k++)
        if (k < K) s += exp32(preds[(int64_t)k * n + i] - mx);
      const int label = (int)yi;
#pragma unroll
      for (int k = 0; k < KM; k++) {
        if (k < K) {
          const float p = exp32(preds[(int64_t)k * n + i] - mx) / s;
          const float ind = k == label ? 1.0f : 0.0f;
          const float gv = wi * (p - ind);
          const float hv = fmaxf(wi * p * (1.0f - p), HESS_MIN);
          g[(int64_t)k * n + i] = gv;
          h[(int64_t)k * n + i] = hv;
          mx_g[k] = max(mx_g[k], abs_bits(gv));
          mx_h[k] = max(mx_h[k], abs_bits(hv));
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < KM; k++) {
        if (k < K) {
          const float p0 = preds[(int64_t)k * n + i];
          float gv, hv;
          objective_pair(objective, op, k, p0, yi, wi, gv, hv);
          g[(int64_t)k * n + i] = gv;
          h[(int64_t)k * n + i] = hv;
          mx_g[k] = max(mx_g[k], abs_bits(gv));
          mx_h[k] = max(mx_h[k], abs_bits(hv));
        }
      }
    }
  }
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < KM; k++) {
    if (k < K) {
      uint32_t a = mx_g[k], b = mx_h[k];
#pragma unroll
      for (int m = WAVE / 2; m >= 1; m >>= 1) {
        a = max(a, (uint32_t)__shfl_xor(a, m, WAVE));
        b = max(b, (uint32_t)__shfl_xor(b, m, WAVE));
      }
      if (lane == 0) {
        s_max[2 * k][wv] = a;
        s_max[2 * k + 1][wv] = b;
      }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * K) {
    uint32_t v = 0u;
    for (int i = 0; i < NT / WAVE; i++) v = max(v, s_max[threadIdx.x][i]);
    atomicMax(&absmax[threadIdx.x], v);
  }
}

// Softmax with more classes than k_gradients keeps in registers
// (classification.rs:141-187): per row the same f32 operations in class order
// (the max, the sum of exps, each class's probability), each class's |g| and
// |h| maxima in LDS, then one atomicMax per value and block.
__global__ __launch_bounds__(kGradThreads) void k_softmax_wide(
    const float* __restrict__ preds, const float* __restrict__ y, const float* __restrict__ w,
    int64_t n, int K, float* __restrict__ g, float* __restrict__ h, uint32_t* __restrict__ absmax) {
  extern __shared__ uint32_t s_wmax[];  // [2 K]
  const float HESS_MIN = 1e-6f;
  for (int i = threadIdx.x; i < 2 * K; i += blockDim.x) s_wmax[i] = 0u;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float wi = w ? w[i] : 1.0f;
    const int label = (int)y[i];
    float mx = -INFINITY;
    for (int k = 0; k < K; k++) mx = fmaxf(mx, preds[(int64_t)k * n + i]);
    float s = 0.0f;
    for (int k = 0; k < K; k++) s += exp32(preds[(int64_t)k * n + i] - mx);
    for (int k = 0; k < K; k++) {
      const float p = exp32(preds[(int64_t)k * n + i] - mx) / s;
      const float ind = k == label ? 1.0f : 0.0f;
      const float gv = wi * (p - ind);
      const float hv = fmaxf(wi * p * (1.0f - p), HESS_MIN);
      g[(int64_t)k * n + i] = gv;
      h[(int64_t)k * n + i] = hv;
      atomicMax(&s_wmax[2 * k], abs_bits(gv));
      atomicMax(&s_wmax[2 * k + 1], abs_bits(hv));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 2 * K; i += blockDim.x)
    if (s_wmax[i]) atomicMax(&absmax[i], s_wmax[i]);
}

void launch_gradients(hipStream_t s, int objective, const ObjParams& op, const float* preds,
                      const float* y, const float* w, int64_t n, int K, float* g, float* h,
                      uint32_t* absmax) {
  // Callers validate 1 <= K <= kMaxOutputs, and K <= kGradMaxK unless softmax.
  if (n <= 0 || K < 1) return;
  // At most 256 blocks: each ends with one atomicMax per value on the same
  // words, and same-address atomics serialise in L2.
  int64_t blocks = (n + kGradThreads - 1) / kGradThreads;
  if (blocks > 256) blocks = 256;
  if (K > kGradMaxK) {
    if (objective != 2) return;
    hipLaunchKernelGGL(k_softmax_wide, dim3((unsigned)blocks), dim3(kGradThreads),
                       (size_t)2 * K * sizeof(uint32_t), s, preds, y, w, n, K, g, h, absmax);
    return;
  }
  if (K == 1 && objective != 2) {
    hipLaunchKernelGGL(k_gradients<1>, dim3((unsigned)blocks), dim3(kGradThreads), 0, s,
                       objective, op, preds, y, w, n, K, g, h, absmax);
  } else {
    constexpr int nt = grad_threads<kGradMaxK>();
    const int64_t b = std::min<int64_t>((n + nt - 1) / nt, 1024);
    hipLaunchKernelGGL(k_gradients<kGradMaxK>, dim3((unsigned)b), dim3(nt), 0, s, objective, op,
                       preds, y, w, n, K, g, h, absmax);
  }
}

// ---------------------------------------------------------------------------
// LambdaRank over ragged query groups (no counterpart in the reference; the
// published LambdaMART formulation with sigma = 1 and pair truncation):
// one block per group.  Documents are ranked by score descending, ties by
// index; a pair of positions (a, b) counts when the labels differ and
// min(a, b) < kRankTrunc;
//   rho = 1 / (1 + exp(s_hi - s_lo))                         (f32)
//   dNDCG = (2^l_hi - 2^l_lo) |disc[a] - disc[b]| / maxDCG    (f64)
// the higher-labelled document gets -rho dNDCG, the other +rho dNDCG, both
// rho (1 - rho) dNDCG.  Each document sums its pairs in f64 in position
// order, so the result does not depend on the launch shape; disc[r] =
// 1 / log2(2 + r) is a host-computed table.
//
// Every counted pair has a top position (< T = min(kRankTrunc, m)), so the
// pair terms of a group form a T x m matrix, walked in tiles of kRankChunk
// positions: all threads evaluate a tile's terms (from the top position's
// side) into LDS, then one thread per top position adds its row to its
// running sums and one thread per lower position (>= T) adds its column —
// the negated term, the same magnitude seen from the other side — and is
// done.  Only the ordered f64 additions are serial; the exp, divisions and
// table reads are spread over the block.  Keys (order-preserving score bits
// above the index) are sorted by rank counting for m <= kRankCountSortMax,
// by a bitonic network above.
constexpr int kRankTrunc = 30;
constexpr int kRankMaxGroup = 4096;
constexpr int kRankThreads = 256;
constexpr int kRankChunk = 32;
constexpr int kRankStride = kRankChunk + 1;  // row stride of the term tile: rows on distinct banks
constexpr int kRankCountSortMax = 2 * kRankThreads;

__device__ __forceinline__ uint32_t ord_f32_asc(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// Score and index held by a sort key ((~ord(score)) << 32 | index).
__device__ __forceinline__ float rank_key_score(unsigned long long key) {
  const uint32_t o = ~(uint32_t)(key >> 32);
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}
__device__ __forceinline__ uint32_t rank_key_index(unsigned long long key) {
  return (uint32_t)(key & 0xFFFFFFFFull);
}

__global__ __launch_bounds__(kRankThreads) void k_lambdarank(
    const float* __restrict__ preds, const float* __restrict__ y, const float* __restrict__ w,
    const int64_t* __restrict__ off, int64_t n_groups, const double* __restrict__ disc,
    float* __restrict__ g, float* __restrict__ h, uint32_t* __restrict__ absmax) {
  __shared__ unsigned long long s_key[kRankMaxGroup];
  __shared__ uint8_t s_l[kRankMaxGroup];
  __shared__ double s_tl[kRankTrunc * kRankStride];
  __shared__ double s_th[kRankTrunc * kRankStride];
  __shared__ uint32_t s_cnt[32];
  __shared__ double s_inv;
  __shared__ uint32_t s_max[2][kRankThreads / WAVE];
  const int tid = threadIdx.x;
  uint32_t mg = 0u, mh = 0u;
  auto emit = [&](uint32_t i, double ga, double ha, int64_t b0) {
    const double wi = w ? (double)w[b0 + i] : 1.0;
    const float gv = (float)(ga * wi), hv = (float)(ha * wi);
    g[b0 + i] = gv;
    h[b0 + i] = hv;
    mg = max(mg, abs_bits(gv));
    mh = max(mh, abs_bits(hv));
  };
  for (int64_t q = blockIdx.x; q < n_groups; q += gridDim.x) {
    const int64_t b0 = off[q];
    const int m = (int)(off[q + 1] - b0);
    const int T = m < kRankTrunc ? m : kRankTrunc;
    const bool count_sort = m <= kRankCountSortMax;
    int P = 1;
    while (P < m) P <<= 1;
    if (tid < 32) s_cnt[tid] = 0u;
    const int n_keys = count_sort ? m : P;
    for (int i = tid; i < n_keys; i += kRankThreads) {
      unsigned long long key = ~0ull;
      if (i < m) {
        const float sv = preds[b0 + i] + 0.0f;  // -0 ranks with +0
        int l = (int)y[b0 + i];
        l = l < 0 ? 0 : (l > 31 ? 31 : l);
        s_l[i] = (uint8_t)l;
        key = ((unsigned long long)(~ord_f32_asc(sv)) << 32) | (uint32_t)i;
      }
      s_key[i] = key;
    }
    __syncthreads();
    for (int i = tid; i < m; i += kRankThreads) atomicAdd(&s_cnt[s_l[i]], 1u);
    if (count_sort) {
      // Keys are distinct (the index is part of the key): a key's position is
      // the number of keys below it.
      const unsigned long long k0 = tid < m ? s_key[tid] : ~0ull;
      const unsigned long long k1 = tid + kRankThreads < m ? s_key[tid + kRankThreads] : ~0ull;
      uint32_t r0 = 0u, r1 = 0u;
      for (int j = 0; j < m; j++) {
        const unsigned long long kj = s_key[j];
        r0 += kj < k0 ? 1u : 0u;
        r1 += kj < k1 ? 1u : 0u;
      }
      __syncthreads();
      if (tid < m) s_key[r0] = k0;
      if (tid + kRankThreads < m) s_key[r1] = k1;
    } else {
      for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int i = tid; i < P; i += kRankThreads) {
            const int ixj = i ^ j;
            if (ixj > i) {
              const unsigned long long a = s_key[i], b = s_key[ixj];
              const bool up = (i & k) == 0;
              if ((a > b) == up) {
                s_key[i] = b;
                s_key[ixj] = a;
              }
            }
          }
          __syncthreads();
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      // maxDCG@T: labels in descending order, ranks ascending.
      double md = 0.0;
      int r = 0;
      for (int l = 31; l >= 0 && r < T; l--)
        for (uint32_t c = 0; c < s_cnt[l] && r < T; c++, r++) md += (ldexp(1.0, l) - 1.0) * disc[r];
      s_inv = md > 0.0 ? 1.0 / md : 0.0;
    }
    __syncthreads();
    const double inv = s_inv;
    if (inv > 0.0) {
      double ga = 0.0, ha = 0.0;  // row sums of top position `tid` (tid < T)
      for (int c0 = 0; c0 < m; c0 += kRankChunk) {
        const int cn = m - c0 < kRankChunk ? m - c0 : kRankChunk;
        for (int t = tid; t < T * kRankChunk; t += kRankThreads) {
          const int a = t / kRankChunk, bb = t % kRankChunk, b = c0 + bb;
          double lam = 0.0, hes = 0.0;
          if (bb < cn && b != a) {
            const unsigned long long ka = s_key[a], kb = s_key[b];
            const int li = s_l[rank_key_index(ka)], lj = s_l[rank_key_index(kb)];
            if (li != lj) {
              const bool ih = li > lj;
              const float si = rank_key_score(ka), sj = rank_key_score(kb);
              const float shi = ih ? si : sj, slo = ih ? sj : si;
              const float rho = 1.0f / (1.0f + exp32(shi - slo));
              const double gap = ldexp(1.0, ih ? li : lj) - ldexp(1.0, ih ? lj : li);
              const double dn = gap * fabs(disc[a] - disc[b]) * inv;
              const double l = (double)rho * dn;
              lam = ih ? -l : l;
              hes = (double)rho * (1.0 - (double)rho) * dn;
            }
          }
          s_tl[a * kRankStride + bb] = lam;
          s_th[a * kRankStride + bb] = hes;
        }
        __syncthreads();
        if (tid < T) {
          // Row of top position tid, in position order; the diagonal and
          // equal-label terms are +0, which leaves a sum that is never -0
          // unchanged.
          const double* rl = s_tl + tid * kRankStride;
          const double* rh = s_th + tid * kRankStride;
          for (int bb = 0; bb < cn; bb++) {
            ga += rl[bb];
            ha += rh[bb];
          }
        } else if (tid >= WAVE && tid < WAVE + cn && c0 + (tid - WAVE) >= T) {
          // Column of lower position b: its pairs are the top positions, in
          // order; its term is the top position's, negated.
          const int bb = tid - WAVE;
          double gb = 0.0, hb = 0.0;
          for (int a = 0; a < T; a++) {
            gb += -s_tl[a * kRankStride + bb];
            hb += s_th[a * kRankStride + bb];
          }
          emit(rank_key_index(s_key[c0 + bb]), gb, hb, b0);
        }
        __syncthreads();
      }
      if (tid < T) emit(rank_key_index(s_key[tid]), ga, ha, b0);
    } else {
      for (int a = tid; a < m; a += kRankThreads) emit(rank_key_index(s_key[a]), 0.0, 0.0, b0);
    }
    __syncthreads();  // LDS is reused by the next group
  }
#pragma unroll
  for (int m2 = WAVE / 2; m2 >= 1; m2 >>= 1) {
    mg = max(mg, (uint32_t)__shfl_xor(mg, m2, WAVE));
    mh = max(mh, (uint32_t)__shfl_xor(mh, m2, WAVE));
  }
  if ((tid & (WAVE - 1)) == 0) {
    s_max[0][tid / WAVE] = mg;
    s_max[1][tid / WAVE] = mh;
  }
  __syncthreads();
  if (tid < 2) {
    uint32_t v = 0u;
    for (int i = 0; i < kRankThreads / WAVE; i++) v = max(v, s_max[tid][i]);
    atomicMax(&absmax[tid], v);
  }
}

void launch_lambdarank(hipStream_t s, const float* preds, const float* y, const float* w,
                       const int64_t* group_off, int64_t n_groups, const double* disc, float* g,
                       float* h, uint32_t* absmax) {
  if (n_groups <= 0) return;
  const int64_t grid = n_groups < 16384 ? n_groups : 16384;
  hipLaunchKernelGGL(k_lambdarank, dim3((unsigned)grid), dim3(kRankThreads), 0, s, preds, y, w,
                     group_off, n_groups, disc, g, h, absmax);
}

// Scale exponent per (tree, g/h): largest e with n * max|x| * 2^e <= 2^62.
__global__ void k_qscale(uint32_t* absmax, int K, int64_t n_rows, int32_t* qexp) {
  for (int i = threadIdx.x; i < 2 * K; i += blockDim.x) {
    qexp[i] = qscale_exponent(__uint_as_float(absmax[i]), n_rows);
    absmax[i] = 0u;  // ready for the next round's gradients
  }
}

void launch_qscale(hipStream_t s, uint32_t* absmax, int K, int64_t n_rows, int32_t* qexp) {
  const int threads = std::min(1024, 64 * ((2 * K + 63) / 64));
  hipLaunchKernelGGL(k_qscale, dim3(1), dim3(threads), 0, s, absmax, K, n_rows, qexp);
}

// idx = row ids (sampled list or 0..m-1); ogh = gradient pairs in that order.
__global__ __launch_bounds__(256) void k_init_rows(uint32_t* __restrict__ idx,
                                                   float2* __restrict__ ogh,
                                                   const float* __restrict__ g,
                                                   const float* __restrict__ h,
                                                   const uint32_t* __restrict__ sampled, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t r = sampled ? sampled[i] : (uint32_t)i;
  idx[i] = r;
  ogh[i] = make_float2(g[r], h[r]);
}

void launch_init_rows(hipStream_t s, uint32_t* idx, float2* ogh, const float* g, const float* h,
                      const uint32_t* sampled, int64_t m) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_init_rows, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, idx, ogh, g,
                     h, sampled, m);
}

// Rows of every tree slot in order (slot t uses output out0 + t) and the
// exact fixed-point root totals, added to node_sum[t * max_nodes] (zeroed
// before).  Grid (row blocks, slots).
constexpr int kInitThreads = 1024;

__global__ __launch_bounds__(kInitThreads) void k_init_rows_sums(uint32_t* __restrict__ idx,
                                                        float2* __restrict__ ogh,
                                                        const float* __restrict__ g,
                                                        const float* __restrict__ h,
                                                        int64_t n, const int32_t* __restrict__ qexp,
                                                        uint32_t max_nodes,
                                                        long long* __restrict__ node_sum) {
  __shared__ long long s_red[kInitThreads / WAVE];
  const int t = blockIdx.y;
  const double sg = ldexp(1.0, qexp[2 * t]), sh = ldexp(1.0, qexp[2 * t + 1]);
  long long ag = 0, ah = 0;
  // kRowBatch rows per thread per pass, their loads issued together.
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n;
       i0 += kRowBatch * stride) {
    float2 gh[kRowBatch];
#pragma unroll
    for (int k = 0; k < kRowBatch; k++) {
      const int64_t i = i0 + k * stride;
      gh[k] = i < n ? make_float2(g[(int64_t)t * n + i], h[(int64_t)t * n + i])
                    : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int k = 0; k < kRowBatch; k++) {
      const int64_t i = i0 + k * stride;
      if (i < n) {
        idx[(int64_t)t * n + i] = (uint32_t)i;
        ogh[(int64_t)t * n + i] = gh[k];
        ag += __double2ll_rn((double)gh[k].x * sg);
        ah += __double2ll_rn((double)gh[k].y * sh);
      }
    }
  }
  ag = wave_sum_ll(ag);
  ah = wave_sum_ll(ah);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_red[threadIdx.x / WAVE] = ag;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long tg = 0;
    for (int i = 0; i < kInitThreads / WAVE; i++) tg += s_red[i];
    atomicAdd((unsigned long long*)&node_sum[2 * (int64_t)t * max_nodes], (unsigned long long)tg);
  }
  __syncthreads();
  if ((threadIdx.x & (WAVE - 1)) == 0) s_red[threadIdx.x / WAVE] = ah;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long th = 0;
    for (int i = 0; i < kInitThreads / WAVE; i++) th += s_red[i];
    atomicAdd((unsigned long long*)&node_sum[2 * (int64_t)t * max_nodes + 1], (unsigned long long)th);
  }
}

void launch_init_rows_sums(hipStream_t s, uint32_t* idx, float2* ogh, const float* g,
                           const float* h, int64_t n, int T, const int32_t* qexp,
                           uint32_t max_nodes, long long* node_sum) {
  if (n <= 0 || T <= 0) return;
  // At most 128 blocks per tree: each adds its totals with one atomic per
  // value on the root's words (same-address atomics serialise in L2).
  int64_t bx = (n + kInitThreads - 1) / kInitThreads;
  if (bx > 128) bx = 128;
  hipLaunchKernelGGL(k_init_rows_sums, dim3((unsigned)bx, (unsigned)T), dim3(kInitThreads), 0, s,
                     idx, ogh, g, h, n, qexp, max_nodes, node_sum);
}

// ===========================================================================
// Histograms
// ===========================================================================
// One block-iteration = one HistItem: a tile of a node's rows x one group of
// FG features.  Lane (threadIdx % FG) owns one feature and each wave holds
// WAVE / FG row slots, so the FG lanes of a row read FG consecutive bytes of
// that row.  The LDS sub-histogram is laid out [bin][FG]: lanes of one row
// always hit FG distinct 8-byte words, so ds_add_u64 is bank-conflict free
// whatever the bins are, and hot bins never collide inside a row.  Global
// layout per slot: [group][bin][FG] of (g, h) int64 pairs, which the flush
// writes with unit stride.
//
// The LDS pipe is the limit of this kernel: per step of a wave (one row per
// slot, FG features each) it performs the two 64-bit atomics (g, h), and
// nothing else goes through LDS.  Rows are staged in registers: the tile's
// rows are dealt to the 16 waves in wave-chunks of WR = slots x RPT rows,
// wave-chunk k going to wave k % 16, and lane j < RPT of each 16-lane DPP
// row of slot s loads row (s, j) of its wave's chunk (row id, (g, h)),
// quantises (g, h) to fixed point once and keeps the row's byte offset and
// the pair in registers.  A DPP row broadcast (row_newbcast) then hands row
// u's offset and pair to every lane of its slot: VALU moves instead of LDS
// reads.  Waves run independently between the tile's zeroing and its flush
// (two block barriers per tile, none per row).  Per wave-chunk a lane issues
// the bin loads of all RPT rows of its slot at once and then adds them, so
// RPT rows of gathers are in flight per lane and the other waves of the SIMD
// cover their latency; the staging loads (row ids, (g, h)) of the next
// wave-chunk are issued before and consumed after, so their latency is
// covered by this wave-chunk's adds.
//
// The add is a few VALU instructions per row and lane: no bounds tests (rows
// past the tile's end are staged as row 0 with (g, h) = 0, padding features
// read zero bytes of the padded row) and no select for the NaN bin
// (bin | hi << 8 is 256 for it; row 256 takes every add with nothing to
// count and is never read).
//
// XCD-aware item order: work-groups are dispatched round-robin over the 8
// XCDs (block b runs on XCD b % 8), each with its own L2.  Position p of the
// grid-stride walk maps to tile (p / (8 NG)) * 8 + p % 8 and group
// (p / 8) % NG, so the NG group blocks of one tile (which read disjoint
// pieces of the same rows) run on one XCD and share its L2 instead of each
// fetching every row from HBM.  The grid is a multiple of 8.
//
// OFF64: row byte offsets need 64 bits (n * S >= 2^32); otherwise they are
// 32-bit offsets from the uniform base, which keeps address registers small.
constexpr int kHistThreads = 1024;
constexpr int kHistWaves = kHistThreads / WAVE;
#ifndef BST_HIST_RPT16
#define BST_HIST_RPT16 8
#endif
// Rows per slot per wave-chunk (at most 16: one staging lane per row in each
// 16-lane DPP row): the bin gathers a lane keeps in flight.
template <int FG>
constexpr int hist_rpt() { return FG == 16 ? BST_HIST_RPT16 : 8; }
template <int FG>
constexpr int hist_waves_per_eu() { return FG == 16 ? 8 : 4; }

// Lane U's value of each 16-lane row, in every lane of that row.
template <int U>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + U, 0xF, 0xF, false);
}
template <int U>
__device__ __forceinline__ long long row_bcast64(long long v) {
  const uint32_t lo = row_bcast<U>((uint32_t)(unsigned long long)v);
  const uint32_t hi = row_bcast<U>((uint32_t)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}
template <typename F, int... U>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, U...>) {
  (f(std::integral_constant<int, U>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// A staged row in its staging lane: byte offset (row id with OFF64) and the
// quantised (g, h).
struct HistStage {
  uint32_t r;
  long long g, h;
};

template <int FG, bool GATHER, bool HAS_HI, bool OFF64>
__global__ __launch_bounds__(kHistThreads)
__attribute__((amdgpu_waves_per_eu(hist_waves_per_eu<FG>()))) void k_hist(
    const HistItem* __restrict__ items, uint32_t n_items_static, int NG,
    const uint32_t* __restrict__ idx, const float2* __restrict__ ogh, int64_t tree_stride,
    const uint8_t* __restrict__ bins_rm, int S, int hi_off, int E,
    const int32_t* __restrict__ qexp, long long* __restrict__ hist, int64_t slot_elems,
    long long* __restrict__ node_sum, long long* __restrict__ partial,
    const uint32_t* __restrict__ dyn, const uint32_t* __restrict__ idx_alt,
    const float2* __restrict__ ogh_alt) {
  // Bins 0..255, plus row 256 that takes every add with nothing to count.
  // One array, g plane then h plane, so that an h add reaches its column
  // with the g address plus an immediate offset.
  __shared__ long long s_gh[2 * 257 * FG];
  long long* const s_g = s_gh;
  long long* const s_h = s_gh + 257 * FG;
  constexpr int RS = WAVE / FG;        // row slots per wave
  constexpr int RPT = hist_rpt<FG>();  // rows per slot per wave-chunk
  constexpr int WR = RS * RPT;         // rows per wave-chunk
  static_assert(RPT <= 16 && RPT % 2 == 0, "wave-chunk shape");
  __shared__ long long s_red[2][kHistWaves];
  const uint32_t n_items = dyn ? *dyn : n_items_static;
  const uint32_t n_tiles = n_items / (uint32_t)NG;
  const uint32_t n_pos = ((n_tiles + 7u) / 8u) * 8u * (uint32_t)NG;
  const int wv = threadIdx.x / WAVE;
  const int ln = threadIdx.x % WAVE;
  const int lf = ln % FG;
  const int rs = ln / FG;
  // Staging lanes: lane j < RPT of each 16-lane DPP row stages row (rs, j);
  // with FG = 32 both DPP rows of a slot stage the same rows (each needs its
  // own copy for the broadcast) and only the first counts it in the totals.
  const int sj = ln & 15;
  const bool stager = sj < RPT;
  const bool counter = stager && lf < 16;
  constexpr uint32_t kPlaneBytes = 257u * FG * 8u;  // g plane to h plane
  char* const lane_g = reinterpret_cast<char*>(s_g + lf);
  for (uint32_t p = blockIdx.x; p < n_pos; p += gridDim.x) {
    const uint32_t tile = (p / (8u * NG)) * 8u + (p & 7u);
    const uint32_t grp = (p >> 3) % (uint32_t)NG;
    if (tile >= n_tiles) continue;  // block-uniform
    const uint32_t item_idx = tile * (uint32_t)NG + grp;
    const HistItem it = items[item_idx];
    const bool direct = (it.tree & 0x80000000u) != 0;  // this tile covers the whole node
    const bool alt = (it.tree & 0x40000000u) != 0;     // rows in the alternate buffer
    const uint32_t tree = it.tree & 0x3FFFFFFFu;
    for (int i = threadIdx.x; i < 256 * FG; i += blockDim.x) {
      s_g[i] = 0;
      s_h[i] = 0;
    }
    // Padding lanes (f >= E) load their own bytes (the row stride covers
    // every group's FG bin bytes and FG / 8 NaN-bit bytes) but add nothing:
    // their columns are never read, and the LDS atomics they skip (12 of 16
    // lanes in the last group of 100 features) are the kernel's bottleneck.
    const int f = (int)it.group * FG + lf;
    const bool live = f < E;
    const int hb = hi_off + (f >> 3), hs = f & 7;
    // The NaN bit (bit hs of its byte) shifted to the bin-256 bit of the
    // byte offset below: bit log2(256 * FG * 8).
    constexpr int kHiBit = FG == 16 ? 15 : 16;
    static_assert((1 << kHiBit) == 256 * FG * 8, "hi bit position");
    const uint32_t hsh = (uint32_t)(kHiBit - hs);
    const double sg = ldexp(1.0, qexp[2 * tree]);
    const double sh = ldexp(1.0, qexp[2 * tree + 1]);
    const uint32_t* tidx = (alt ? idx_alt : idx) + (int64_t)tree * tree_stride;
    const float2* togh = (alt ? ogh_alt : ogh) + (int64_t)tree * tree_stride;
    // This wave's wave-chunks: k = 0 .. nk - 1 are the tile's wave-chunks
    // k * 16 + wv.
    const uint32_t n_wc = (it.count + WR - 1) / WR;
    const uint32_t nk = n_wc > (uint32_t)wv ? (n_wc - (uint32_t)wv + kHistWaves - 1) / kHistWaves : 0u;
    long long accg = 0, acch = 0;
    // Raw staging registers of a staging lane: rows past the tile's end stage
    // row 0 with (0, 0), so the adds need no bounds checks (they add zeros).
    uint32_t st_r = 0;
    float2 st_gh = make_float2(0.0f, 0.0f);
    bool st_ok = false;
    auto stage_load = [&](uint32_t k) {
      const uint32_t q = (k * kHistWaves + (uint32_t)wv) * WR + (uint32_t)(rs * RPT + sj);
      st_ok = stager && q < it.count;
      if (st_ok) {
        const uint32_t pos = it.begin + q;
        st_r = GATHER ? tidx[pos] : pos;
        st_gh = togh[pos];
      }
    };
    auto stage_quant = [&](HistStage& st) {
      st.r = st_ok ? (OFF64 ? st_r : st_r * (uint32_t)S) : 0u;
      st.g = st_ok ? __double2ll_rn((double)st_gh.x * sg) : 0ll;
      st.h = st_ok ? __double2ll_rn((double)st_gh.y * sh) : 0ll;
      if (counter) {
        accg += st.g;
        acch += st.h;
      }
    };
    // Bin loads of a wave-chunk: row u's offset broadcast from its staging
    // lane.  The loaded bytes stay uint8_t until they are used: widened where
    // they are loaded, the compiler would widen them in another block and
    // wait for the loads there.
    auto load_bins = [&](const HistStage& st, uint8_t* bv, uint8_t* hv) {
      static_for<RPT>([&](auto U) {
        constexpr int u = decltype(U)::value;
        const uint32_t r = row_bcast<u>(st.r);
        if (OFF64) {
          const uint8_t* rp = bins_rm + (uint64_t)r * (uint64_t)S;
          bv[u] = rp[f];
          hv[u] = HAS_HI ? rp[hb] : 0u;
        } else {
          bv[u] = bins_rm[r + (uint32_t)f];
          hv[u] = HAS_HI ? bins_rm[r + (uint32_t)hb] : 0u;
        }
      });
    };
    auto add_rows = [&](const HistStage& st, const uint8_t (&cb)[RPT], const uint8_t (&ch)[RPT]) {
      static_for<RPT>([&](auto U) {
        constexpr int u = decltype(U)::value;
        const long long g = row_bcast64<u>(st.g);
        const long long h = row_bcast64<u>(st.h);
        // Byte offset of (bin, lf) in the g plane: bin * FG * 8 with
        // bin = byte | NaN bit << 8, formed as two shifted adds.
        uint32_t off = (uint32_t)cb[u] * (uint32_t)(FG * 8);
        if (HAS_HI) off |= ((uint32_t)ch[u] << hsh) & (256u * FG * 8u);
        char* const gp = lane_g + off;
        if (live) {
          atomicAdd((unsigned long long*)gp, (unsigned long long)g);
          atomicAdd((unsigned long long*)(gp + kPlaneBytes), (unsigned long long)h);
        }
      });
    };
    __syncthreads();  // the LDS zeroing is done before any wave adds
    HistStage sc{0u, 0ll, 0ll};
    if (nk > 0) {
      stage_load(0);
      stage_quant(sc);
    }
    for (uint32_t k = 0; k < nk; k++) {
      uint8_t cb[RPT], ch[RPT];
      load_bins(sc, cb, ch);
      if (k + 1 < nk) stage_load(k + 1);
      add_rows(sc, cb, ch);
      if (k + 1 < nk) stage_quant(sc);
    }
    const bool do_sum = node_sum != nullptr && it.group == 0;
    if (do_sum) {
      accg = wave_sum_ll(accg);
      acch = wave_sum_ll(acch);
      if (ln == 0) {
        s_red[0][wv] = accg;
        s_red[1][wv] = acch;
      }
    }
    __syncthreads();  // every wave's adds (and node-sum parts) are done
    if (do_sum && threadIdx.x == 0) {
      long long tg = 0, th = 0;
      for (int w = 0; w < kHistWaves; w++) {
        tg += s_red[0][w];
        th += s_red[1][w];
      }
      if (direct) {
        node_sum[2 * it.node] = tg;
        node_sum[2 * it.node + 1] = th;
      } else {
        atomicAdd((unsigned long long*)&node_sum[2 * it.node], (unsigned long long)tg);
        atomicAdd((unsigned long long*)&node_sum[2 * it.node + 1], (unsigned long long)th);
      }
    }
    long long* dst = hist + (int64_t)it.slot * slot_elems + (int64_t)it.group * 256 * FG * 2;
    if (direct || partial) {
      // Whole node: the slot itself.  One tile of several: this item's own
      // partial, summed by k_hist_finish (no global atomics).
      longlong2* out = direct ? (longlong2*)dst
                              : (longlong2*)(partial + (int64_t)item_idx * 256 * FG * 2);
      for (int i = threadIdx.x; i < 256 * FG; i += blockDim.x) {
        longlong2 v;
        v.x = s_g[i];
        v.y = s_h[i];
        out[i] = v;
      }
    } else {
      for (int i = threadIdx.x; i < 256 * FG; i += blockDim.x) {
        const long long gv = s_g[i], hv = s_h[i];
        if (gv | hv) {
          atomicAdd((unsigned long long*)&dst[2 * i], (unsigned long long)gv);
          atomicAdd((unsigned long long*)&dst[2 * i + 1], (unsigned long long)hv);
        }
      }
    }
    __syncthreads();  // the LDS sub-histogram is reused by the next item
  }
}

// The same items built with global-memory privatisation instead of LDS: the
// block owns its item's section in global memory — the node's slot for a
// tile that covers the whole node, the item's partial otherwise — zeroes it
// and adds each (row, feature) into it with two device atomics at L2; with
// no partial buffer a tile adds into the shared, pre-zeroed slot.  No LDS
// sub-histogram, so no zeroing and flushing of one per item and no LDS
// limit on the block count, but every (row, feature) pays two L2 atomics
// instead of two LDS atomics.  The sums are the same integers as k_hist's,
// bin 256 (the NaN bin of a 257-bin feature) is dropped the same way, and
// the node totals are formed the same way, so the outputs are identical; the
// autotuner measures both and keeps the faster per data shape.
template <int FG, bool GATHER, bool HAS_HI>
__global__ __launch_bounds__(kHistThreads) void k_hist_gpriv(
    const HistItem* __restrict__ items, uint32_t n_items_static, int NG,
    const uint32_t* __restrict__ idx, const float2* __restrict__ ogh, int64_t tree_stride,
    const uint8_t* __restrict__ bins_rm, int S, int hi_off, const int32_t* __restrict__ qexp,
    long long* __restrict__ hist, int64_t slot_elems, long long* __restrict__ node_sum,
    long long* __restrict__ partial, const uint32_t* __restrict__ dyn,
    const uint32_t* __restrict__ idx_alt, const float2* __restrict__ ogh_alt) {
  __shared__ long long s_red[2][kHistWaves];
  constexpr int RS = WAVE / FG;  // rows per wave per pass, one lane per feature
  const uint32_t n_items = dyn ? *dyn : n_items_static;
  const uint32_t n_tiles = n_items / (uint32_t)NG;
  const uint32_t n_pos = ((n_tiles + 7u) / 8u) * 8u * (uint32_t)NG;
  const int wv = threadIdx.x / WAVE;
  const int ln = threadIdx.x % WAVE;
  const int lf = ln % FG;
  const int rs = ln / FG;
  for (uint32_t p = blockIdx.x; p < n_pos; p += gridDim.x) {
    const uint32_t tile = (p / (8u * NG)) * 8u + (p & 7u);
    const uint32_t grp = (p >> 3) % (uint32_t)NG;
    if (tile >= n_tiles) continue;  // block-uniform
    const uint32_t item_idx = tile * (uint32_t)NG + grp;
    const HistItem it = items[item_idx];
    const bool direct = (it.tree & 0x80000000u) != 0;
    const bool alt = (it.tree & 0x40000000u) != 0;
    const uint32_t tree = it.tree & 0x3FFFFFFFu;
    long long* const slot_sec =
        hist + (int64_t)it.slot * slot_elems + (int64_t)it.group * 256 * FG * 2;
    const bool own = direct || partial != nullptr;
    long long* const sec = direct || !partial ? slot_sec : partial + (int64_t)item_idx * 256 * FG * 2;
    if (own)
      for (int i = threadIdx.x; i < 256 * FG * 2; i += blockDim.x) sec[i] = 0;
    __syncthreads();  // the section is zero before any wave adds into it
    const int f = (int)it.group * FG + lf;
    const int hb = hi_off + (f >> 3), hs = f & 7;
    const double sg = ldexp(1.0, qexp[2 * tree]);
    const double sh = ldexp(1.0, qexp[2 * tree + 1]);
    const uint32_t* tidx = (alt ? idx_alt : idx) + (int64_t)tree * tree_stride;
    const float2* togh = (alt ? ogh_alt : ogh) + (int64_t)tree * tree_stride;
    long long accg = 0, acch = 0;
    for (uint32_t q = (uint32_t)(wv * RS + rs); q < it.count; q += (uint32_t)(kHistWaves * RS)) {
      const uint32_t pos = it.begin + q;
      const uint32_t r = GATHER ? tidx[pos] : pos;
      const float2 v = togh[pos];
      const long long g = __double2ll_rn((double)v.x * sg);
      const long long h = __double2ll_rn((double)v.y * sh);
      if (lf == 0) {
        accg += g;
        acch += h;
      }
      const uint8_t* rp = bins_rm + (uint64_t)r * (uint64_t)S;
      uint32_t b = rp[f];
      if (HAS_HI) b |= ((uint32_t)(rp[hb] >> hs) & 1u) << 8;
      if (b < 256u) {
        long long* const e = sec + 2 * ((int64_t)b * FG + lf);
        atomicAdd((unsigned long long*)e, (unsigned long long)g);
        atomicAdd((unsigned long long*)(e + 1), (unsigned long long)h);
      }
    }
    const bool do_sum = node_sum != nullptr && it.group == 0;
    if (do_sum) {
      accg = wave_sum_ll(accg);
      acch = wave_sum_ll(acch);
      if (ln == 0) {
        s_red[0][wv] = accg;
        s_red[1][wv] = acch;
      }
    }
    __syncthreads();
    if (do_sum && threadIdx.x == 0) {
      long long tg = 0, th = 0;
      for (int w = 0; w < kHistWaves; w++) {
        tg += s_red[0][w];
        th += s_red[1][w];
      }
      if (direct) {
        node_sum[2 * it.node] = tg;
        node_sum[2 * it.node + 1] = th;
      } else {
        atomicAdd((unsigned long long*)&node_sum[2 * it.node], (unsigned long long)tg);
        atomicAdd((unsigned long long*)&node_sum[2 * it.node + 1], (unsigned long long)th);
      }
    }
    __syncthreads();  // s_red is reused by the next item
  }
}

// Finish the histograms of one step: for each built node (small child, or a
// root) and group, the node histogram is the sum of its tiles' partials (or
// was stored directly by a single tile), and the larger sibling, which holds
// the parent's histogram in its slot, becomes parent - small
// (histograms/pool.rs subtraction trick).  fin[4 i] = {small slot, large slot
// or ~0, first item index, tile count}; work = (node, group, 1024-value chunk).
// A node with one tile was stored directly; with none it is all zeros.  The
// device-driven schedule passes no large slot: its split does the
// subtraction while it stages the bins (launch_split's `minus`).
constexpr int kFinThreads = 256;
constexpr int kFinChunk = kFinThreads * 2;  // int64 values per work item

__global__ __launch_bounds__(kFinThreads) void k_hist_finish(
    const uint32_t* __restrict__ fin, uint32_t n_static, int NG, int FG,
    const long long* __restrict__ partial, long long* __restrict__ hist, int64_t slot_elems,
    const uint32_t* __restrict__ dyn) {
  const uint32_t V = 256u * (uint32_t)FG * 2u;  // int64 values per group section
  const uint32_t chunks = V / kFinChunk;
  const uint32_t per_node = (uint32_t)NG * chunks;
  // The block's first entry is read with the count when it lies inside the
  // list (n_static is the list's length, or its capacity with dyn).
  const uint32_t fi0 = blockIdx.x / per_node;
  uint4 e0 = make_uint4(0u, 0u, 0u, 0u);
  if (fi0 < n_static)
    e0 = make_uint4(fin[4 * fi0], fin[4 * fi0 + 1], fin[4 * fi0 + 2], fin[4 * fi0 + 3]);
  const uint32_t n = dyn ? *dyn : n_static;
  const uint32_t n_work = n * per_node;
  for (uint32_t w = blockIdx.x; w < n_work; w += gridDim.x) {
    const uint32_t fi = w / per_node, rem = w % per_node;
    const uint32_t g = rem / chunks, c = rem % chunks;
    const bool first = w == blockIdx.x;
    const uint32_t small = first ? e0.x : fin[4 * fi], large = first ? e0.y : fin[4 * fi + 1];
    const uint32_t h0 = first ? e0.z : fin[4 * fi + 2], nt = first ? e0.w : fin[4 * fi + 3];
    if (nt == 1 && large == 0xFFFFFFFFu) continue;  // stored directly, nothing to subtract
    const uint32_t v0 = c * kFinChunk + threadIdx.x * 2u;
    longlong2* sd = (longlong2*)(hist + (int64_t)small * slot_elems + (int64_t)g * V + v0);
    longlong2 a;
    if (nt == 1) {
      a = *sd;
    } else {  // nt == 0 (no rows) stores zeros
      long long x0 = 0, y0 = 0, x1 = 0, y1 = 0, x2 = 0, y2 = 0, x3 = 0, y3 = 0;
      const longlong2* src = (const longlong2*)(partial + (int64_t)(h0 + g) * V + v0);
      const int64_t step = (int64_t)NG * V / 2;  // longlong2 between a node's tiles
      uint32_t k = 0;
      for (; k + 4 <= nt; k += 4) {  // four independent loads in flight
        const longlong2 p0 = src[(k + 0) * step], p1 = src[(k + 1) * step];
        const longlong2 p2 = src[(k + 2) * step], p3 = src[(k + 3) * step];
        x0 += p0.x;
        y0 += p0.y;
        x1 += p1.x;
        y1 += p1.y;
        x2 += p2.x;
        y2 += p2.y;
        x3 += p3.x;
        y3 += p3.y;
      }
      for (; k < nt; k++) {
        const longlong2 p0 = src[k * step];
        x0 += p0.x;
        y0 += p0.y;
      }
      a = make_longlong2(x0 + x1 + x2 + x3, y0 + y1 + y2 + y3);
      *sd = a;
    }
    if (large != 0xFFFFFFFFu) {
      longlong2* ld = (longlong2*)(hist + (int64_t)large * slot_elems + (int64_t)g * V + v0);
      longlong2 p = *ld;
      p.x -= a.x;
      p.y -= a.y;
      *ld = p;
    }
  }
}

void launch_hist_finish(hipStream_t s, const uint32_t* fin, int n, int NG, int FG,
                        const long long* partial, long long* hist, int64_t slot_elems,
                        const uint32_t* dyn) {
  if (n <= 0) return;
  // Work items are (node, group, chunk): NG x 16 per node.  A grid sized for
  // the worst case would mostly start blocks that find no work in the
  // device-side count (a cost at every step); 512 covers the steps that have
  // real work (nodes of several tiles).
  int64_t grid = (int64_t)n * NG * (256 * FG * 2 / kFinChunk);
  if (grid > 512) grid = 512;
  hipLaunchKernelGGL(k_hist_finish, dim3((unsigned)grid), dim3(kFinThreads), 0, s, fin,
                     (uint32_t)n, NG, FG, partial, hist, slot_elems, dyn);
}

// Grid for a histogram launch: a multiple of 8 (one slot per XCD), at most
// kHistMaxGrid blocks; with a device-side count the caller's bound is used.
constexpr int kHistMaxGrid = 2048;

template <int FG>
static void launch_hist_fg(hipStream_t s, const DatasetDev& d, const HistItem* items, int n_items,
                           const uint32_t* idx, const float2* ogh, int64_t tree_stride,
                           const int32_t* qexp, long long* hist, int64_t slot_elems,
                           long long* node_sum, long long* partial, bool gather,
                           const uint32_t* dyn, const uint32_t* idx_alt, const float2* ogh_alt) {
  const bool hi = d.hi_off >= 0;
  int64_t want;
  if (dyn) {
    want = n_items;  // capacity bound of the device-side item list
  } else {
    const int64_t n_tiles = n_items / d.NG;
    want = (n_tiles + 7) / 8 * 8 * d.NG;
  }
  int64_t grid = want < kHistMaxGrid ? want : kHistMaxGrid;
  grid = (grid + 7) / 8 * 8;
#define BST_HIST_LAUNCH(G, H, O)                                                                \
  hipLaunchKernelGGL((k_hist<FG, G, H, O>), dim3((unsigned)grid), dim3(kHistThreads), 0, s, items, \
                     (uint32_t)n_items, d.NG, idx, ogh, tree_stride, d.bins_rm, d.S, d.hi_off,  \
                     d.E, qexp, hist, slot_elems, node_sum, partial, dyn, idx_alt, ogh_alt)
  if (d.hist_priv == 2) {
#define BST_GPRIV_LAUNCH(G, H)                                                                  \
  hipLaunchKernelGGL((k_hist_gpriv<FG, G, H>), dim3((unsigned)grid), dim3(kHistThreads), 0, s,    \
                     items, (uint32_t)n_items, d.NG, idx, ogh, tree_stride, d.bins_rm, d.S,      \
                     d.hi_off, qexp, hist, slot_elems, node_sum, partial, dyn, idx_alt, ogh_alt)
    if (gather) {
      if (hi) BST_GPRIV_LAUNCH(true, true); else BST_GPRIV_LAUNCH(true, false);
    } else {
      if (hi) BST_GPRIV_LAUNCH(false, true); else BST_GPRIV_LAUNCH(false, false);
    }
#undef BST_GPRIV_LAUNCH
    return;
  }
  const bool off64 = (uint64_t)d.n * (uint64_t)d.S >= (1ull << 32);
  if (off64) {
    if (gather) {
      if (hi) BST_HIST_LAUNCH(true, true, true); else BST_HIST_LAUNCH(true, false, true);
    } else {
      if (hi) BST_HIST_LAUNCH(false, true, true); else BST_HIST_LAUNCH(false, false, true);
    }
  } else {
    if (gather) {
      if (hi) BST_HIST_LAUNCH(true, true, false); else BST_HIST_LAUNCH(true, false, false);
    } else {
      if (hi) BST_HIST_LAUNCH(false, true, false); else BST_HIST_LAUNCH(false, false, false);
    }
  }
#undef BST_HIST_LAUNCH
}

void launch_hist(hipStream_t s, const DatasetDev& d, const HistItem* items, int n_items,
                 const uint32_t* idx, const float2* ogh, int64_t tree_stride, const int32_t* qexp,
                 long long* hist, int64_t slot_elems, long long* node_sum, bool gather,
                 int /*rows_per_block_hint*/, const uint32_t* dyn, long long* partial,
                 const uint32_t* idx_alt, const float2* ogh_alt) {
  if (n_items <= 0) return;
  if (!hist_layout_padded(d, d.FG))
    throw std::runtime_error("histogram rows must hold every group's padding features");
  if (d.FG == 16)
    launch_hist_fg<16>(s, d, items, n_items, idx, ogh, tree_stride, qexp, hist, slot_elems,
                       node_sum, partial, gather, dyn, idx_alt, ogh_alt);
  else
    launch_hist_fg<32>(s, d, items, n_items, idx, ogh, tree_stride, qexp, hist, slot_elems,
                       node_sum, partial, gather, dyn, idx_alt, ogh_alt);
}

// Exact node totals over row segments (used where no histogram is needed).
__global__ __launch_bounds__(256) void k_node_sums(const Segment* __restrict__ segs,
                                                   const float2* __restrict__ ogh,
                                                   int64_t tree_stride,
                                                   const int32_t* __restrict__ qexp,
                                                   long long* __restrict__ node_sum,
                                                   const uint32_t* __restrict__ dyn) {
  __shared__ long long s_red[2][256 / WAVE];
  const uint32_t n_work = dyn ? *dyn : gridDim.x;
  for (uint32_t wi = blockIdx.x; wi < n_work; wi += gridDim.x) {
  const Segment sg = segs[wi];
  const double a = ldexp(1.0, qexp[2 * sg.tree]), b = ldexp(1.0, qexp[2 * sg.tree + 1]);
  const float2* t = ogh + (int64_t)sg.tree * tree_stride;
  long long g = 0, h = 0;
  for (uint32_t p = threadIdx.x; p < sg.count; p += blockDim.x) {
    const float2 v = t[sg.begin + p];
    g += __double2ll_rn((double)v.x * a);
    h += __double2ll_rn((double)v.y * b);
  }
  g = wave_sum_ll(g);
  h = wave_sum_ll(h);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    s_red[0][threadIdx.x / WAVE] = g;
    s_red[1][threadIdx.x / WAVE] = h;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long tg = 0, th = 0;
    for (int i = 0; i < 256 / WAVE; i++) {
      tg += s_red[0][i];
      th += s_red[1][i];
    }
    atomicAdd((unsigned long long*)&node_sum[2 * sg.node], (unsigned long long)tg);
    atomicAdd((unsigned long long*)&node_sum[2 * sg.node + 1], (unsigned long long)th);
  }
  __syncthreads();
  }
}

void launch_node_sums(hipStream_t s, const Segment* segs, int n_segs, const float2* ogh,
                      int64_t tree_stride, const int32_t* qexp, long long* node_sum,
                      const uint32_t* dyn) {
  if (n_segs <= 0) return;
  hipLaunchKernelGGL(k_node_sums, dim3(n_segs), dim3(256), 0, s, segs, ogh, tree_stride, qexp,
                     node_sum, dyn);
}

// Subtraction trick: for each (large_slot, small_slot, large_node, small_node)
// hist[large] -= hist[small], node_sum[large] -= node_sum[small].
__global__ __launch_bounds__(256) void k_subtract(const uint32_t* __restrict__ quads,
                                                  long long* __restrict__ hist,
                                                  int64_t slot_elems,
                                                  long long* __restrict__ node_sum,
                                                  const uint32_t* __restrict__ dyn) {
  const uint32_t n_work = dyn ? *dyn : gridDim.y;
  for (uint32_t wi = blockIdx.y; wi < n_work; wi += gridDim.y) {
    const uint32_t* q = quads + 4 * wi;
    longlong2* dst = (longlong2*)(hist + (int64_t)q[0] * slot_elems);
    const longlong2* src = (const longlong2*)(hist + (int64_t)q[1] * slot_elems);
    const int64_t m = slot_elems / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
      longlong2 a = dst[i];
      const longlong2 b = src[i];
      a.x -= b.x;
      a.y -= b.y;
      dst[i] = a;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2)
      node_sum[2 * q[2] + threadIdx.x] -= node_sum[2 * q[3] + threadIdx.x];
  }
}

void launch_subtract(hipStream_t s, const uint32_t* quads, int n, long long* hist,
                     int64_t slot_elems, long long* node_sum, const uint32_t* dyn) {
  if (n <= 0) return;
  int64_t bx = (slot_elems / 2 + 255) / 256;
  if (bx > 64) bx = 64;
  if (bx < 1) bx = 1;
  hipLaunchKernelGGL(k_subtract, dim3((unsigned)bx, n), dim3(256), 0, s, quads, hist, slot_elems,
                     node_sum, dyn);
}

__global__ __launch_bounds__(256) void k_zero_slots(const uint32_t* __restrict__ slots,
                                                    long long* __restrict__ hist,
                                                    int64_t slot_elems,
                                                    const uint32_t* __restrict__ dyn) {
  const uint32_t n_work = dyn ? *dyn : gridDim.y;
  for (uint32_t wi = blockIdx.y; wi < n_work; wi += gridDim.y) {
    longlong2* dst = (longlong2*)(hist + (int64_t)slots[wi] * slot_elems);
    const int64_t m = slot_elems / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
      dst[i] = make_longlong2(0, 0);
  }
}

void launch_zero_slots(hipStream_t s, const uint32_t* slots, int n, long long* hist,
                       int64_t slot_elems, const uint32_t* dyn) {
  if (n <= 0) return;
  int64_t bx = (slot_elems / 2 + 255) / 256;
  if (bx > 64) bx = 64;
  hipLaunchKernelGGL(k_zero_slots, dim3((unsigned)bx, n), dim3(256), 0, s, slots, hist,
                     slot_elems, dyn);
}

// ===========================================================================
// Split finding: one wave per (node, feature)
// ===========================================================================
struct GainCtx {  // NodeGainContext, gain.rs:16-52
  double lambda, offset;
  __device__ GainCtx(double pg, double ph, const SplitParams& p) {
    lambda = p.lambda;
    const double ps = pg * pg / (ph + lambda);
    offset = 0.5 * ps + p.min_gain;
  }
  __device__ __forceinline__ float gain(double gl, double hl, double gr, double hr) const {
    const double sl = gl * gl / (hl + lambda);
    const double sr = gr * gr / (hr + lambda);
    return (float)(0.5 * (sl + sr) - offset);
  }
};

__device__ __forceinline__ bool valid_split(const SplitParams& p, double hl, double hr,
                                            uint32_t cl, uint32_t cr) {
  return hl >= p.min_child_weight && hr >= p.min_child_weight && cl >= p.min_samples_leaf &&
         cr >= p.min_samples_leaf;
}
__device__ __forceinline__ uint32_t sat_sub(uint32_t a, uint32_t b) { return a > b ? a - b : 0u; }

// Candidate key inside a feature: max gain, then earliest in scan order
// (forward bins ascending, then backward bins descending; strict '>' in the
// reference keeps the first maximum).
__device__ __forceinline__ unsigned long long cand_key(float gain, uint32_t rank) {
  return ((unsigned long long)ord_f32(gain) << 32) | (0xFFFFFFFFu - rank);
}

constexpr int kSplitWaves = 8;  // features per block

// Wave-wide best candidate key; the winning lane's left-child sums go to
// out.lg / out.lh (keys are unique per candidate, so one lane matches).
__device__ __forceinline__ unsigned long long wave_best(unsigned long long local, long long lg,
                                                        long long lh, DevSplit& out) {
  const unsigned long long wb = wave_max_ull(local);
  const unsigned long long who = __ballot(wb != 0ull && local == wb);
  const int src = who ? __ffsll((unsigned long long)who) - 1 : 0;
  out.lg = (long long)(((unsigned long long)(unsigned)__shfl((int)(lg >> 32), src, WAVE) << 32) |
                       (unsigned)__shfl((int)(lg & 0xFFFFFFFF), src, WAVE));
  out.lh = (long long)(((unsigned long long)(unsigned)__shfl((int)(lh >> 32), src, WAVE) << 32) |
                       (unsigned)__shfl((int)(lh & 0xFFFFFFFF), src, WAVE));
  return wb;
}

// Best split of one feature of one node, one wave (find.rs:286-573): bin b
// comes from `load(b)` as its (g, h) fixed-point pair; `ratio` / `order` are
// the wave's 256-entry scratch for the sorted categorical partition.  Lane 0
// writes the feature's candidate; every lane returns the feature's packed
// node key (gain, then the tie rule's feature order), 0 if nothing is valid.
template <bool SORTED, class BinLoad>
__device__ unsigned long long split_feature(BinLoad load, const FeatMeta& fm, uint32_t node,
                                            uint32_t f, long long G, long long H, uint32_t cnt,
                                            double ig, double ih, const SplitParams& p,
                                            const GainCtx& ctx, double* ratio, uint16_t* order,
                                            int E, DevSplit* __restrict__ cand) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint32_t nb = fm.n_bins;
  // Bins 4*lane .. 4*lane+3.
  long long vg[4], vh[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int b = 4 * lane + j;
    const bool in = (uint32_t)b < nb;
    const longlong2 v = in ? load((uint32_t)b) : make_longlong2(0, 0);
    vg[j] = v.x;
    vh[j] = v.y;
  }
  unsigned long long best = 0;  // 0 = no valid candidate
  long long bl_g = 0, bl_h = 0;  // left-child sums of this lane's best candidate
  DevSplit out;
  out.gain = -INFINITY;
  out.lg = 0;
  out.lh = 0;
  out.feature = f;
  out.bin = 0;
  out.flags = 0;
#pragma unroll
  for (int w = 0; w < kMaxCatWords; w++) out.cats[w] = 0;

  if (!fm.is_cat) {
    if (nb >= 2) {
      long long lg = vg[0] + vg[1] + vg[2] + vg[3], lh = vh[0] + vh[1] + vh[2] + vh[3];
      const long long ig_incl = wave_incl_scan(lg), ih_incl = wave_incl_scan(lh);
      long long pgv = ig_incl - lg, phv = ih_incl - lh;  // exclusive prefix before bin 4*lane
      // X[j]: prefix before bin 4*lane+j; P[j]: through it (X[j+1] = P[j]).
      const long long X0 = pgv, Y0 = phv;
      long long P[4], Q[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        pgv += vg[j];
        phv += vh[j];
        P[j] = pgv;
        Q[j] = phv;
      }
      // Backward scan bound: bins [1, max_excl) accumulate right; the missing
      // (last) bin stays left.
      const uint32_t max_excl = fm.has_missing ? nb - 1 : nb;
      const int me = (int)max_excl - 1;
      // Totals over bins 0..me: every loaded bin (bins past nb, and bin 256,
      // are zero: the NaN bin of a 257-bin feature is never accumulated)
      // minus the missing bin when it is one of bins 0..255.  (Selecting
      // P[me & 3] from the lane holding bin me would index registers at run
      // time, which the compiler does in scratch.)
      long long Pme = (long long)((unsigned long long)(unsigned)__shfl((int)(ig_incl >> 32), WAVE - 1, WAVE) << 32) |
                      (long long)(unsigned)__shfl((int)(ig_incl & 0xFFFFFFFF), WAVE - 1, WAVE);
      long long Qme = (long long)((unsigned long long)(unsigned)__shfl((int)(ih_incl >> 32), WAVE - 1, WAVE) << 32) |
                      (long long)(unsigned)__shfl((int)(ih_incl & 0xFFFFFFFF), WAVE - 1, WAVE);
      if (fm.has_missing && nb - 1 < 256u) {
        const longlong2 mv = load(nb - 1);
        Pme -= mv.x;
        Qme -= mv.y;
      }
      // Exclusive prefix (bins < b) for the backward candidates.
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t b = 4 * lane + j;
        // Forward: bins 0..b left (missing right), b in [0, nb-2].
        if (b + 1 < nb) {
          const long long GL = P[j], HL = Q[j];
          const uint32_t cl = b + 1, cr = sat_sub(cnt, cl);
          const double gl = (double)GL * ig, hl = (double)HL * ih;
          const double gr = (double)(G - GL) * ig, hr = (double)(H - HL) * ih;
          if (valid_split(p, hl, hr, cl, cr)) {
            const float gn = ctx.gain(gl, hl, gr, hr);
            const unsigned long long k = cand_key(gn, b);
            if (k > best) {
              best = k;
              bl_g = GL;
              bl_h = HL;
            }
          }
        }
        // Backward: bins b..me right, split at b-1, missing left; b in [1, me].
        if (fm.has_missing && b >= 1 && (int)b <= me) {
          const long long excl_g = j == 0 ? X0 : P[j - 1], excl_h = j == 0 ? Y0 : Q[j - 1];
          const long long GR = Pme - excl_g, HR = Qme - excl_h;
          const uint32_t cr = max_excl - b, cl = sat_sub(cnt, cr);
          const double gr = (double)GR * ig, hr = (double)HR * ih;
          const double gl = (double)(G - GR) * ig, hl = (double)(H - HR) * ih;
          if (valid_split(p, hl, hr, cl, cr)) {
            const float gn = ctx.gain(gl, hl, gr, hr);
            const unsigned long long k = cand_key(gn, 256u + (uint32_t)(me - (int)b));
            if (k > best) {
              best = k;
              bl_g = G - GR;
              bl_h = H - HR;
            }
          }
        }
      }
      best = wave_best(best, bl_g, bl_h, out);
      if (best) {
        const uint32_t rank = 0xFFFFFFFFu - (uint32_t)(best & 0xFFFFFFFFu);
        out.gain = unord_f32((uint32_t)(best >> 32));
        if (rank < 256u) {
          out.bin = rank;
          out.flags = 0;
        } else {
          const uint32_t b = (uint32_t)me - (rank - 256u);
          out.bin = b - 1;
          out.flags = 1;  // default_left
        }
      }
    }
  } else if (nb >= 2) {
    const uint32_t nc = nb;
    if (nc <= p.max_onehot_cats) {
      // One-hot: each category alone goes left (find.rs:472-508).
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t k = 4 * lane + j;
        if (k < nc) {
          const double gl = (double)vg[j] * ig, hl = (double)vh[j] * ih;
          const double gr = (double)(G - vg[j]) * ig, hr = (double)(H - vh[j]) * ih;
          if (valid_split(p, hl, hr, 1u, sat_sub(cnt, 1u))) {
            const float gn = ctx.gain(gl, hl, gr, hr);
            const unsigned long long key = cand_key(gn, k);
            if (key > best) {
              best = key;
              bl_g = vg[j];
              bl_h = vh[j];
            }
          }
        }
      }
      best = wave_best(best, bl_g, bl_h, out);
      if (best) {
        const uint32_t k = 0xFFFFFFFFu - (uint32_t)(best & 0xFFFFFFFFu);
        out.gain = unord_f32((uint32_t)(best >> 32));
        out.flags = 2;
        // Word index known at run time: written word by word so that `out`
        // stays in registers.
#pragma unroll
        for (int w = 0; w < kMaxCatWords; w++) out.cats[w] = (uint32_t)w == (k >> 5) ? 1u << (k & 31) : 0u;
      }
    } else if constexpr (SORTED) {
      // Sorted ratio partition (find.rs:511-573): categories with hess > 1e-10
      // sorted stably by -G/(H+1e-6), then a prefix scan in that order.  (A
      // launch without SORTED has no feature with more than max_onehot_cats
      // categories.)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t k = 4 * lane + j;
        if (k < 256) {
          const double hv = (double)vh[j] * ih;
          const bool valid = k < nc && hv > 1e-10;
          ratio[k] = valid ? -((double)vg[j] * ig) / (hv + 1e-6) : NAN;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      uint32_t m = 0;
      for (uint32_t k = 0; k < nc; k++) m += isnan(ratio[k]) ? 0u : 1u;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t k = 4 * lane + j;
        if (k < nc && !isnan(ratio[k])) {
          const double rk = ratio[k];
          uint32_t pos = 0;
          for (uint32_t q = 0; q < nc; q++) {
            const double rq = ratio[q];
            if (!isnan(rq) && (rq < rk || (rq == rk && q < k))) pos++;
          }
          order[pos] = (uint16_t)k;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (m >= 2) {
        // Positions 4*lane+j < m-1 are candidates; gather bin sums in sorted order.
        long long og[4], oh[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t q = 4 * lane + j;
          og[j] = 0;
          oh[j] = 0;
          if (q < m) {
            const uint32_t k = order[q];
            const longlong2 v = load(k);
            og[j] = v.x;
            oh[j] = v.y;
          }
        }
        const long long sgl = og[0] + og[1] + og[2] + og[3], shl = oh[0] + oh[1] + oh[2] + oh[3];
        long long pg_ = wave_incl_scan(sgl) - sgl, ph_ = wave_incl_scan(shl) - shl;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t q = 4 * lane + j;
          pg_ += og[j];
          ph_ += oh[j];
          if (q + 1 < m) {
            const uint32_t cl = q + 1, cr = sat_sub(cnt, cl);
            const double gl = (double)pg_ * ig, hl = (double)ph_ * ih;
            const double gr = (double)(G - pg_) * ig, hr = (double)(H - ph_) * ih;
            if (valid_split(p, hl, hr, cl, cr)) {
              const float gn = ctx.gain(gl, hl, gr, hr);
              const unsigned long long key = cand_key(gn, q);
              if (key > best) {
                best = key;
                bl_g = pg_;
                bl_h = ph_;
              }
            }
          }
        }
        best = wave_best(best, bl_g, bl_h, out);
        if (best) {
          const uint32_t qb = 0xFFFFFFFFu - (uint32_t)(best & 0xFFFFFFFFu);
          out.gain = unord_f32((uint32_t)(best >> 32));
          out.flags = 2;
          for (uint32_t q = 0; q <= qb; q++) {
            const uint32_t k = order[q];
#pragma unroll
            for (int w = 0; w < kMaxCatWords; w++)
              out.cats[w] |= (uint32_t)w == (k >> 5) ? 1u << (k & 31) : 0u;
          }
        }
      }
    }
  }
  // A categorical column with a missing bin (an exclusive feature bundle:
  // code 0, every feature of the bundle at its default) routes that bin by
  // default_left, which a categorical split leaves false (find.rs:472-573),
  // so its rows go right even when the category set counted them left.  The
  // left sums kept for the children are those of the rows that do go left,
  // as the reference takes the children's totals from their own rows
  // (grower.rs:424-440); the gain keeps the set's sums, as the reference's.
  if (fm.is_cat && fm.has_missing && best && out.flags == 2u) {
    const uint32_t mb = fm.default_bin;
    // The set's word holding bin mb, selected word by word so that `out`
    // stays in registers (an index known only at run time would put it in
    // scratch memory).
    uint32_t mword = 0u;
#pragma unroll
    for (int w = 0; w < kMaxCatWords; w++) mword = (uint32_t)w == (mb >> 5) ? out.cats[w] : mword;
    if (mb < 256u && ((mword >> (mb & 31)) & 1u)) {
      const longlong2 v = load(mb);
      out.lg -= v.x;
      out.lh -= v.y;
    }
  }
  if (lane == 0) cand[(int64_t)node * E + f] = out;
  if (!best) return 0ull;
  const uint32_t tb = p.rightmost ? f : (0xFFFFFFFFu - f);
  return ((unsigned long long)ord_f32(out.gain) << 32) | tb;
}

// Staged bin b of the block's feature w: features 257 entries apart and the
// low two bits of b rotated by bits 3-4 and 5-6 of b, so that the staging
// stores (8 consecutive lanes: 8 features of one bin) and the scans' loads
// (lane i: bins 4i .. 4i + 3 of one feature) both spread each 8 or 16
// consecutive lanes over distinct 16-byte bank groups (unpadded, they took 8
// and 4 passes).
__device__ __forceinline__ int split_lds_index(int w, int b) {
  return w * 257 + (b ^ (((b >> 3) + (b >> 5)) & 3));
}

// One block = one node x kSplitWaves consecutive candidate features.  The
// block first stages the features' 256 bins into LDS ([feature][bin]): with
// the [group][bin][FG] histogram layout the features' pairs of one bin are
// adjacent, so consecutive threads read consecutive 16-byte pairs and each
// 128-byte line is fetched once per block instead of once per wave.  Then
// wave w evaluates feature w from LDS.  Each wave's own inputs (feature
// meta, node totals and count, scales) are loaded before the staging so
// that their latency overlaps it.  SORTED: some categorical feature has more
// categories than max_onehot_cats, so waves need the 256-entry ratio / order
// scratch of the sorted partition; without it the block's LDS is the staged
// bins only and more blocks fit a CU.
template <bool SORTED>
__global__ __launch_bounds__(kSplitWaves * 64) void k_split(
    const uint32_t* __restrict__ nodes, const uint32_t* __restrict__ slots,
    const uint32_t* __restrict__ feats, int n_feats, const FeatMeta* __restrict__ fmeta, int FG,
    int NG, int E, long long* __restrict__ hist, int64_t slot_elems,
    const long long* __restrict__ node_sum, const uint32_t* __restrict__ node_count,
    const int32_t* __restrict__ qexp, uint32_t max_nodes, SplitParams p,
    unsigned long long* __restrict__ node_key, DevSplit* __restrict__ cand,
    const uint32_t* __restrict__ dyn, int feat_stride, const uint32_t* __restrict__ minus) {
  constexpr int kScratch = SORTED ? 256 : 1;
  __shared__ double s_ratio[kSplitWaves][kScratch];
  __shared__ uint16_t s_order[kSplitWaves][kScratch];
  __shared__ longlong2 s_bins[kSplitWaves * 257];
  __shared__ unsigned long long s_key[kSplitWaves];
  constexpr int kThreads = kSplitWaves * 64;
  const int wv = threadIdx.x / WAVE;
  const int fi = blockIdx.x * kSplitWaves + wv;
  // The first entry is read with the count (the grid never exceeds the
  // entry lists).
  const uint32_t node0 = nodes[blockIdx.y], slot0 = slots[blockIdx.y];
  const uint32_t mraw0 = minus ? minus[blockIdx.y] : 0xFFFFFFFFu;
  const uint32_t n_work = dyn ? *dyn : gridDim.y;
  for (uint32_t wy = blockIdx.y; wy < n_work; wy += gridDim.y) {  // block-uniform
    const bool first = wy == blockIdx.y;
    const uint32_t node = first ? node0 : nodes[wy];
    const uint32_t slot = first ? slot0 : slots[wy];
    const uint32_t mraw = first ? mraw0 : minus ? minus[wy] : 0xFFFFFFFFu;
    const uint32_t mslot = mraw & ~kMinusNoWriteback;
    const bool wback = (mraw & kMinusNoWriteback) == 0u;
    const uint32_t tree = node / max_nodes;
    long long* base = hist + (int64_t)slot * slot_elems;
    const long long* mbase =
        mraw != 0xFFFFFFFFu ? hist + (int64_t)mslot * slot_elems : nullptr;
    const uint32_t* fl = feats ? feats + (int64_t)tree * feat_stride : nullptr;
    // This wave's feature and inputs.
    const bool fok = fi < n_feats;
    const uint32_t f = fok ? (fl ? fl[fi] : (uint32_t)fi) : 0u;
    const FeatMeta fm = fmeta[f];
    const int32_t qg = qexp[2 * tree], qh = qexp[2 * tree + 1];
    const long long G = node_sum[2 * node], H = node_sum[2 * node + 1];
    const uint32_t cnt = node_count[node];
    // Stage: thread t loads bin t / kSplitWaves + pass * (kThreads / kSplitWaves)
    // of the block's feature t % kSplitWaves.  With a sibling to subtract,
    // the slot holds the parent's histogram: stage parent - sibling and
    // write it back, so the slot holds this node's histogram from now on
    // (the block's features only; no other block touches them), unless no
    // later step reads it.
    {
      const int w = threadIdx.x % kSplitWaves;
      const int fw = blockIdx.x * kSplitWaves + w;
      if (fw < n_feats) {
        const uint32_t fs = fl ? fl[fw] : (uint32_t)fw;
        const int grp = (int)fs / FG, lf = (int)fs % FG;
        if (mbase) {
          for (int b = threadIdx.x / kSplitWaves; b < 256; b += kThreads / kSplitWaves) {
            const int64_t e = ((int64_t)(grp * 256 + b) * FG + lf) * 2;
            longlong2 v = *(const longlong2*)(base + e);
            const longlong2 m = *(const longlong2*)(mbase + e);
            v.x -= m.x;
            v.y -= m.y;
            if (wback) *(longlong2*)(base + e) = v;
            s_bins[split_lds_index(w, b)] = v;
          }
        } else {
          for (int b = threadIdx.x / kSplitWaves; b < 256; b += kThreads / kSplitWaves) {
            const int64_t e = ((int64_t)(grp * 256 + b) * FG + lf) * 2;
            s_bins[split_lds_index(w, b)] = *(const longlong2*)(base + e);
          }
        }
      }
    }
    __syncthreads();
    unsigned long long key = 0ull;
    if (fok) {
      const double ig = ldexp(1.0, -qg), ih = ldexp(1.0, -qh);
      const GainCtx ctx((double)G * ig, (double)H * ih, p);
      auto load = [&](uint32_t b) { return s_bins[split_lds_index(wv, (int)b)]; };
      key = split_feature<SORTED>(load, fm, node, f, G, H, cnt, ig, ih, p, ctx, s_ratio[wv],
                                  s_order[wv], E, cand);
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) s_key[wv] = key;
    __syncthreads();  // also: s_bins is restaged for the block's next node
    // One atomic per (node, block): the node's key word is shared by every
    // feature block of the node, so per-feature atomics serialise on it.
    if (threadIdx.x == 0) {
      unsigned long long m = 0ull;
      for (int w = 0; w < kSplitWaves; w++) m = s_key[w] > m ? s_key[w] : m;
      if (m) atomicMax(&node_key[node], m);
    }
  }
}

__global__ void k_split_finalize(const uint32_t* __restrict__ nodes, int n_nodes, int E,
                                 int rightmost, unsigned long long* __restrict__ node_key,
                                 const DevSplit* __restrict__ cand,
                                 DevSplit* __restrict__ node_split,
                                 const uint32_t* __restrict__ dyn) {
  const int n_work = dyn ? (int)*dyn : n_nodes;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_work; i += gridDim.x * blockDim.x) {
  const uint32_t node = nodes[i];
  const unsigned long long k = node_key[node];
  node_key[node] = 0;
  DevSplit s;
  if (k == 0) {
    s.gain = -INFINITY;
    s.feature = 0;
    s.bin = 0;
    s.flags = 0;
    for (int w = 0; w < kMaxCatWords; w++) s.cats[w] = 0;
    s.lg = 0;
    s.lh = 0;
  } else {
    const uint32_t lo = (uint32_t)(k & 0xFFFFFFFFu);
    const uint32_t f = rightmost ? lo : (0xFFFFFFFFu - lo);
    s = cand[(int64_t)node * E + f];
  }
  node_split[node] = s;
  }
}

void launch_split(hipStream_t s, const DatasetDev& d, const uint32_t* nodes, const uint32_t* slots,
                  int n_nodes, const uint32_t* feats, int n_feats, const long long* hist,
                  int64_t slot_elems, const long long* node_sum, const uint32_t* node_count,
                  const int32_t* qexp, uint32_t max_nodes, SplitParams p,
                  unsigned long long* node_key, DevSplit* cand, DevSplit* node_split,
                  const uint32_t* dyn, bool finalize, int feat_stride, const uint32_t* minus) {
  long long* hw = const_cast<long long*>(hist);  // written only through `minus` entries
  if (n_nodes <= 0) return;
  dim3 grid((unsigned)((n_feats + kSplitWaves - 1) / kSplitWaves), (unsigned)n_nodes);
  if (d.max_cat_bins > (int32_t)p.max_onehot_cats)
    hipLaunchKernelGGL(k_split<true>, grid, dim3(kSplitWaves * WAVE), 0, s, nodes, slots, feats,
                       n_feats, d.meta, d.FG, d.NG, d.E, hw, slot_elems, node_sum, node_count,
                       qexp, max_nodes, p, node_key, cand, dyn, feat_stride, minus);
  else
    hipLaunchKernelGGL(k_split<false>, grid, dim3(kSplitWaves * WAVE), 0, s, nodes, slots, feats,
                       n_feats, d.meta, d.FG, d.NG, d.E, hw, slot_elems, node_sum, node_count,
                       qexp, max_nodes, p, node_key, cand, dyn, feat_stride, minus);
  if (!finalize) return;  // the caller reduces node_key / cand itself
  hipLaunchKernelGGL(k_split_finalize, dim3((n_nodes + 63) / 64), dim3(64), 0, s, nodes, n_nodes,
                     d.E, p.rightmost, node_key, cand, node_split, dyn);
}

// ===========================================================================
// Stable partition (ballot + prefix sums), three passes
// ===========================================================================
__device__ __forceinline__ bool goes_left(const DevSplit& sp, const FeatMeta& fm, uint32_t bin) {
  if (fm.has_missing && bin == fm.default_bin) return (sp.flags & 1u) != 0;
  if (sp.flags & 2u) return bin < 256u && ((sp.cats[bin >> 5] >> (bin & 31)) & 1u);
  return bin <= sp.bin;
}

constexpr int kPartThreads = 256;

// Pass 1: per tile, count rows that go left; leave a flag per position.
__global__ __launch_bounds__(kPartThreads) void k_part_count(
    const PartItem* __restrict__ items, const PartTile* __restrict__ tiles,
    const DevSplit* __restrict__ node_split, const FeatMeta* __restrict__ fmeta,
    const uint8_t* __restrict__ bins_cm, const uint32_t* __restrict__ hi_cm, int64_t n,
    int64_t hi_words, const uint32_t* __restrict__ idx, uint8_t* __restrict__ flag,
    uint32_t* __restrict__ tile_left, const uint32_t* __restrict__ dyn_tiles) {
  __shared__ uint32_t s_cnt[kPartThreads / WAVE];
  const uint32_t n_work = dyn_tiles ? *dyn_tiles : gridDim.x;
  for (uint32_t wi = blockIdx.x; wi < n_work; wi += gridDim.x) {
    const PartTile t = tiles[wi];
    const PartItem it = items[t.item];
    const DevSplit sp = node_split[it.node];
    const FeatMeta fm = fmeta[it.feature];
    const uint8_t* col = bins_cm + (int64_t)it.feature * n;
    const uint32_t* hcol = hi_cm ? hi_cm + (int64_t)it.feature * hi_words : nullptr;
    uint32_t c = 0;
    for (uint32_t p = threadIdx.x; p < t.count; p += kPartThreads) {
      const uint32_t pos = t.begin + p;
      const uint32_t row = idx[pos];
      uint32_t bin = col[row];
      if (hcol) bin |= ((hcol[row >> 5] >> (row & 31)) & 1u) << 8;
      const bool l = goes_left(sp, fm, bin);
      flag[pos] = l ? 1 : 0;
      c += l ? 1u : 0u;
    }
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) s_cnt[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int i = 0; i < kPartThreads / WAVE; i++) s += s_cnt[i];
      tile_left[wi] = s;
    }
    __syncthreads();
  }
}

// Pass 2: exclusive scan of left counts over the tiles of each node.
__global__ void k_part_scan(const PartItem* __restrict__ items, int n_items,
                            const uint32_t* __restrict__ tile_left, uint32_t* __restrict__ tile_off,
                            uint32_t* __restrict__ item_left, const uint32_t* __restrict__ dyn_items) {
  if (dyn_items) n_items = (int)*dyn_items;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += gridDim.x * blockDim.x) {
    const PartItem it = items[i];
    uint32_t s = 0;
    for (uint32_t k = 0; k < it.n_tiles; k++) {
      tile_off[it.tile0 + k] = s;
      s += tile_left[it.tile0 + k];
    }
    item_left[i] = s;
  }
}

// Pass 3: scatter each tile's rows to their stable destinations in tmp, then
// (pass 4) copy the node ranges back.
__global__ __launch_bounds__(kPartThreads) void k_part_scatter(
    const PartItem* __restrict__ items, const PartTile* __restrict__ tiles,
    const uint32_t* __restrict__ tile_off, const uint32_t* __restrict__ item_left,
    const uint8_t* __restrict__ flag, const uint32_t* __restrict__ idx,
    const float2* __restrict__ ogh, uint32_t* __restrict__ idx_tmp,
    float2* __restrict__ ogh_tmp, const uint32_t* __restrict__ dyn_tiles) {
  __shared__ uint32_t s_wave[kPartThreads / WAVE];
  __shared__ uint32_t s_carry;
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  const uint32_t n_work = dyn_tiles ? *dyn_tiles : gridDim.x;
  for (uint32_t wi = blockIdx.x; wi < n_work; wi += gridDim.x) {
    const PartTile t = tiles[wi];
    const PartItem it = items[t.item];
    const uint32_t L = item_left[t.item];
    const uint32_t lbefore = tile_off[wi];
    const uint32_t rbefore = (t.begin - it.begin) - lbefore;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t p0 = 0; p0 < t.count; p0 += kPartThreads) {
      const uint32_t p = p0 + threadIdx.x;
      const bool in = p < t.count;
      const bool l = in && flag[t.begin + p];
      const unsigned long long bal = __ballot(l);
      const uint32_t below = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) s_wave[wv] = __popcll(bal);
      __syncthreads();
      uint32_t woff = s_carry;
      for (int k = 0; k < wv; k++) woff += s_wave[k];
      if (in) {
        const uint32_t lrank = woff + below;  // lefts before p in this tile
        const uint32_t rrank = p - lrank;     // rights before p in this tile
        const uint32_t dst = l ? it.begin + lbefore + lrank : it.begin + L + rbefore + rrank;
        idx_tmp[dst] = idx[t.begin + p];
        ogh_tmp[dst] = ogh[t.begin + p];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t s = s_carry;
        for (int k = 0; k < kPartThreads / WAVE; k++) s += s_wave[k];
        s_carry = s;
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kPartThreads) void k_part_copyback(
    const PartTile* __restrict__ tiles, const uint32_t* __restrict__ idx_tmp,
    const float2* __restrict__ ogh_tmp, uint32_t* __restrict__ idx, float2* __restrict__ ogh,
    const uint32_t* __restrict__ dyn_tiles) {
  const uint32_t n_work = dyn_tiles ? *dyn_tiles : gridDim.x;
  for (uint32_t wi = blockIdx.x; wi < n_work; wi += gridDim.x) {
    const PartTile t = tiles[wi];
    for (uint32_t p = threadIdx.x; p < t.count; p += kPartThreads) {
      idx[t.begin + p] = idx_tmp[t.begin + p];
      ogh[t.begin + p] = ogh_tmp[t.begin + p];
    }
  }
}

// Children bookkeeping after the scan: counts of both children, the parent
// total copied to the right child, and the smaller child's total cleared
// (it is rebuilt from its rows; the larger one is parent - smaller).
__global__ void k_part_children(const PartItem* __restrict__ items, int n_items,
                                const uint32_t* __restrict__ item_left,
                                uint32_t* __restrict__ node_count,
                                long long* __restrict__ node_sum,
                                const uint32_t* __restrict__ dyn_items) {
  if (dyn_items) n_items = (int)*dyn_items;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += gridDim.x * blockDim.x) {
    const PartItem it = items[i];
    const uint32_t L = item_left[i], R = it.count - L;
    node_count[it.node] = L;
    node_count[it.right_node] = R;
    const long long g = node_sum[2 * it.node], h = node_sum[2 * it.node + 1];
    node_sum[2 * it.right_node] = g;
    node_sum[2 * it.right_node + 1] = h;
    const uint32_t small = L <= R ? it.node : it.right_node;
    node_sum[2 * small] = 0;
    node_sum[2 * small + 1] = 0;
  }
}

__device__ __forceinline__ long long block_sum_ll(long long v, long long* s_red) {
  v = wave_sum_ll(v);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_red[threadIdx.x / WAVE] = v;
  __syncthreads();
  long long t = 0;
  for (int i = 0; i < (int)(blockDim.x / WAVE); i++) t += s_red[i];
  __syncthreads();
  return t;
}

// One-pass stable partition of a tile of a node's rows (device-driven
// schedule; partition.rs:73-145 keeps both sides in row order).
//
// Every tile a block waits on is one of its node's earlier tiles and has a
// lower index.  When the grid covers every tile, block b takes tile b: work
// groups start in index order (each XCD walks its share of the grid in
// order), so the lowest unfinished tile always belongs to a running block.
// Otherwise tiles are claimed in order through a ticket counter, so every
// tile waited on was claimed by a running block.  Each tile counts its left
// rows (ballots + a block scan over (row pass, wave)), publishes that count
// at once, and then finds its exclusive prefix among the tiles of its node
// by a decoupled look-back: one wave reads up to 64 preceding tiles' status
// words (four windows of 64 per round trip), adds their counts back to the
// nearest tile that has published its
// inclusive prefix, and waits only for tiles that have not yet published a
// count (which they do before any waiting of their own).  With the prefix,
// every row's destination follows from its rank: left rows fill the node's
// range from its start in row order; right rows fill it from its end
// backwards, so the right child holds its rows in row order read from its
// last position down (kTileRev marks such a node as a source: its tiles
// start at their highest position and walk down).  The last tile of a node
// records the node's left count.  Rows are read strided (row j *
// kFusedThreads + thread), so every load is coalesced and each thread has
// kFusedRows independent row loads in flight.  The children's totals are not
// summed here: the split recorded the exact left sums.
//
// The look-back wait is bounded: past kLookbackSpins rounds without
// progress the tile sets *err and goes on (the trainer then fails the
// round), so no launch can hang.
constexpr int kFusedThreads = 1024;
constexpr int kFusedRows = kFusedTileRows / kFusedThreads;  // rows per thread
constexpr int kFusedWaves = kFusedThreads / WAVE;
static_assert(kFusedRows * kFusedWaves <= WAVE, "one prefix value per lane");
constexpr unsigned long long kStatAgg = 1ull << 62;     // tile count published
constexpr unsigned long long kStatPrefix = 2ull << 62;  // inclusive prefix published
constexpr uint32_t kLookbackSpins = 1u << 22;

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += (uint32_t)__shfl_xor((int)v, m, WAVE);
  return v;
}

__global__ __launch_bounds__(kFusedThreads) void k_part_fused(
    const PartItem* __restrict__ items, const PartTile* __restrict__ tiles,
    const DevSplit* __restrict__ node_split, const FeatMeta* __restrict__ fmeta,
    const uint8_t* __restrict__ bins_cm, const uint32_t* __restrict__ hi_cm, int64_t n,
    int64_t hi_words, uint32_t* __restrict__ idx0, float2* __restrict__ ogh0,
    uint32_t* __restrict__ idx1, float2* __restrict__ ogh1, uint32_t* __restrict__ cursor,
    unsigned long long* __restrict__ tstat, uint32_t* __restrict__ ticket,
    uint32_t* __restrict__ err, const uint32_t* __restrict__ dyn_tiles,
    const uint32_t* __restrict__ tree_output, float* __restrict__ preds, int64_t n_rows,
    uint32_t max_nodes) {
  __shared__ uint32_t s_l[kFusedRows * kFusedWaves];  // lefts per (row pass, wave)
  __shared__ uint32_t s_v[kFusedRows * kFusedWaves];  // rows per (row pass, wave)
  __shared__ uint32_t s_base[2];
  __shared__ uint32_t s_wi;
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  // The block's own tile is read with the count (the grid never exceeds the
  // tile buffer's capacity).
  const PartTile first = tiles[blockIdx.x];
  const uint32_t n_work = *dyn_tiles;
  if (blockIdx.x >= n_work) return;  // no tile, no ticket
  const bool direct = n_work <= gridDim.x;
  for (uint32_t round = 0;; round++) {
    uint32_t wi;
    if (direct) {
      if (round > 0) break;
      wi = blockIdx.x;
    } else {
      __syncthreads();  // the previous tile's shared values have been read
      if (threadIdx.x == 0) s_wi = atomicAdd(ticket, 1u);
      __syncthreads();
      wi = s_wi;
      if (wi >= n_work) break;  // block-uniform
    }
    const PartTile t = direct ? first : tiles[wi];
    const bool src1 = (t.src & kTileAlt) != 0;
    const bool rev = (t.src & kTileRev) != 0;  // source rows run down from t.begin
    const bool leaves = (t.src & kTileLeaves) != 0;  // block-uniform
    const uint32_t* sidx = src1 ? idx1 : idx0;
    const float2* sogh = src1 ? ogh1 : ogh0;
    uint32_t* didx = src1 ? idx0 : idx1;
    float2* dogh = src1 ? ogh0 : ogh1;
    uint32_t rid[kFusedRows];
    float2 gh[kFusedRows];
#pragma unroll
    for (int j = 0; j < kFusedRows; j++) {
      const uint32_t p = (uint32_t)j * kFusedThreads + threadIdx.x;
      // Positions past the tile's rows are not read: they stand for row 0
      // (which exists), so the bin loads below stay in bounds.
      const bool in = p < t.count;
      const uint32_t pos = rev ? t.begin - p : t.begin + p;
      rid[j] = in ? sidx[pos] : 0u;
      gh[j] = in && !leaves ? sogh[pos] : make_float2(0.0f, 0.0f);
    }
    const PartItem it = items[t.item];
    // The split decision comes with the item (goes_left on it): only a
    // categorical split reads its category words from the node's split.
    const uint32_t* cats = (it.split_flags & 2u) ? node_split[it.node].cats : nullptr;
    const uint8_t* col = bins_cm + (int64_t)it.feature * n;
    const uint32_t* hcol = hi_cm ? hi_cm + (int64_t)it.feature * hi_words : nullptr;
    uint32_t bins[kFusedRows];
#pragma unroll
    for (int j = 0; j < kFusedRows; j++) {
      uint32_t b = col[rid[j]];
      if (hcol) b |= ((hcol[rid[j] >> 5] >> (rid[j] & 31)) & 1u) << 8;
      bins[j] = b;
    }
    if (leaves) {
      // The children are leaves: add their values to the rows' predictions
      // (the rows of one tree are distinct, so the updates are independent;
      // the same float add as the leaf-range update).  No row moves.
      float* pr = preds + (int64_t)tree_output[it.node / max_nodes] * n_rows;
      float x[kFusedRows];
      bool l[kFusedRows];
#pragma unroll
      for (int j = 0; j < kFusedRows; j++) {
        const uint32_t p = (uint32_t)j * kFusedThreads + threadIdx.x;
        const uint32_t bn = bins[j];
        l[j] = bn == it.miss_bin ? (it.split_flags & 1u) != 0
               : cats ? bn < 256u && ((cats[bn >> 5] >> (bn & 31)) & 1u)
                      : bn <= it.split_bin;
        x[j] = p < t.count ? pr[rid[j]] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < kFusedRows; j++) {
        const uint32_t p = (uint32_t)j * kFusedThreads + threadIdx.x;
        if (p < t.count) pr[rid[j]] = x[j] + (l[j] ? it.leaf_l : it.leaf_r);
      }
      continue;  // block-uniform
    }
    unsigned long long lbal[kFusedRows], vbal[kFusedRows];
#pragma unroll
    for (int j = 0; j < kFusedRows; j++) {
      const uint32_t p = (uint32_t)j * kFusedThreads + threadIdx.x;
      const bool valid = p < t.count;
      const uint32_t bn = bins[j];
      const bool l =
          valid && (bn == it.miss_bin ? (it.split_flags & 1u) != 0
                    : cats ? bn < 256u && ((cats[bn >> 5] >> (bn & 31)) & 1u)
                           : bn <= it.split_bin);
      lbal[j] = __ballot(l);
      vbal[j] = __ballot(valid);
      if (lane == 0) {
        s_l[j * kFusedWaves + wv] = (uint32_t)__popcll(lbal[j]);
        s_v[j * kFusedWaves + wv] = (uint32_t)__popcll(vbal[j]);
      }
    }
    __syncthreads();
    // Exclusive prefix over (row pass, wave) in position order, then the
    // tile's place among its node's tiles; one wave.
    if (wv == 0) {
      constexpr int NS = kFusedRows * kFusedWaves;  // 64 = one value per lane
      uint32_t lv = lane < NS ? s_l[lane] : 0u, vv = lane < NS ? s_v[lane] : 0u;
      uint32_t li = lv, vi = vv;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t a = __shfl_up(li, d, 64), b = __shfl_up(vi, d, 64);
        if (lane >= d) {
          li += a;
          vi += b;
        }
      }
      const uint32_t TL = __shfl(li, 63, 64);
      if (lane < NS) {
        s_l[lane] = li - lv;
        s_v[lane] = vi - vv;
      }
      const uint32_t q = wi - it.tile0;  // the tile's index within its node
      if (lane == 0)
        __hip_atomic_store(&tstat[wi], (q == 0 ? kStatPrefix : kStatAgg) | TL, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      uint32_t excl = 0;
      if (q > 0) {
        constexpr int kWin = 4;  // windows of 64 tiles read per round trip
        int64_t j = (int64_t)wi - 1;
        const int64_t lo = (int64_t)it.tile0;
        uint32_t spins = 0;
        bool done = false;
        while (!done) {
          unsigned long long v[kWin];
#pragma unroll
          for (int w = 0; w < kWin; w++) {
            const int64_t jj = j - w * WAVE - lane;
            v[w] = jj >= lo ? __hip_atomic_load(&tstat[jj], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : 0ull;
          }
          bool retry = false;
#pragma unroll
          for (int w = 0; w < kWin; w++) {
            if (done || retry) continue;  // wave-uniform
            const bool valid = j - w * WAVE - lane >= lo;
            const uint32_t fl = (uint32_t)(v[w] >> 62);
            // Nearest preceding tile with its prefix (the node's first tile
            // always publishes one), and whether every tile up to it has a
            // count.
            const unsigned long long pm = __ballot(valid && fl == 2u);
            const int pl = pm ? __ffsll((long long)pm) - 1 : WAVE;
            const unsigned long long zm = __ballot(valid && fl == 0u && lane <= pl);
            if (zm != 0ull) {
              retry = true;  // read again from this window
              j -= w * WAVE;
            } else {
              excl += wave_sum_u32(valid && lane <= pl ? (uint32_t)v[w] : 0u);
              done = pl < WAVE;  // else: 64 counts, no prefix: the next window
            }
          }
          if (done) break;
          if (!retry) {
            j -= kWin * WAVE;
          } else if (++spins > kLookbackSpins) {
            if (lane == 0) atomicOr(err, 1u);
            break;
          }
        }
      }
      if (lane == 0) {
        const uint32_t incl = excl + TL;
        if (q > 0)
          __hip_atomic_store(&tstat[wi], kStatPrefix | incl, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        if (q + 1 == it.n_tiles) cursor[2 * t.item] = incl;  // the node's left count
        // Rows of the node before this tile (in row order), left ones among them.
        const uint32_t before = rev ? it.begin + it.count - 1u - t.begin : t.begin - it.begin;
        s_base[0] = excl;
        s_base[1] = before - excl;
      }
    }
    __syncthreads();
    const uint32_t lbase = it.begin + s_base[0];
    const uint32_t rtop = it.begin + it.count - 1u - s_base[1];  // next right row goes here, down
#pragma unroll
    for (int j = 0; j < kFusedRows; j++) {
      const uint32_t p = (uint32_t)j * kFusedThreads + threadIdx.x;
      if (p < t.count) {
        const bool l = (lbal[j] >> lane) & 1ull;
        const uint32_t lrank = s_l[j * kFusedWaves + wv] + (uint32_t)__popcll(lbal[j] & lt_mask);
        const uint32_t vrank = s_v[j * kFusedWaves + wv] + (uint32_t)__popcll(vbal[j] & lt_mask);
        const uint32_t dst = l ? lbase + lrank : rtop - (vrank - lrank);
        didx[dst] = rid[j];
        dogh[dst] = gh[j];
      }
    }
  }
}

void launch_partition_fused(hipStream_t s, const DatasetDev& d, const PartItem* items,
                            const PartTile* tiles, int grid_tiles, const DevSplit* node_split,
                            uint32_t* idx0, float2* ogh0, uint32_t* idx1, float2* ogh1,
                            uint32_t* cursor, unsigned long long* tstat, uint32_t* ticket,
                            uint32_t* err, const uint32_t* dyn_tiles,
                            const uint32_t* tree_output, float* preds, int64_t n_rows,
                            uint32_t max_nodes) {
  if (grid_tiles <= 0) return;
  hipLaunchKernelGGL(k_part_fused, dim3(grid_tiles), dim3(kFusedThreads), 0, s, items, tiles,
                     node_split, d.meta, d.bins_cm, d.hi_cm, d.n, d.hi_words, idx0, ogh0, idx1,
                     ogh1, cursor, tstat, ticket, err, dyn_tiles, tree_output, preds, n_rows,
                     max_nodes);
}

void launch_partition(hipStream_t s, const DatasetDev& d, const PartItem* items, int n_items,
                      const PartTile* tiles, int n_tiles, const DevSplit* node_split,
                      uint32_t* idx, float2* ogh, uint32_t* idx_tmp, float2* ogh_tmp,
                      uint8_t* flag, uint32_t* tile_left, uint32_t* tile_off,
                      uint32_t* item_left, uint32_t* node_count, long long* node_sum,
                      const uint32_t* dyn_items, const uint32_t* dyn_tiles) {
  if (n_items <= 0) return;
  hipLaunchKernelGGL(k_part_count, dim3(n_tiles), dim3(kPartThreads), 0, s, items, tiles,
                     node_split, d.meta, d.bins_cm, d.hi_cm, d.n, d.hi_words, idx, flag,
                     tile_left, dyn_tiles);
  hipLaunchKernelGGL(k_part_scan, dim3((n_items + 255) / 256), dim3(256), 0, s, items, n_items,
                     tile_left, tile_off, item_left, dyn_items);
  hipLaunchKernelGGL(k_part_scatter, dim3(n_tiles), dim3(kPartThreads), 0, s, items, tiles,
                     tile_off, item_left, flag, idx, ogh, idx_tmp, ogh_tmp, dyn_tiles);
  hipLaunchKernelGGL(k_part_copyback, dim3(n_tiles), dim3(kPartThreads), 0, s, tiles, idx_tmp,
                     ogh_tmp, idx, ogh, dyn_tiles);
  hipLaunchKernelGGL(k_part_children, dim3((n_items + 63) / 64), dim3(64), 0, s, items, n_items,
                     item_left, node_count, node_sum, dyn_items);
}

// ===========================================================================
// Prediction updates and traversal
// ===========================================================================
__global__ __launch_bounds__(256) void k_leaf_update(const Segment* __restrict__ segs,
                                                     const float* __restrict__ leaf_val,
                                                     const uint32_t* __restrict__ idx,
                                                     int64_t tree_stride, float* __restrict__ preds,
                                                     const uint32_t* __restrict__ tree_output,
                                                     int64_t n_rows,
                                                     const uint32_t* __restrict__ dyn,
                                                     const uint32_t* __restrict__ idx_alt) {
  const uint32_t n_work = dyn ? *dyn : gridDim.x;
  for (uint32_t wi = blockIdx.x; wi < n_work; wi += gridDim.x) {
    const Segment sg = segs[wi];
    const float v = leaf_val[sg.node];
    const uint32_t tree = sg.tree & 0x7FFFFFFFu;
    const uint32_t* t = ((sg.tree >> 31) ? idx_alt : idx) + (int64_t)tree * tree_stride;
    float* pr = preds + (int64_t)tree_output[tree] * n_rows;
    // A segment's rows are distinct, so the read-modify-writes of a batch
    // are independent: load kLeafBatch row ids, then their predictions,
    // then store, keeping the batch's gathers in flight together.
    constexpr int kLeafBatch = 8;
    for (uint32_t p0 = 0; p0 < sg.count; p0 += blockDim.x * kLeafBatch) {
      uint32_t r[kLeafBatch];
      float x[kLeafBatch];
#pragma unroll
      for (int k = 0; k < kLeafBatch; k++) {
        const uint32_t p = p0 + (uint32_t)k * blockDim.x + threadIdx.x;
        r[k] = p < sg.count ? t[sg.begin + p] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int k = 0; k < kLeafBatch; k++) x[k] = r[k] != 0xFFFFFFFFu ? pr[r[k]] : 0.0f;
#pragma unroll
      for (int k = 0; k < kLeafBatch; k++)
        if (r[k] != 0xFFFFFFFFu) pr[r[k]] = x[k] + v;
    }
  }
}

void launch_leaf_update(hipStream_t s, const Segment* segs, int n_segs, const float* leaf_val,
                        const uint32_t* idx, int64_t tree_stride, float* preds,
                        const uint32_t* tree_output, int64_t n_rows, const uint32_t* dyn,
                        const uint32_t* idx_alt) {
  if (n_segs <= 0) return;
  hipLaunchKernelGGL(k_leaf_update, dim3(n_segs), dim3(256), 0, s, segs, leaf_val, idx,
                     tree_stride, preds, tree_output, n_rows, dyn, idx_alt);
}

__device__ __forceinline__ float sigmoid32(float x) {  // transform.rs:106-115
  // f32::clamp keeps NaN (fminf/fmaxf would not).
  float c = x;
  if (c < -500.0f) c = -500.0f;
  if (c > 500.0f) c = 500.0f;
  if (c >= 0.0f) return 1.0f / (1.0f + exp32(-c));
  const float e = exp32(c);
  return e / (1.0f + e);
}

constexpr int kPredMaxK = 16;
struct PredGroup8 {  // eight consecutive per-tree words, read in one load
  uint32_t v[8];
};
constexpr int kPredChunk16 = kPredChunkNodes / 2;          // uint4 per node buffer
constexpr int kPredPer = (kPredChunk16 + kPredMaxRows - 1) / kPredMaxRows;  // staged uint4 per lane

// Categorical test of one raw value (tree_view.rs:131-146): (u32)x, Rust's
// saturating cast, in the bitset -> right child.
__device__ __forceinline__ bool pred_cat_right(float x, uint32_t start, uint32_t size,
                                               const uint32_t* __restrict__ cat_words) {
  const uint32_t cat = x <= 0.0f ? 0u : (x >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)x);
  const uint32_t w = cat >> 5;
  return size != 0 && w < size && ((cat_words[start + w] >> (cat & 31)) & 1u);
}

// Output transform of one row's accumulators (transform.rs:83-115).
__device__ __forceinline__ void pred_transform(float* acc, int K, int transform) {
  if (transform == 1) {
#pragma unroll
    for (int k = 0; k < kPredMaxK; k++)
      if (k < K) acc[k] = sigmoid32(acc[k]);
  } else if (transform == 2) {  // softmax, max-subtracted
    float mx = -INFINITY;
#pragma unroll
    for (int k = 0; k < kPredMaxK; k++)
      if (k < K) mx = fmaxf(mx, acc[k]);
    float ssum = 0.0f;
#pragma unroll
    for (int k = 0; k < kPredMaxK; k++)
      if (k < K) {
        acc[k] = exp32(acc[k] - mx);
        ssum += acc[k];
      }
    if (ssum > 0.0f) {
#pragma unroll
      for (int k = 0; k < kPredMaxK; k++)
        if (k < K) acc[k] /= ssum;
    }
  }
}

// Traversal with the forest staged in LDS, one lane per row.
//
// A block owns blockDim.x rows (staged feature-major in LDS when STAGE_X, as
// order keys behind a row of zeros, so a node test reads s_x[f + 1][lane]:
// consecutive lanes, conflict free) and walks the chunks [c0, c1) of the
// forest in order.  An LDS chunk's compact nodes are copied into one of two
// LDS buffers: the next chunk's global loads are issued before the current
// chunk is traversed and written to the other buffer after it, one barrier
// per chunk.  A lane follows kPredILP trees of a chunk at once (independent
// chains of dependent LDS loads in flight) for as many steps as the deepest
// of them has levels, each step one compare and one add (engine.h), then
// adds their leaves to its accumulators in forest order, so every row's f32
// sum is the reference's (predictor.rs:227-229,280-301).  Trees kept in the
// 16-byte layout are traversed from global memory the same way.
//
// MODE 0: out = transform(base + sum); MODE 1: out += sum (training updates);
// MODE 2: each tree's leaf value to leaf_out[t][row] (two-phase launch; the
// block's chunk range is blockIdx.y's).
extern __shared__ uint4 s_pred_mem[];  // [2][kPredChunk16] node buffers, then [F + 1][rows] keys

// The rows [row0, row0 + blockDim.x) of a block through the chunks [c0, c1):
// the body of k_predict_lds and of the serving kernel's tiles.
template <bool STAGE_X, int MODE, bool CAT>
__device__ __forceinline__ void pred_block(
    PredModel m, const float* __restrict__ X, int64_t n, int64_t x_stride, int64_t row0, int c0,
    int c1, int transform, float* __restrict__ out, int64_t out_stride,
    uint32_t* __restrict__ leaf_idx, int64_t leaf_stride, float* __restrict__ leaf_out) {
  uint4* s_mem = s_pred_mem;
  const int R = blockDim.x;
  uint32_t* s_x = reinterpret_cast<uint32_t*>(s_mem + 2 * kPredChunk16);
  const int lane = threadIdx.x;
  const int64_t row = row0 + lane;
  const bool in = row < n;
  const uint4* __restrict__ gn16 = reinterpret_cast<const uint4*>(m.nodes);
  const uint4* __restrict__ gc16 = reinterpret_cast<const uint4*>(m.cnodes);
  const int64_t* __restrict__ tree_off = m.tree_off;
  const uint32_t* __restrict__ ctree_off = m.ctree_off;
  const uint32_t* __restrict__ cdepth = m.cdepth;
  const uint32_t* __restrict__ cat_ref = m.cat_ref;
  const uint32_t* __restrict__ tree_group = m.tree_group;
  const uint32_t* __restrict__ cat_start = m.cat_start;
  const uint32_t* __restrict__ cat_size = m.cat_size;
  const uint32_t* __restrict__ cat_words = m.cat_words;
  const PredChunk* __restrict__ chunks = m.chunks;
  const int K = m.Kw;         // outputs accumulated here: m.g0 .. m.g0 + K - 1
  const bool one = m.K == 1;  // the model has one output (every tree adds to it)
  // Rows are read from X only when the model tests features (a forest of
  // leaves may come with no feature rows at all).
  const bool xok = in && m.n_features > 0;
  // The next LDS chunk's nodes, in flight during the current chunk.  Loads
  // past the chunk's end read its first pair instead (always in range);
  // only the chunk's own pairs are stored.
  // Seven 16-byte pairs per lane cover a 1792-pair buffer at 256 lanes; with
  // fewer lanes (R = 128 or 64 rows) the pairs past 7 R are copied when the
  // buffer is written.
  static_assert(kPredPer == 7, "seven staged pairs per lane");
  uint4 p0, p1, p2, p3, p4, p5, p6;
#define BST_PRED_L1(pj, j)                                                          \
  {                                                                                 \
    const uint32_t i_ = (uint32_t)(lane + (j) * R);                                 \
    pj = gc16[b16_ + (i_ < n16_ ? i_ : 0u)];                                        \
  }
#define BST_PRED_S1(pj, j, dst)                                                     \
  {                                                                                 \
    const uint32_t i_ = (uint32_t)(lane + (j) * R);                                 \
    if (i_ < n16_) (dst)[i_] = pj;                                                  \
  }
  // Loads of a chunk's pairs (a global-memory chunk, n_nodes 0, loads its
  // first pair and stores nothing); loads past the end read the first pair.
#define BST_PRED_LOAD(ch)                                                           \
  do {                                                                              \
    const uint32_t n16_ = ((ch).n_nodes + 1) / 2, b16_ = (ch).node_begin / 2;       \
    BST_PRED_L1(p0, 0) BST_PRED_L1(p1, 1) BST_PRED_L1(p2, 2) BST_PRED_L1(p3, 3)     \
    BST_PRED_L1(p4, 4) BST_PRED_L1(p5, 5) BST_PRED_L1(p6, 6)                         \
  } while (0)
#define BST_PRED_STORE(ch, dst)                                                     \
  do {                                                                              \
    const uint32_t n16_ = ((ch).n_nodes + 1) / 2, b16_ = (ch).node_begin / 2;       \
    BST_PRED_S1(p0, 0, dst) BST_PRED_S1(p1, 1, dst) BST_PRED_S1(p2, 2, dst)         \
    BST_PRED_S1(p3, 3, dst) BST_PRED_S1(p4, 4, dst) BST_PRED_S1(p5, 5, dst)         \
    BST_PRED_S1(p6, 6, dst)                                                         \
    for (uint32_t i_ = (uint32_t)(lane + kPredPer * R); i_ < n16_; i_ += (uint32_t)R) \
      (dst)[i_] = gc16[b16_ + i_];                                                  \
  } while (0)
  float acc[kPredMaxK];
  if (MODE != 2) {
#pragma unroll
    for (int k = 0; k < kPredMaxK; k++)
      acc[k] = (k < K && in) ? (MODE == 1 ? out[(int64_t)k * out_stride + row] : 0.0f) : 0.0f;
    if (MODE == 0) {
      // The base scores are padded past the last output, so the window's
      // loads need no condition (and issue together).
      float b[kPredMaxK];
#pragma unroll
      for (int k = 0; k < kPredMaxK; k++) b[k] = m.base[m.g0 + k];
#pragma unroll
      for (int k = 0; k < kPredMaxK; k++) acc[k] = (k < K && in) ? b[k] : 0.0f;
    }
  }
  // The first chunk's loads are issued before the rows are staged.
  if (c0 < c1) BST_PRED_LOAD(chunks[c0]);
  if (STAGE_X) {
    // Staged row 0: zeros (what leaves read); row f + 1: feature f's keys.
    // 32 feature rows' loads in flight at a time (with the first chunk's).
    const int nf = m.n_features;
    const int64_t r = in ? row : 0;
    s_x[lane] = 0u;
    uint32_t* sx1 = s_x + R;
    int f = 0;
    for (; f + 32 <= nf; f += 32) {
      float v[32];
#pragma unroll
      for (int q = 0; q < 32; q++) v[q] = X[(int64_t)(f + q) * x_stride + r];
#pragma unroll
      for (int q = 0; q < 32; q++)
        sx1[(f + q) * R + lane] = in ? pred_order_key(__float_as_uint(v[q])) : 0u;
    }
    for (; f + 8 <= nf; f += 8) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = X[(int64_t)(f + q) * x_stride + r];
#pragma unroll
      for (int q = 0; q < 8; q++)
        sx1[(f + q) * R + lane] = in ? pred_order_key(__float_as_uint(v[q])) : 0u;
    }
    for (; f < nf; f++)
      sx1[f * R + lane] = in ? pred_order_key(__float_as_uint(X[(int64_t)f * x_stride + r])) : 0u;
  }
  if (c0 < c1) BST_PRED_STORE(chunks[c0], s_mem);
  __syncthreads();
  // The current chunk's descriptor is carried over from the previous
  // iteration's next one: one descriptor load per chunk.
  PredChunk ch = chunks[c0 < c1 ? c0 : 0];
  for (int c = c0; c < c1; c++) {
    const bool next = c + 1 < c1;
    const PredChunk chn = chunks[next ? c + 1 : c];
    if (next) BST_PRED_LOAD(chn);
    const uint2* sn = reinterpret_cast<const uint2*>(s_mem + ((c - c0) & 1) * kPredChunk16);
    for (uint32_t t0 = ch.first_tree; t0 < ch.end_tree; t0 += kPredILP) {
      // Chain j follows tree t0 + j from its root; a chain past the chunk's
      // last tree starts (and stays) on the sentinel leaf.
      uint32_t leafid[kPredILP];
      float leafv[kPredILP];
      if (ch.n_nodes) {
        uint32_t pos[kPredILP];  // node index in the chunk
        uint32_t steps = 0;      // levels of the group's deepest tree (uniform)
        // The group's roots and levels as two 8-entry loads (the arrays are
        // padded by 8 entries past the last tree).
        static_assert(kPredILP <= 8, "one 8-entry load per array");
        const PredGroup8 roots = *reinterpret_cast<const PredGroup8*>(ctree_off + t0);
        const PredGroup8 levels = *reinterpret_cast<const PredGroup8*>(cdepth + t0);
#pragma unroll
        for (int j = 0; j < kPredILP; j++) {
          const bool ok = t0 + j < ch.end_tree;
          pos[j] = ok ? roots.v[j] - ch.node_begin : ch.n_nodes - 1;
          steps = max(steps, ok ? levels.v[j] : 0u);
        }
        // Every chain takes the same number of steps; a chain on a leaf
        // reads staged row 0 (key 0) and stays (engine.h).
        for (uint32_t d = 0; d < steps; d++) {
          uint2 v[kPredILP];
#pragma unroll
          for (int j = 0; j < kPredILP; j++) v[j] = sn[pos[j]];
          uint32_t kx[kPredILP];
#pragma unroll
          for (int j = 0; j < kPredILP; j++) {
            const uint32_t f = v[j].x & 0xFFFFu;  // staged row
            if (STAGE_X)
              kx[j] = s_x[f * R + lane];
            else
              kx[j] = f != 0u && xok
                          ? pred_order_key(__float_as_uint(X[(int64_t)(f - 1u) * x_stride + row]))
                          : 0u;
          }
#pragma unroll
          for (int j = 0; j < kPredILP; j++) {
            const uint32_t w = v[j].x;
            uint32_t r = kx[j] + ((w >> 17) & 1u) >= v[j].y ? 1u : 0u;
            if (CAT && (w & kCCat) && kx[j] != 0xFFFFFFFFu) {
              const uint32_t c = v[j].y - 1u;
              r = pred_cat_right(__uint_as_float(pred_key_bits(kx[j])), cat_ref[2 * c],
                                 cat_ref[2 * c + 1], cat_words) ? 1u : 0u;
            }
            pos[j] = (w >> kCShift) + r;
          }
        }
#pragma unroll
        for (int j = 0; j < kPredILP; j++) {
          leafv[j] = __uint_as_float(~sn[pos[j]].y);
          leafid[j] = pos[j];  // the original id is looked up below when wanted
        }
      } else {
        // One tree in the 16-byte layout, from global memory.
        const int64_t b = tree_off[t0];
        uint32_t nd = 0;
        uint4 v = gn16[b];
        while (!(v.x & kNodeLeaf)) {
          const uint32_t f = v.x & 0xFFFFFFu;
          const float x = STAGE_X ? __uint_as_float(pred_key_bits(s_x[(f + 1u) * R + lane]))
                                  : (xok ? X[(int64_t)f * x_stride + row] : 0.0f);
          bool left;
          if (isnan(x))
            left = (v.x & kNodeDefLeft) != 0;
          else if (!(v.x & kNodeCat))
            left = x < __uint_as_float(v.y);
          else
            left = !pred_cat_right(x, cat_start[b + nd], cat_size[b + nd], cat_words);
          nd = left ? v.z : v.w;
          v = gn16[b + nd];
        }
#pragma unroll
        for (int j = 0; j < kPredILP; j++) {
          leafv[j] = 0.0f;
          leafid[j] = 0;
        }
        leafv[0] = __uint_as_float(v.y);
        leafid[0] = nd;
      }
      const uint32_t t_end = ch.n_nodes ? ch.end_tree : t0 + 1;
      // The group's output groups in one load (the array is padded by 8).
      PredGroup8 tg{};
      if (MODE != 2 && !one) tg = *reinterpret_cast<const PredGroup8*>(tree_group + t0);
#pragma unroll
      for (int j = 0; j < kPredILP; j++) {
        const uint32_t t = t0 + j;
        if (t < t_end) {
          if (in && leaf_idx)
            leaf_idx[(int64_t)t * leaf_stride + row] =
                ch.n_nodes ? m.corig[ch.node_begin + min(leafid[j], ch.n_nodes - 1u)]
                           : leafid[j];
          if (MODE == 2) {
            if (in) leaf_out[(int64_t)t * n + row] = leafv[j];
          } else if (one) {
            acc[0] += leafv[j];
          } else {
            // Trees of other windows match no accumulator (the difference
            // wraps for groups below g0).
            const uint32_t g = tg.v[j] - (uint32_t)m.g0;
#pragma unroll
            for (int k = 0; k < kPredMaxK; k++)
              if ((uint32_t)k == g) acc[k] += leafv[j];
          }
        }
      }
      if (!ch.n_nodes) break;  // a global-memory chunk holds one tree
    }
    if (next) BST_PRED_STORE(chn, s_mem + ((c + 1 - c0) & 1) * kPredChunk16);
    __syncthreads();
    ch = chn;
  }
#undef BST_PRED_L1
#undef BST_PRED_S1
#undef BST_PRED_LOAD
#undef BST_PRED_STORE
  if (MODE == 2 || !in) return;
  if (MODE == 0) pred_transform(acc, K, transform);
#pragma unroll
  for (int k = 0; k < kPredMaxK; k++)
    if (k < K) out[(int64_t)k * out_stride + row] = acc[k];
}

template <bool STAGE_X, int MODE, bool CAT>
__global__ __launch_bounds__(kPredMaxRows) void k_predict_lds(
    PredModel m, const float* __restrict__ X, int64_t n, int64_t x_stride, int transform,
    float* __restrict__ out, int64_t out_stride, uint32_t* __restrict__ leaf_idx,
    int64_t leaf_stride, float* __restrict__ leaf_out, int chunks_per_block) {
  const int c0 = MODE == 2 ? (int)blockIdx.y * chunks_per_block : 0;
  const int c1 = MODE == 2 ? min(c0 + chunks_per_block, m.n_chunks) : m.n_chunks;
  pred_block<STAGE_X, MODE, CAT>(m, X, n, x_stride, (int64_t)blockIdx.x * blockDim.x, c0, c1,
                                 transform, out, out_stride, leaf_idx, leaf_stride, leaf_out);
}

// Second phase of the two-phase launch: per row, base + every tree's leaf in
// forest order (the one-phase kernel's exact f32 sum), then the transform.
// leaf_in is [n_trees][n]; also the serving kernel's sum units.
__device__ __forceinline__ void pred_sum_row(const PredModel& m, const float* __restrict__ leaf_in,
                                             int64_t n, int64_t row, int transform,
                                             float* __restrict__ out, int64_t out_stride) {
  float acc[kPredMaxK];
#pragma unroll
  for (int k = 0; k < kPredMaxK; k++) acc[k] = k < m.Kw ? m.base[m.g0 + k] : 0.0f;
  if (m.K == 1) {
    for (int64_t t = 0; t < m.n_trees; t++) acc[0] += leaf_in[t * n + row];
  } else {
    for (int64_t t = 0; t < m.n_trees; t++) {
      const uint32_t g = m.tree_group[t] - (uint32_t)m.g0;
      const float v = leaf_in[t * n + row];
#pragma unroll
      for (int k = 0; k < kPredMaxK; k++)
        if ((uint32_t)k == g) acc[k] += v;
    }
  }
  pred_transform(acc, m.Kw, transform);
#pragma unroll
  for (int k = 0; k < kPredMaxK; k++)
    if (k < m.Kw) out[(int64_t)k * out_stride + row] = acc[k];
}

// ---- serving kernel --------------------------------------------------------
// Claims work units in order, lane 0 (small batches) before lane 1 (a
// compare-and-swap on the lane's next unit, only below its published count,
// so a block never holds a unit that is not yet described),
// finds the unit's batch among the ring's descriptors and runs it with the
// prediction body: a whole tile (raw margins), a leaf unit (a range of the
// forest's chunks, leaf values to the slot's leaf buffer) or a sum unit (the
// tile's rows summed over every tree in forest order, after every leaf unit
// of the batch is done: those were claimed before it, by running blocks that
// wait for nothing).  It makes the outputs visible to the host and counts the
// unit done; the block that finishes a batch's last unit publishes the
// batch's seq.  Every wave leaves the loop on the stop flag, once no block
// has claimed a unit for idle_ticks of the constant-rate clock, or after a
// bounded number of polls.
template <bool CAT>
__global__ __launch_bounds__(kPredMaxRows) void k_serve(PredModel m, ServeHostState* h,
                                                        ServeDevState* d,
                                                        const float* __restrict__ in_base,
                                                        float* __restrict__ out_base,
                                                        float* __restrict__ leaf_base,
                                                        uint64_t idle_ticks) {
  __shared__ uint64_t s_tile;
  __shared__ uint32_t s_slot;
  constexpr uint32_t kMaxPolls = 1u << 20;  // 2^20 polls of several us: a bound on any wait
  uint32_t cursor = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      uint64_t got = ~0ull;
      const uint64_t t0 = wall_clock64();
      // The poll bound counts polls since any block last claimed a unit.
      uint64_t seen_last = 0;
      uint32_t glane = 0;
      uint32_t hp = blockIdx.x;  // counts every poll (the bound's count restarts)
      for (uint32_t poll = 0; poll < kMaxPolls; poll++) {
        bool lost = false;  // a compare-and-swap lost a race: look again at once
        // A block reads the host's published counts on one poll in 8 and
        // folds them into a device copy that the other polls read: far fewer
        // reads across the host link while many blocks wait for work.
        const bool host_poll = (hp++ & 7u) == 0u;
        for (uint32_t ln = 0; ln < 2u && got == ~0ull && !lost; ln++) {
          uint64_t pub;
          if (host_poll) {
            pub = __hip_atomic_load(&h->published[ln], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_fetch_max(&d->seen_pub[ln], pub, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          } else {
            pub = __hip_atomic_load(&d->seen_pub[ln], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          }
          uint64_t cur =
              __hip_atomic_load(&d->next_unit[ln], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (cur < pub) {
            if (__hip_atomic_compare_exchange_strong(&d->next_unit[ln], &cur, cur + 1,
                                                     __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
              got = cur;
              glane = ln;
            } else {
              lost = true;
            }
          }
        }
        if (got != ~0ull) {
          __hip_atomic_store(&d->last_claim, wall_clock64(), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        if (lost) continue;
        // The stop flag is read with the published counts (host-link reads).
        if (host_poll && __hip_atomic_load(&h->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
          break;
        // Idle: leave only when no block has claimed a unit for idle_ticks
        // (a block that finds nothing while others work stays resident).
        const uint64_t now = wall_clock64();
        const uint64_t last =
            __hip_atomic_load(&d->last_claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (now - (last > t0 ? last : t0) > idle_ticks) break;
        if (last != seen_last) {
          seen_last = last;
          poll = 0;
        }
        __builtin_amdgcn_s_sleep(127);  // ~8k cycles between polls of host memory
      }
      uint32_t slot = kServeRing;
      if (got != ~0ull) {
        for (uint32_t k = 0; k < (uint32_t)kServeRing; k++) {
          const uint32_t i = (cursor + k) % kServeRing;
          const uint64_t rg = __hip_atomic_load(&h->desc[i].range, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          const uint64_t b = (rg >> 16) & kServeUnitMask, nt = rg & 0xFFFFu;
          if ((uint32_t)(rg >> 63) == glane && got >= b && got < b + nt) {
            slot = i;
            break;
          }
        }
        if (slot == kServeRing) {
          __hip_atomic_store(&h->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          got = ~0ull;
        } else {
          cursor = slot;
        }
      }
      s_tile = got;
      s_slot = slot;
    }
    __syncthreads();
    const uint64_t g = s_tile;
    const uint32_t slot = s_slot;
    if (g == ~0ull) return;
    const ServeDesc* dp = &h->desc[slot];
    const uint64_t tile_begin =
        (__hip_atomic_load(&dp->range, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) >> 16) &
        kServeUnitMask;
    const uint64_t in_off = __hip_atomic_load(&dp->in_off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint64_t out_off = __hip_atomic_load(&dp->out_off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint64_t target = __hip_atomic_load(&dp->done_target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint32_t rows = __hip_atomic_load(&dp->rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint32_t seq = __hip_atomic_load(&dp->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint32_t split = __hip_atomic_load(&dp->split, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint64_t u = g - tile_begin;  // the unit within its batch
    if (split == 0u) {
      const int64_t row0 = (int64_t)u * blockDim.x;
      pred_block<true, 0, CAT>(m, in_base + in_off, rows, rows, row0, 0, m.n_chunks, 0,
                               out_base + out_off, rows, nullptr, 0, nullptr);
      __threadfence_system();  // this thread's outputs reach the host before the count
    } else {
      const uint32_t tiles = __hip_atomic_load(&dp->tiles, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      const uint64_t leaf_off =
          __hip_atomic_load(&dp->leaf_off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      float* leaf = leaf_base + leaf_off;
      if (u < (uint64_t)tiles * split) {
        // Leaf unit: tile u / split, chunk range u % split of split.
        // The host picks split so that every range holds chunks.
        const int per = (m.n_chunks + (int)split - 1) / (int)split;
        const int c0 = (int)(u % split) * per, c1 = min(c0 + per, m.n_chunks);
        if (c0 < c1)  // block-uniform
          pred_block<true, 2, CAT>(m, in_base + in_off, rows, rows,
                                   (int64_t)(u / split) * blockDim.x, c0, c1, 0, nullptr, 0,
                                   nullptr, 0, leaf);
        __threadfence();  // leaf values visible on the device before the count
      } else {
        // Sum unit of tile u - tiles * split.
        const uint64_t leaf_target =
            __hip_atomic_load(&dp->leaf_target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (threadIdx.x == 0) {
          uint32_t poll = 0;
          while (__hip_atomic_load(&d->tiles_done[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
                 leaf_target) {
            if (++poll > kMaxPolls) {
              __hip_atomic_store(&h->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
              break;
            }
            __builtin_amdgcn_s_sleep(8);
          }
        }
        __syncthreads();
        __threadfence();  // every leaf unit's values are visible to this block
        const int64_t row = (int64_t)(u - (uint64_t)tiles * split) * blockDim.x + threadIdx.x;
        if (row < (int64_t)rows) pred_sum_row(m, leaf, rows, row, 0, out_base + out_off, rows);
        __threadfence_system();  // this thread's outputs reach the host before the count
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t old = __hip_atomic_fetch_add(&d->tiles_done[slot], 1ull, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_SYSTEM);
      if (old + 1 == target)
        __hip_atomic_store(&h->done_seq[slot], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// The serving kernel's LDS: the prediction body's, within the limit less
// room for its own claim words.
constexpr size_t kServeLdsReserve = 256;

int serve_tile_rows(int F, size_t lds_limit) {
  bool stage = false;
  int R = kPredMaxRows;
  (void)pred_lds_bytes(F, lds_limit - kServeLdsReserve, &stage, &R);
  return stage ? R : 0;
}

// Second phase of the two-phase launch (pred_sum_row per row).
__global__ __launch_bounds__(256) void k_predict_sum(PredModel m, const float* __restrict__ leaf_in,
                                                     int64_t n, int transform,
                                                     float* __restrict__ out, int64_t out_stride) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  pred_sum_row(m, leaf_in, n, row, transform, out, out_stride);
}

// Output transform of rows whose outputs were accumulated in windows
// (transform.rs:83-115): the same per-row f32 operations as pred_transform,
// over any number of outputs.
__global__ __launch_bounds__(256) void k_transform_rows(float* __restrict__ out, int64_t n,
                                                        int64_t stride, int K, int transform) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  float* o = out + row;
  if (transform == 1) {
    for (int k = 0; k < K; k++) o[k * stride] = sigmoid32(o[k * stride]);
  } else if (transform == 2) {
    float mx = -INFINITY;
    for (int k = 0; k < K; k++) mx = fmaxf(mx, o[k * stride]);
    float ssum = 0.0f;
    for (int k = 0; k < K; k++) {
      const float e = exp32(o[k * stride] - mx);
      o[k * stride] = e;
      ssum += e;
    }
    if (ssum > 0.0f)
      for (int k = 0; k < K; k++) o[k * stride] /= ssum;
  }
}

size_t pred_lds_bytes(int F, size_t lds_limit, bool* stage_rows, int* rows_per_block) {
  const size_t nodes = (size_t)2 * kPredChunk16 * sizeof(uint4);
  for (int r = kPredMaxRows; r >= 64; r >>= 1) {
    // The features' keys behind the row of zeros that leaves read.
    const size_t rows = (size_t)(std::max(F, 0) + 1) * r * sizeof(uint32_t);
    if (nodes + rows <= lds_limit) {
      if (stage_rows) *stage_rows = true;
      if (rows_per_block) *rows_per_block = r;
      return nodes + rows;
    }
  }
  if (stage_rows) *stage_rows = false;
  if (rows_per_block) *rows_per_block = kPredMaxRows;
  return nodes;
}

namespace {
// Raise a kernel's dynamic-LDS limit (on the current device) when a launch
// needs more than the default 64 KiB.
template <class Kern>
void pred_allow_lds(Kern k, size_t bytes) {
  if (bytes <= 65536) return;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <bool STAGE_X, int MODE, bool CAT>
void launch_pred_lds3(hipStream_t s, const PredModel& m, dim3 grid, int R, size_t lds,
                      const float* X, int64_t n, int64_t x_stride, int transform, float* out,
                      int64_t out_stride, uint32_t* leaf_idx, int64_t leaf_stride,
                      float* leaf_out, int chunks_per_block) {
  pred_allow_lds(k_predict_lds<STAGE_X, MODE, CAT>, lds);
  hipLaunchKernelGGL((k_predict_lds<STAGE_X, MODE, CAT>), grid, dim3(R), lds, s, m, X, n,
                     x_stride, transform, out, out_stride, leaf_idx, leaf_stride, leaf_out,
                     chunks_per_block);
}

template <int MODE>
void launch_pred_lds(hipStream_t s, const PredModel& m, unsigned gy, const float* X, int64_t n,
                     int64_t x_stride, int transform, float* out, int64_t out_stride,
                     uint32_t* leaf_idx, int64_t leaf_stride, float* leaf_out,
                     int chunks_per_block, size_t lds_limit) {
  bool stage = false;
  int R = kPredMaxRows;
  const size_t lds = pred_lds_bytes(m.n_features, lds_limit, &stage, &R);
  const dim3 grid((unsigned)((n + R - 1) / R), gy);
  const bool cat = m.has_cat != 0;
#define BST_PRED_GO(ST, CT)                                                                   \
  launch_pred_lds3<ST, MODE, CT>(s, m, grid, R, lds, X, n, x_stride, transform, out,          \
                                 out_stride, leaf_idx, leaf_stride, leaf_out, chunks_per_block)
  if (stage && cat) BST_PRED_GO(true, true);
  else if (stage) BST_PRED_GO(true, false);
  else if (cat) BST_PRED_GO(false, true);
  else BST_PRED_GO(false, false);
#undef BST_PRED_GO
}

// Blocks that keep every CU busy on the two-phase path.
int pred_target_blocks(int n_cu) { return 4 * std::max(n_cu, 1); }
}  // namespace

bool pred_two_phase(const PredModel& m, int64_t n, int n_cu, size_t lds_limit,
                    int64_t* scratch_floats) {
  int R = kPredMaxRows;
  (void)pred_lds_bytes(m.n_features, lds_limit, nullptr, &R);
  const int64_t tiles = (n + R - 1) / R;
  const bool two = m.n_chunks > 1 && tiles < (int64_t)std::max(n_cu, 1);
  if (scratch_floats) *scratch_floats = two ? m.n_trees * n : 0;
  return two;
}

// Every tree's leaf value per row into leaf_scratch [n_trees][n] (the first
// phase of a two-phase launch: (row tile, chunk range) blocks).
static void launch_leaf_values(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                               int64_t x_stride, uint32_t* leaf_idx, int64_t leaf_stride,
                               float* leaf_scratch, size_t lds_limit, int n_cu) {
  int R = kPredMaxRows;
  (void)pred_lds_bytes(m.n_features, lds_limit, nullptr, &R);
  const int64_t tiles = (n + R - 1) / R;
  const int want = (int)std::min<int64_t>(m.n_chunks,
                                          (pred_target_blocks(n_cu) + tiles - 1) / tiles);
  const int per = (m.n_chunks + want - 1) / want;
  const unsigned groups = (unsigned)((m.n_chunks + per - 1) / per);
  launch_pred_lds<2>(s, m, groups, X, n, x_stride, 0, nullptr, 0, leaf_idx, leaf_stride,
                     leaf_scratch, per, lds_limit);
}

void launch_predict_rows(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                         int64_t x_stride, int F, int transform, float* out, int64_t out_stride,
                         uint32_t* leaf_idx, int64_t leaf_stride, float* leaf_scratch,
                         size_t lds_limit, int n_cu) {
  (void)F;  // rows stage the model's features (F >= m.n_features is checked by the caller)
  if (n <= 0) return;
  const bool two = leaf_scratch && m.n_chunks > 1;
  const unsigned sum_grid = (unsigned)((n + 255) / 256);
  if (m.K <= kPredMaxK) {
    if (two) {
      launch_leaf_values(s, m, X, n, x_stride, leaf_idx, leaf_stride, leaf_scratch, lds_limit,
                         n_cu);
      hipLaunchKernelGGL(k_predict_sum, dim3(sum_grid), dim3(256), 0, s, m, leaf_scratch, n,
                         transform, out, out_stride);
    } else {
      launch_pred_lds<0>(s, m, 1, X, n, x_stride, transform, out, out_stride, leaf_idx,
                         leaf_stride, nullptr, m.n_chunks, lds_limit);
    }
    return;
  }
  // More outputs than one launch holds: windows of kPredMaxK outputs, each
  // adding only its own trees (in forest order), raw; then the transform
  // over all outputs of each row.  A two-phase launch takes the leaf values
  // once and sums them per window.
  if (two)
    launch_leaf_values(s, m, X, n, x_stride, leaf_idx, leaf_stride, leaf_scratch, lds_limit, n_cu);
  for (int g0 = 0; g0 < m.K; g0 += kPredMaxK) {
    PredModel w = m;
    w.g0 = g0;
    w.Kw = std::min(kPredMaxK, m.K - g0);
    float* wout = out + (int64_t)g0 * out_stride;
    if (two)
      hipLaunchKernelGGL(k_predict_sum, dim3(sum_grid), dim3(256), 0, s, w, leaf_scratch, n, 0,
                         wout, out_stride);
    else
      launch_pred_lds<0>(s, w, 1, X, n, x_stride, 0, wout, out_stride,
                         g0 == 0 ? leaf_idx : nullptr, leaf_stride, nullptr, m.n_chunks,
                         lds_limit);
  }
  if (transform)
    hipLaunchKernelGGL(k_transform_rows, dim3(sum_grid), dim3(256), 0, s, out, n, out_stride, m.K,
                       transform);
}

void launch_predict(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                    int64_t x_stride, int F, int transform, float* out, int64_t out_stride,
                    uint32_t* leaf_idx, float* leaf_scratch, size_t lds_limit, int n_cu) {
  launch_predict_rows(s, m, X, n, x_stride, F, transform, out, out_stride, leaf_idx, n,
                      leaf_scratch, lds_limit, n_cu);
}

void launch_serve(hipStream_t s, const PredModel& m, ServeHostState* h, ServeDevState* d,
                  const float* in_base, float* out_base, float* leaf_base, uint64_t idle_ticks,
                  size_t lds_limit, int n_cu) {
  bool stage = false;
  int R = kPredMaxRows;
  const size_t lds = pred_lds_bytes(m.n_features, lds_limit - kServeLdsReserve, &stage, &R);
  if (m.has_cat) {
    pred_allow_lds(k_serve<true>, lds);
    hipLaunchKernelGGL(k_serve<true>, dim3((unsigned)std::max(n_cu, 1)), dim3(R), lds, s, m, h, d,
                       in_base, out_base, leaf_base, idle_ticks);
  } else {
    pred_allow_lds(k_serve<false>, lds);
    hipLaunchKernelGGL(k_serve<false>, dim3((unsigned)std::max(n_cu, 1)), dim3(R), lds, s, m, h, d,
                       in_base, out_base, leaf_base, idle_ticks);
  }
}

void launch_predict_accum(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                          int64_t x_stride, int F, float* out, int64_t out_stride,
                          size_t lds_limit) {
  (void)F;
  if (n <= 0) return;
  for (int g0 = 0; g0 < m.K; g0 += kPredMaxK) {  // one window unless K > kPredMaxK
    PredModel w = m;
    w.g0 = g0;
    w.Kw = std::min(kPredMaxK, m.K - g0);
    launch_pred_lds<1>(s, w, 1, X, n, x_stride, 0, out + (int64_t)g0 * out_stride, out_stride,
                       nullptr, 0, nullptr, m.n_chunks, lds_limit);
  }
}

// Plain 16-byte streaming copy, the calibration point for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy16(hipStream_t s, void* dst, const void* src, int64_t bytes) {
  const int64_t n16 = bytes / 16;
  int64_t blocks = (n16 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4*)src,
                     (uint4*)dst, n16);
}

// l * log(q) + (1 - l) * log(1 - q).  For a label of exactly 0 or 1 one
// product is a signed zero, which leaves the other term's bits unchanged, so
// one log of the selected argument gives the same value (and a wave of such
// labels takes one log per lane instead of two).
__device__ __forceinline__ double logloss_term(double q, double l) {
  const bool hard = l == 0.0 || l == 1.0;
  const double lx = log(l == 0.0 ? 1.0 - q : q);
  return hard ? lx : l * lx + (1.0 - l) * log(1.0 - q);
}

// Transformed prediction of output k for row i (OutputTransform,
// transform.rs:23-115): sigmoid, max-subtracted softmax over the K outputs
// (f32, summed in output order), or the raw score.
__device__ __forceinline__ float metric_pred(int transform, const float* __restrict__ preds, int K,
                                             int64_t n, int64_t i, int k) {
  if (transform == BST_TRANSFORM_SIGMOID) return sigmoid32(preds[(int64_t)k * n + i]);
  if (transform == BST_TRANSFORM_SOFTMAX) {
    float mx = -INFINITY;
    for (int c = 0; c < K; c++) mx = fmaxf(mx, preds[(int64_t)c * n + i]);
    float ssum = 0.0f, ek = 0.0f;
    for (int c = 0; c < K; c++) {
      const float e = exp32(preds[(int64_t)c * n + i] - mx);
      ssum += e;
      if (c == k) ek = e;
    }
    return ssum > 0.0f ? ek / ssum : ek;
  }
  return preds[(int64_t)k * n + i];
}

// One row's (weighted loss, weight) contribution for metrics other than the
// batched RMSE / logloss path (metrics/regression.rs:17-250,
// classification.rs:13-360); k is the output for the per-output metrics.
__device__ double metric_term(int metric, int transform, const MetricParams& mp,
                              const float* __restrict__ preds, int K, int64_t n, int64_t i, int k,
                              float yi) {
  const double EPS = 1e-15;
  if (metric == BST_METRIC_RMSE) {
    const double d = (double)metric_pred(transform, preds, K, n, i, k) - (double)yi;
    return d * d;
  }
  if (metric == BST_METRIC_LOGLOSS) {
    const float p = metric_pred(transform, preds, K, n, i, 0);
    const double q = fmin(fmax((double)p, EPS), 1.0 - EPS);
    return -logloss_term(q, (double)yi);
  }
  if (metric == BST_METRIC_MLOGLOSS) {
    const int lab = (int)rint((double)yi);
    const float p = metric_pred(transform, preds, K, n, i, lab);
    const double q = fmin(fmax((double)p, EPS), 1.0 - EPS);
    return -log(q);
  }
  if (metric == BST_METRIC_MAE)
    return fabs((double)metric_pred(transform, preds, K, n, i, k) - (double)yi);
  if (metric == BST_METRIC_MAPE) {
    const double pp = metric_pred(transform, preds, K, n, i, 0), l = yi;
    return fabs(pp - l) / fmax(fabs(l), EPS);
  }
  if (metric == BST_METRIC_QUANTILE) {
    const double a = (double)mp.alphas[k];
    const double r = (double)yi - (double)metric_pred(transform, preds, K, n, i, k);
    return r >= 0.0 ? a * r : (1.0 - a) * (-r);
  }
  if (metric == BST_METRIC_HUBER) {
    const double r = fabs((double)metric_pred(transform, preds, K, n, i, 0) - (double)yi);
    const double d = mp.huber_delta;
    return r <= d ? 0.5 * r * r : d * (r - 0.5 * d);
  }
  if (metric == BST_METRIC_POISSON_DEVIANCE) {
    const double E9 = 1e-9;
    const double yy = yi, mu = fmax((double)metric_pred(transform, preds, K, n, i, 0), E9);
    return yy > E9 ? 2.0 * (yy * log(yy / mu) - (yy - mu)) : 2.0 * mu;
  }
  if (metric == BST_METRIC_ACCURACY || metric == BST_METRIC_MARGIN_ACCURACY) {
    const float t = metric == BST_METRIC_ACCURACY ? mp.accuracy_threshold : 0.0f;
    const float pc = metric_pred(transform, preds, K, n, i, 0) >= t ? 1.0f : 0.0f;
    return fabsf(pc - yi) < 0.5f ? 1.0 : 0.0;
  }
  // multiclass accuracy: the rounded score (one output) or the arg-max, the
  // last of equal maxima (Iterator::max_by; NaN compares Less)
  float pc;
  if (K == 1) {
    pc = roundf(metric_pred(transform, preds, K, n, i, 0));
  } else {
    int best = 0;
    float vb = metric_pred(transform, preds, K, n, i, 0);
    for (int c = 1; c < K; c++) {
      const float vc = metric_pred(transform, preds, K, n, i, c);
      if (isnan(vb) || isnan(vc) || !(vb > vc)) {
        best = c;
        vb = vc;
      }
    }
    pc = (float)best;
  }
  return fabsf(pc - yi) < 0.5f ? 1.0 : 0.0;
}

// Metric partial sums per block: (sum w*loss, sum w), f64.  RMSE, MAE and
// the quantile metric get one row of partials per output; the others read
// output 0 (or every output, for the multiclass ones) and use one row.
// Logloss and multiclass logloss clamp the transformed probability to
// [1e-15, 1 - 1e-15] (metrics/classification.rs:13-40, 332-360).
__global__ __launch_bounds__(256) void k_metric(int metric, int transform, MetricParams mp,
                                                const float* __restrict__ preds, int K,
                                                const float* __restrict__ y,
                                                const float* __restrict__ w, int64_t n,
                                                double* __restrict__ partial) {
  __shared__ double s_a[256 / WAVE], s_b[256 / WAVE];
  const int k = blockIdx.y;
  const double EPS = 1e-15;
  double a = 0.0, b = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i_start = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool fast = (metric == BST_METRIC_RMSE && transform == BST_TRANSFORM_IDENTITY) ||
                    (metric == BST_METRIC_LOGLOSS && transform != BST_TRANSFORM_SOFTMAX);
  if (fast) {
    // kRowBatch rows per thread per pass, loads issued together.
    for (; i_start < n; i_start += kRowBatch * stride) {
      float pv[kRowBatch], yv[kRowBatch], wv_[kRowBatch];
#pragma unroll
      for (int q = 0; q < kRowBatch; q++) {
        const int64_t i = i_start + q * stride;
        const bool in = i < n;
        pv[q] = in ? preds[(metric == BST_METRIC_RMSE ? (int64_t)k * n : 0) + i] : 0.0f;
        yv[q] = in ? y[i] : 0.0f;
        wv_[q] = in ? (w ? w[i] : 1.0f) : 0.0f;
      }
#pragma unroll
      for (int q = 0; q < kRowBatch; q++) {
        if (i_start + q * stride >= n) continue;
        if (metric == BST_METRIC_RMSE) {
          const double d = (double)pv[q] - (double)yv[q];
          a += (double)wv_[q] * d * d;
        } else {
          const float p = transform == BST_TRANSFORM_SIGMOID ? sigmoid32(pv[q]) : pv[q];
          const double qq = fmin(fmax((double)p, EPS), 1.0 - EPS);
          a += (double)wv_[q] * -logloss_term(qq, (double)yv[q]);
        }
        b += (double)wv_[q];
      }
    }
  }
  for (int64_t i = i_start; i < n; i += stride) {
    const float wi = w ? w[i] : 1.0f;
    a += (double)wi * metric_term(metric, transform, mp, preds, K, n, i, k, y[i]);
    b += (double)wi;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    a += __shfl_xor(a, m, WAVE);
    b += __shfl_xor(b, m, WAVE);
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    s_a[threadIdx.x / WAVE] = a;
    s_b[threadIdx.x / WAVE] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int i = 0; i < 256 / WAVE; i++) {
      ta += s_a[i];
      tb += s_b[i];
    }
    partial[((int64_t)k * gridDim.x + blockIdx.x) * 2] = ta;
    partial[((int64_t)k * gridDim.x + blockIdx.x) * 2 + 1] = tb;
  }
}

void launch_metric(hipStream_t s, int metric, int transform, const MetricParams& mp,
                   const float* preds, int K, const float* y, const float* w, int64_t n,
                   double* partial, int n_blocks) {
  const bool per_output = metric == BST_METRIC_RMSE || metric == BST_METRIC_MAE ||
                          metric == BST_METRIC_QUANTILE;
  const int ky = per_output ? K : 1;
  hipLaunchKernelGGL(k_metric, dim3(n_blocks, ky), dim3(256), 0, s, metric, transform, mp, preds,
                     K, y, w, n, partial);
}

}  // namespace bst
