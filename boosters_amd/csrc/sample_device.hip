// sample_device.hip — row sampling (uniform, GOSS), column sampling and the
// raw-feature prediction update on the device (see sample_device.h).
#include <math.h>

#include "sample_device.h"
#include "sample_keys.h"

namespace bst {

namespace {

constexpr int kWave = 64;
constexpr int kSelThreads = 256;
constexpr int kCompThreads = 1024;
constexpr int kCompPasses = kSampleBlockRows / kCompThreads;
enum : int { kValMag = 0, kValKeySmall = 1, kValKeyAll = 2 };

__device__ __forceinline__ uint32_t mag_bits(const float* g, const float* h, int64_t o) {
  return __float_as_uint(fabsf(g[o] * h[o]));  // |g h| >= 0: bit order = value order
}

// Exclusive scan over a block of kCompThreads threads; *total gets the sum.
__device__ uint32_t comp_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = __shfl_up(x, d, kWave);
    if (lane >= d) x += o;
  }
  if (lane == kWave - 1) s_w[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int i = 0; i < kCompThreads / kWave; i++) {
      const uint32_t t = s_w[i];
      s_w[i] = s;
      s += t;
    }
    s_w[kCompThreads / kWave] = s;
  }
  __syncthreads();
  const uint32_t r = x - v + s_w[wv];
  *total = s_w[kCompThreads / kWave];
  __syncthreads();
  return r;
}

// The candidate value of row i in slot t for a selection pass.
template <int WHAT>
__device__ __forceinline__ bool cand_value(const SampleParams& P, const float* g, const float* h,
                                           uint32_t thr_mag, uint64_t base, uint32_t t,
                                           uint32_t i, uint32_t* v) {
  if (WHAT == kValKeyAll) {
    *v = sample_key(base, i);
    return true;
  }
  const uint32_t mb = mag_bits(g, h, (int64_t)t * P.n + i);
  if (WHAT == kValMag) {
    *v = mb;
    return true;
  }
  if (mb >= thr_mag) return false;  // a large-gradient row, always kept
  *v = sample_key(base, i);
  return true;
}

}  // namespace

__global__ void k_sel_init(SampleState* st, int which, uint32_t k) {
  SampleState& S = st[blockIdx.x];
  S.hist[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    S.prefix = 0u;
    S.mask = 0u;
    S.k = k;
    S.count = 0u;
    S.done = 0u;
    if (which == 0) S.thr_mag = 0xFFFFFFFFu;
  }
}

// Digit histogram of the candidates matching the prefix so far.
template <int WHAT>
__global__ __launch_bounds__(kSelThreads) void k_sel_hist(SampleParams P, const float* __restrict__ g,
                                                          const float* __restrict__ h,
                                                          SampleState* __restrict__ st, int shift,
                                                          const uint32_t* __restrict__ round) {
  __shared__ uint32_t lh[256];
  const uint32_t t = blockIdx.y;
  const uint32_t done = st[t].done, prefix = st[t].prefix, mask = st[t].mask;
  const uint32_t thr_mag = st[t].thr_mag;
  if (done) return;  // uniform over the block
  lh[threadIdx.x] = 0u;
  __syncthreads();
  const uint64_t base = sample_key_base(P.seed, WHAT == kValKeyAll ? kStreamRows : kStreamGoss,
                                        *round, t);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v;
    if (cand_value<WHAT>(P, g, h, thr_mag, base, t, (uint32_t)i, &v) && (v & mask) == prefix)
      atomicAdd(&lh[(v >> shift) & 255u], 1u);
  }
  __syncthreads();
  const uint32_t c = lh[threadIdx.x];
  if (c) atomicAdd(&st[t].hist[threadIdx.x], c);
}

// Pick the digit holding rank k; on the last digit, publish the threshold.
__global__ __launch_bounds__(kSelThreads) void k_sel_pick(SampleState* __restrict__ st, int shift,
                                                          int first, int which) {
  __shared__ uint32_t s_w[kSelThreads / kWave + 1];
  __shared__ uint32_t s_d, s_before, s_total;
  SampleState& S = st[blockIdx.x];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const uint32_t hv = S.hist[threadIdx.x];
  uint32_t x = hv;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = __shfl_up(x, d, kWave);
    if (lane >= d) x += o;
  }
  if (lane == kWave - 1) s_w[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int i = 0; i < kSelThreads / kWave; i++) {
      const uint32_t tt = s_w[i];
      s_w[i] = s;
      s += tt;
    }
    s_total = s;
    s_d = 0xFFFFFFFFu;
  }
  __syncthreads();
  const uint32_t incl = x + s_w[wv], excl = incl - hv;
  const uint32_t k = S.k;
  const bool done = S.done != 0u;
  if (!done && hv && excl <= k && k < incl) {
    s_d = threadIdx.x;
    s_before = excl;
  }
  __syncthreads();
  S.hist[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    if (first) S.count = s_total;
    bool all = done;
    if (!all && first && k >= s_total) all = true;  // fewer candidates than k + 1
    if (all) {
      S.done = 1u;
    } else if (s_d != 0xFFFFFFFFu) {
      S.k = k - s_before;
      S.prefix |= s_d << shift;
      S.mask |= 0xFFu << shift;
    }
    if (shift == 0) {
      if (which == 0) {
        S.thr_mag = S.prefix;
      } else if (S.done) {
        S.thr_key = 0x80000000u;  // above every 31-bit key: take every candidate
        S.need = 0u;
      } else {
        S.thr_key = S.prefix;
        S.need = S.k + 1u;  // S.k = candidates equal to the cut ranked before it
      }
    }
  }
}

namespace {

// Row flags for the compaction: sure = drawn by the threshold alone, tie =
// candidate whose key equals the cut (drawn by index order up to `need`).
template <int MODE>
__device__ __forceinline__ void row_flags(const SampleParams& P, const float* g, const float* h,
                                          const SampleState& S, uint64_t base, uint32_t t,
                                          uint32_t i, bool* sure, bool* tie, bool* amp) {
  *amp = false;
  if (MODE == kSampleGoss) {
    const uint32_t mb = mag_bits(g, h, (int64_t)t * P.n + i);
    if (mb >= S.thr_mag) {
      *sure = true;
      *tie = false;
      return;
    }
    *amp = true;
  }
  const uint32_t key = sample_key(base, i);
  *sure = key < S.thr_key;
  *tie = key == S.thr_key;
}

}  // namespace

template <int MODE>
__global__ __launch_bounds__(kCompThreads) void k_samp_count(
    SampleParams P, const float* __restrict__ g, const float* __restrict__ h,
    const SampleState* __restrict__ st, const uint32_t* __restrict__ round,
    uint32_t* __restrict__ bcount) {
  __shared__ uint32_t s_w[kCompThreads / kWave + 1];
  const uint32_t t = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const SampleState S = st[t];
  const uint64_t base = sample_key_base(P.seed, MODE == kSampleGoss ? kStreamGoss : kStreamRows,
                                        *round, t);
  uint32_t cs = 0, ct = 0;
  for (int pass = 0; pass < kCompPasses; pass++) {
    const int64_t i = (int64_t)b * kSampleBlockRows + pass * kCompThreads + threadIdx.x;
    if (i < P.n) {
      bool sure, tie, amp;
      row_flags<MODE>(P, g, h, S, base, t, (uint32_t)i, &sure, &tie, &amp);
      cs += sure ? 1u : 0u;
      ct += tie ? 1u : 0u;
    }
  }
  uint32_t ts, tt;
  (void)comp_scan(cs, s_w, &ts);
  (void)comp_scan(ct, s_w, &tt);
  if (threadIdx.x == 0) {
    bcount[4 * ((int64_t)t * nb + b) + 0] = ts;
    bcount[4 * ((int64_t)t * nb + b) + 1] = tt;
  }
}

// Per slot: tie offsets, drawn ties and output offsets of every block.
__global__ __launch_bounds__(kCompThreads) void k_samp_scan(const SampleState* __restrict__ st,
                                                            uint32_t* __restrict__ bcount, int nb,
                                                            uint32_t* __restrict__ rows,
                                                            uint32_t* __restrict__ absmax) {
  __shared__ uint32_t s_w[kCompThreads / kWave + 1];
  const uint32_t t = blockIdx.x;
  const uint32_t need = st[t].need;
  uint32_t* bc = bcount + 4 * (int64_t)t * nb;
  uint32_t tie_carry = 0, out_carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kCompThreads) {
    const int b = b0 + threadIdx.x;
    const bool in = b < nb;
    const uint32_t sure = in ? bc[4 * b] : 0u, tie = in ? bc[4 * b + 1] : 0u;
    uint32_t ttot;
    const uint32_t toff = comp_scan(tie, s_w, &ttot) + tie_carry;
    const uint32_t drawn = toff >= need ? 0u : min(tie, need - toff);
    uint32_t otot;
    const uint32_t ooff = comp_scan(sure + drawn, s_w, &otot) + out_carry;
    if (in) {
      bc[4 * b + 2] = toff;
      bc[4 * b + 3] = ooff;
    }
    tie_carry += ttot;
    out_carry += otot;
  }
  if (threadIdx.x == 0) {
    rows[t] = out_carry;
    absmax[2 * t] = 0u;
    absmax[2 * t + 1] = 0u;
  }
}

template <int MODE>
__global__ __launch_bounds__(kCompThreads) void k_samp_write(
    SampleParams P, const float* __restrict__ g, const float* __restrict__ h,
    const SampleState* __restrict__ st, const uint32_t* __restrict__ round,
    const uint32_t* __restrict__ bcount, uint32_t* __restrict__ idx, float2* __restrict__ ogh,
    uint32_t* __restrict__ absmax) {
  __shared__ uint32_t s_w[kCompThreads / kWave + 1];
  __shared__ uint32_t s_max[2][kCompThreads / kWave];
  const uint32_t t = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const SampleState S = st[t];
  const uint64_t base = sample_key_base(P.seed, MODE == kSampleGoss ? kStreamGoss : kStreamRows,
                                        *round, t);
  const uint32_t* bc = bcount + 4 * ((int64_t)t * nb + b);
  uint32_t tie_carry = bc[2], out_carry = bc[3];
  uint32_t mg = 0u, mh = 0u;
  for (int pass = 0; pass < kCompPasses; pass++) {
    const int64_t i = (int64_t)b * kSampleBlockRows + pass * kCompThreads + threadIdx.x;
    bool sure = false, tie = false, amp = false;
    if (i < P.n) row_flags<MODE>(P, g, h, S, base, t, (uint32_t)i, &sure, &tie, &amp);
    uint32_t ttot;
    const uint32_t trank = comp_scan(tie ? 1u : 0u, s_w, &ttot) + tie_carry;
    const bool sel = sure || (tie && trank < S.need);
    uint32_t stot;
    const uint32_t pos = comp_scan(sel ? 1u : 0u, s_w, &stot) + out_carry;
    if (sel) {
      const int64_t o = (int64_t)t * P.n + i;
      float gv = g[o], hv = h[o];
      if (MODE == kSampleGoss && amp) {
        gv *= P.mult;
        hv *= P.mult;
      }
      idx[(int64_t)t * P.stride + pos] = (uint32_t)i;
      ogh[(int64_t)t * P.stride + pos] = make_float2(gv, hv);
      mg = max(mg, __float_as_uint(fabsf(gv)));
      mh = max(mh, __float_as_uint(fabsf(hv)));
    }
    tie_carry += ttot;
    out_carry += stot;
  }
  for (int m = kWave / 2; m >= 1; m >>= 1) {
    mg = max(mg, (uint32_t)__shfl_xor(mg, m, kWave));
    mh = max(mh, (uint32_t)__shfl_xor(mh, m, kWave));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_max[0][threadIdx.x / kWave] = mg;
    s_max[1][threadIdx.x / kWave] = mh;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t v = 0u;
    for (int i = 0; i < kCompThreads / kWave; i++) v = max(v, s_max[threadIdx.x][i]);
    atomicMax(&absmax[2 * t + threadIdx.x], v);
  }
}

void launch_sample_rows(hipStream_t s, const SampleParams& p, const uint32_t* round,
                        const float* g, const float* h, SampleState* ws, uint32_t* bcount,
                        uint32_t* idx, float2* ogh, uint32_t* rows, uint32_t* absmax) {
  if (p.T == 0 || p.n <= 0) return;
  int64_t hb = (p.n + kSelThreads - 1) / kSelThreads;
  if (hb > 1024) hb = 1024;
  const dim3 hgrid((unsigned)hb, p.T);
  auto select = [&](int which, uint32_t k) {
    hipLaunchKernelGGL(k_sel_init, dim3(p.T), dim3(256), 0, s, ws, which, k);
    for (int shift = 24; shift >= 0; shift -= 8) {
      if (which == 0)
        hipLaunchKernelGGL((k_sel_hist<kValMag>), hgrid, dim3(kSelThreads), 0, s, p, g, h, ws,
                           shift, round);
      else if (p.mode == kSampleGoss)
        hipLaunchKernelGGL((k_sel_hist<kValKeySmall>), hgrid, dim3(kSelThreads), 0, s, p, g, h,
                           ws, shift, round);
      else
        hipLaunchKernelGGL((k_sel_hist<kValKeyAll>), hgrid, dim3(kSelThreads), 0, s, p, g, h, ws,
                           shift, round);
      hipLaunchKernelGGL(k_sel_pick, dim3(p.T), dim3(kSelThreads), 0, s, ws, shift,
                         shift == 24 ? 1 : 0, which);
    }
  };
  if (p.mode == kSampleGoss) select(0, p.k_mag);
  select(1, p.k_key);
  const int nb = (int)((p.n + kSampleBlockRows - 1) / kSampleBlockRows);
  const dim3 cgrid((unsigned)nb, p.T);
  if (p.mode == kSampleGoss) {
    hipLaunchKernelGGL((k_samp_count<kSampleGoss>), cgrid, dim3(kCompThreads), 0, s, p, g, h, ws,
                       round, bcount);
  } else {
    hipLaunchKernelGGL((k_samp_count<kSampleUniform>), cgrid, dim3(kCompThreads), 0, s, p, g, h,
                       ws, round, bcount);
  }
  hipLaunchKernelGGL(k_samp_scan, dim3(p.T), dim3(kCompThreads), 0, s, ws, bcount, nb, rows,
                     absmax);
  if (p.mode == kSampleGoss) {
    hipLaunchKernelGGL((k_samp_write<kSampleGoss>), cgrid, dim3(kCompThreads), 0, s, p, g, h, ws,
                       round, bcount, idx, ogh, absmax);
  } else {
    hipLaunchKernelGGL((k_samp_write<kSampleUniform>), cgrid, dim3(kCompThreads), 0, s, p, g, h,
                       ws, round, bcount, idx, ogh, absmax);
  }
}

// One block per slot: the k effective columns with the smallest keys, in
// column order (bitonic sort of (key, column) in LDS, then a scan).
constexpr int kColThreads = 1024;
constexpr int kColMax = 1024;

__global__ __launch_bounds__(kColThreads) void k_sample_columns(uint64_t seed,
                                                                const uint32_t* __restrict__ round,
                                                                int E, int k,
                                                                uint32_t* __restrict__ feats) {
  __shared__ unsigned long long s_key[kColMax];
  __shared__ uint8_t s_sel[kColMax];
  __shared__ uint32_t s_w[kColThreads / kWave + 1];
  const uint32_t t = blockIdx.x;
  int P = 1;
  while (P < E) P <<= 1;
  const uint64_t base = sample_key_base(seed, kStreamCols, *round, t);
  for (int f = threadIdx.x; f < P; f += blockDim.x) {
    s_key[f] = f < E ? (((unsigned long long)sample_key(base, (uint32_t)f) << 32) | (uint32_t)f)
                     : ~0ull;
    if (f < kColMax) s_sel[f] = 0;
  }
  __syncthreads();
  for (int kk = 2; kk <= P; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long a = s_key[i], b = s_key[ixj];
          const bool up = (i & kk) == 0;
          if ((a > b) == up) {
            s_key[i] = b;
            s_key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int j = threadIdx.x; j < k; j += blockDim.x) s_sel[(uint32_t)(s_key[j] & 0xFFFFFFFFull)] = 1;
  __syncthreads();
  uint32_t carry = 0;
  for (int f0 = 0; f0 < E; f0 += kColThreads) {
    const int f = f0 + threadIdx.x;
    const uint32_t sel = f < E ? s_sel[f] : 0u;
    uint32_t tot;
    const uint32_t r = comp_scan(sel, s_w, &tot) + carry;
    if (sel) feats[(int64_t)t * E + r] = (uint32_t)f;
    carry += tot;
  }
}

void launch_sample_columns(hipStream_t s, uint64_t seed, const uint32_t* round, int T, int E,
                           int k, uint32_t* feats) {
  if (T <= 0 || E <= 0) return;
  hipLaunchKernelGGL(k_sample_columns, dim3((unsigned)T), dim3(kColThreads), 0, s, seed, round, E,
                     k, feats);
}

// tree.predict_into on the raw features (inference rules, tree_view.rs:119-160:
// NaN -> default child; numeric x < thr -> left; categorical: (u32)x in the
// stored set -> right).
__global__ __launch_bounds__(256) void k_tree_update_raw(
    uint32_t MNT, const uint32_t* __restrict__ split, const float* __restrict__ thr,
    const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
    const uint8_t* __restrict__ dleft, const uint8_t* __restrict__ isleaf,
    const uint8_t* __restrict__ stype, const uint32_t* __restrict__ cats,
    const float* __restrict__ leaf, const float* __restrict__ X, int64_t n,
    const uint32_t* __restrict__ tree_output, float* __restrict__ preds) {
  const uint32_t t = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = (int64_t)t * MNT;
  uint32_t node = 0;
  while (!isleaf[b + node]) {
    const float x = X[(int64_t)split[b + node] * n + i];
    bool go_left;
    if (isnan(x)) {
      go_left = dleft[b + node] != 0;
    } else if (!stype[b + node]) {
      go_left = x < thr[b + node];
    } else {
      const uint32_t c = x <= 0.0f ? 0u : (x >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)x);
      const bool in_set = c < 32u * kMaxCatWords &&
                          ((cats[(b + node) * kMaxCatWords + (c >> 5)] >> (c & 31)) & 1u);
      go_left = !in_set;
    }
    node = go_left ? left[b + node] : right[b + node];
  }
  preds[(int64_t)tree_output[t] * n + i] += leaf[b + node];
}

void launch_tree_update_raw(hipStream_t s, int T, uint32_t MNT, const uint32_t* split,
                            const float* thr, const uint32_t* left, const uint32_t* right,
                            const uint8_t* dleft, const uint8_t* isleaf, const uint8_t* stype,
                            const uint32_t* cats, const float* leaf, const float* X, int64_t n,
                            const uint32_t* tree_output, float* preds) {
  if (T <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_tree_update_raw, dim3((unsigned)((n + 255) / 256), (unsigned)T), dim3(256),
                     0, s, MNT, split, thr, left, right, dleft, isleaf, stype, cats, leaf, X, n,
                     tree_output, preds);
}

__global__ void k_round_advance(uint32_t* round) { *round += 1u; }

void launch_round_advance(hipStream_t s, uint32_t* round) {
  hipLaunchKernelGGL(k_round_advance, dim3(1), dim3(1), 0, s, round);
}

}  // namespace bst
