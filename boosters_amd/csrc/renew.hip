// renew.hip — leaf renewal for the quantile objectives (absolute, pinball).
//
// The reference replaces each leaf's Newton value by the alpha-quantile of
// the residuals y - pred of the leaf's rows (trainer.rs:265-291,
// objectives/mod.rs:357-405) with utils.rs:17-126:
//   unweighted: the k-th smallest residual, k = ceil(alpha * n as f32) - 1,
//               clamped to [0, n - 1];
//   weighted:   the rows sorted by residual (a stable sort: ties keep the
//               leaf's order), total weight summed in f32 in the leaf's
//               order, and the first row in sorted order whose running f32
//               weight reaches total * alpha.
// Here one 64-bit key per row — the residual's order-preserving bits above
// the row's position in the partition — is sorted per leaf with a segmented
// radix sort, so equal residuals keep the partition order (the host-driven
// partition is stable, as the reference's is).  One thread per leaf then
// takes the k-th key, or walks the sorted keys with the two f32 sums in the
// reference's order (sequential: weighted quantile regression of very large
// leaves is the slow case).
#include "engine.h"

#include <hipcub/hipcub.hpp>

#include <math.h>
#include <stdexcept>

namespace bst {

namespace {

// Order-preserving unsigned image of an f32 (ascending floats map to
// ascending keys); -0.0 is taken as +0.0, which the reference's
// partial_cmp treats as equal.
__device__ __forceinline__ uint32_t f32_key(float v) {
  uint32_t u = __float_as_uint(v == 0.0f ? 0.0f : v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_f32(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __uint_as_float(u);
}

constexpr int kRenewThreads = 256;

// Keys of every position of every leaf: one block per leaf.
__global__ __launch_bounds__(kRenewThreads) void k_renew_keys(
    const RenewLeaf* __restrict__ leaves, const uint32_t* __restrict__ idx,
    const float* __restrict__ y, const float* __restrict__ preds, int64_t n,
    unsigned long long* __restrict__ keys) {
  const RenewLeaf L = leaves[blockIdx.x];
  const float* pk = preds + (int64_t)L.output * n;
  for (uint32_t j = threadIdx.x; j < L.count; j += blockDim.x) {
    const uint32_t pos = L.begin + j;
    const uint32_t row = idx[pos];
    const float r = y[row] - pk[row];
    keys[pos] = ((unsigned long long)f32_key(r) << 32) | pos;
  }
}

__global__ __launch_bounds__(kRenewThreads) void k_renew_pick(
    const RenewLeaf* __restrict__ leaves, int n_leaves, const uint32_t* __restrict__ idx,
    const float* __restrict__ w, const unsigned long long* __restrict__ sorted,
    float* __restrict__ out) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= n_leaves) return;
  const RenewLeaf L = leaves[li];
  if (L.count == 0) {
    out[li] = NAN;
    return;
  }
  const unsigned long long* sk = sorted + L.begin;
  if (!w) {
    int64_t k = (int64_t)ceilf(L.alpha * (float)L.count) - 1;
    k = k < 0 ? 0 : k;
    k = k > (int64_t)L.count - 1 ? (int64_t)L.count - 1 : k;
    out[li] = key_f32((uint32_t)(sk[k] >> 32));
    return;
  }
  if (L.count == 1) {
    out[li] = key_f32((uint32_t)(sk[0] >> 32));
    return;
  }
  float total = 0.0f;
  for (uint32_t j = 0; j < L.count; j++) total += w[idx[L.begin + j]];
  const float threshold = total * L.alpha;
  float cum = 0.0f;
  float v = key_f32((uint32_t)(sk[L.count - 1] >> 32));
  for (uint32_t j = 0; j < L.count; j++) {
    const uint32_t pos = (uint32_t)(sk[j] & 0xFFFFFFFFull);
    cum += w[idx[pos]];
    if (cum >= threshold) {
      v = key_f32((uint32_t)(sk[j] >> 32));
      break;
    }
  }
  out[li] = v;
}

__global__ void k_renew_offsets(const RenewLeaf* __restrict__ leaves, int n_leaves,
                                uint32_t* __restrict__ seg_begin, uint32_t* __restrict__ seg_end) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= n_leaves) return;
  seg_begin[li] = leaves[li].begin;
  seg_end[li] = leaves[li].begin + leaves[li].count;
}

}  // namespace

size_t renew_temp_bytes(int64_t n_pos, int n_leaves) {
  size_t bytes = 0;
  const hipError_t e = hipcub::DeviceSegmentedRadixSort::SortKeys(
      nullptr, bytes, (const unsigned long long*)nullptr, (unsigned long long*)nullptr,
      (int)n_pos, n_leaves, (const uint32_t*)nullptr, (const uint32_t*)nullptr, 0, 64, 0);
  if (e != hipSuccess) throw std::runtime_error("segmented sort size query failed");
  return bytes;
}

void renew_leaves(hipStream_t s, const RenewLeaf* leaves, int n_leaves, int64_t n_pos,
                  const uint32_t* idx, const float* y, const float* preds, const float* w,
                  int64_t n, unsigned long long* keys, unsigned long long* keys_sorted,
                  uint32_t* seg_begin, uint32_t* seg_end, void* temp, size_t temp_bytes,
                  float* out) {
  if (n_leaves <= 0) return;
  if (n_pos > (int64_t)0x7FFFFFFF) throw std::runtime_error("leaf renewal: too many positions");
  hipLaunchKernelGGL(k_renew_keys, dim3((unsigned)n_leaves), dim3(kRenewThreads), 0, s, leaves,
                     idx, y, preds, n, keys);
  const unsigned ob = (unsigned)((n_leaves + kRenewThreads - 1) / kRenewThreads);
  hipLaunchKernelGGL(k_renew_offsets, dim3(ob), dim3(kRenewThreads), 0, s, leaves, n_leaves,
                     seg_begin, seg_end);
  size_t bytes = temp_bytes;
  const hipError_t e = hipcub::DeviceSegmentedRadixSort::SortKeys(
      temp, bytes, (const unsigned long long*)keys, keys_sorted, (int)n_pos, n_leaves,
      (const uint32_t*)seg_begin, (const uint32_t*)seg_end, 0, 64, s);
  if (e != hipSuccess) throw std::runtime_error("segmented sort failed");
  hipLaunchKernelGGL(k_renew_pick, dim3(ob), dim3(kRenewThreads), 0, s, leaves, n_leaves, idx, w,
                     keys_sorted, out);
}

}  // namespace bst
