// limits.h — shape limits and counting rules of the trainer (host only, no HIP).
#pragma once
#include <stdint.h>

#include <cmath>

namespace bst {

// The class trees of a softmax round grow together (class batching) only
// while every row position t * n + i of the T trees fits a 32-bit index (the
// partition's row positions, histogram items and leaf segments are u32).
// At most kBatchTrees trees grow together (the device controllers' limit; a
// round of more class trees grows them in batches of that many).
// Returns how many trees of a round grow together: min(K, kBatchTrees) when
// batching is wanted and the positions fit, else 1 (one class tree at a
// time, positions < n).
constexpr int kBatchTrees = 16;
inline int class_batch_trees(int K, int64_t n_rows, bool want_batching) {
  if (!want_batching || K <= 1) return 1;
  const int T = K < kBatchTrees ? K : kBatchTrees;
  if ((uint64_t)T * (uint64_t)n_rows >= (1ull << 32)) return 1;
  return T;
}

// Columns a tree keeps under column sampling, ColSampler::sample_tree
// (training/sampling/column.rs:180-186): ceil(F * rate) with the product
// taken in f32, clamped to [1, F].  The f32 product matters: with F = 100 and
// rate = 0.07 it is exactly 7 (f64 would give 7.0000000298 and 8 columns).
inline uint32_t colsample_count(uint32_t n_features, float rate) {
  const float prod = (float)n_features * rate;
  uint64_t k = (uint64_t)std::ceil(prod);
  if (k < 1) k = 1;
  if (k > n_features) k = n_features;
  return (uint32_t)k;
}

}  // namespace bst
