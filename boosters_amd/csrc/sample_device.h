// sample_device.h — per-round row sampling (uniform, GOSS), per-tree column
// sampling and the raw-feature prediction update, on the device.
//
// Draws are "the k candidates with the smallest (key, index)", key = a 31-bit
// splitmix64 hash of (seed, stream, round, output, candidate index)
// (sample_keys.h); rows are returned in index order.  GOSS (row.rs:211-270)
// keeps every row whose |g h| is at least the (n - top_k)-th smallest
// magnitude, draws other_k of the others (all of them if fewer) and scales
// their (g, h) by (n - top_k) / other_k.  The k-th smallest values are found
// by an exact radix select (four 8-bit passes).
#pragma once
#include "engine.h"

namespace bst {

enum : int { kSampleNone = 0, kSampleUniform = 1, kSampleGoss = 2 };

struct SampleParams {
  uint32_t T;       // tree slots (slot t samples output out0 + t)
  int64_t n;        // rows
  int64_t stride;   // positions per slot in idx / ogh
  uint64_t seed;
  int32_t mode;     // kSampleUniform or kSampleGoss
  uint32_t k_mag;   // GOSS: rank (0-based) of the magnitude threshold, n - top_k
  uint32_t k_key;   // rank of the key cut: other_k - 1 (GOSS) or target - 1 (uniform)
  float mult;       // GOSS amplification (n - top_k) / other_k
};

// Per slot selection state (device memory).
struct SampleState {
  uint32_t prefix, mask, k, count, done;
  uint32_t thr_mag;   // GOSS: rows with magnitude bits >= thr_mag are kept
  uint32_t thr_key;   // candidates with key < thr_key are drawn ...
  uint32_t need;      // ... plus the first `need` (index order) with key == thr_key
  uint32_t hist[256];
};

constexpr int kSampleBlockRows = 4096;  // rows per compaction block

// Writes idx[t * stride + j] / ogh (amplified for GOSS) for the sampled rows
// of every slot, rows[t] = their count, absmax[2t..2t+1] over them.
// `round` is a device word (the boosting round).  ws: >= T SampleState,
// bcount: >= T * 4 * ceil(n / kSampleBlockRows) words.
void launch_sample_rows(hipStream_t s, const SampleParams& p, const uint32_t* round,
                        const float* g, const float* h, SampleState* ws, uint32_t* bcount,
                        uint32_t* idx, float2* ogh, uint32_t* rows, uint32_t* absmax);

// feats[t * E + j], j < k: the k sampled effective columns of slot t, sorted.
void launch_sample_columns(hipStream_t s, uint64_t seed, const uint32_t* round, int T, int E,
                           int k, uint32_t* feats);

// preds[tree_output[t]][i] += leaf of tree t for every row i, traversing the
// raw features X [F][n] (tree.predict_into, used when training rows were
// sampled).  Tree arrays [T][MNT] as the device grower writes them.
void launch_tree_update_raw(hipStream_t s, int T, uint32_t MNT, const uint32_t* split,
                            const float* thr, const uint32_t* left, const uint32_t* right,
                            const uint8_t* dleft, const uint8_t* isleaf, const uint8_t* stype,
                            const uint32_t* cats, const float* leaf, const float* X, int64_t n,
                            const uint32_t* tree_output, float* preds);

// *round += 1.
void launch_round_advance(hipStream_t s, uint32_t* round);

}  // namespace bst
