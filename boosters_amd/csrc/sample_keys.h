// sample_keys.h — counter-based sampling keys (host and device).
//
// Every random draw of training (row subsample, GOSS small-gradient rows,
// column subsets) is "the k candidates with the smallest (key, index)", with
// key = a 31-bit splitmix64 hash of (seed, stream, round, output, candidate).
// The reference's draws come from rand 0.8.5 streams (sampling/row.rs:106,
// 185,256) or OS entropy (sampling/column.rs:137-142) that nothing pins; these
// keys give the same uniform k-subset distributions and are reproducible by
// any number of threads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bst {

constexpr uint32_t kStreamRows = 1, kStreamGoss = 2, kStreamCols = 3, kStreamLevel = 4;

__host__ __device__ inline uint64_t splitmix64_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t sample_key_base(uint64_t seed, uint32_t stream,
                                                    uint32_t round, uint32_t output) {
  return splitmix64_mix(seed ^ splitmix64_mix(((uint64_t)round << 32) |
                                              ((uint64_t)output << 8) | stream));
}

// Keys of a tree's level subset at `depth` (column sampling by level).
__host__ __device__ inline uint64_t level_key_base(uint64_t seed, uint32_t round, uint32_t output,
                                                   uint32_t depth) {
  return sample_key_base(seed, kStreamLevel, round, output | (depth << 12));
}

__host__ __device__ inline uint32_t sample_key(uint64_t base, uint32_t i) {
  return (uint32_t)(splitmix64_mix(base ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull)) >> 33);
}

}  // namespace bst
