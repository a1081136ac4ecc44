// bst_oracle.cpp — TEST INFRASTRUCTURE ONLY (parity checker + CPU baseline).
//
// A CPU restatement of egordm/boosters' histogram GBDT path.  Paths in the
// citations are relative to crates/boosters/src/.  Build flags matter for
// parity: -O2 -ffp-contract=off (no FMA contraction, as rustc), no -ffast-math,
// glibc expf/logf (what Rust's f32::exp lowers to on x86_64-linux-gnu).
//
// Deliberate differences from the reference, each numerically neutral or
// documented:
//   * No LRU histogram pool (pool.rs): every parent histogram is kept, so the
//     larger child is always parent - smaller.  The reference rebuilds some
//     children from rows when the 8-slot pool evicted the parent; that only
//     changes f64 summation order.
//   * Sparse-storage features are binned and histogrammed like dense ones
//     (the reference skips sparse features in its gathered histogram path,
//     ops.rs:158-160); no in-scope workload has a feature with <=10% nonzeros.
//   * Row subsampling, GOSS small-gradient rows and column subsets are drawn
//     as "the k candidates with the smallest counter-based key" (smallest_keys
//     below) instead of the reference's rand 0.8.5 streams (row.rs:106,185,
//     256; column seeded from OS entropy, column.rs:137-142): the same uniform
//     k-subset distributions, reproducible by parallel hardware.  GOSS keeps
//     the reference's threshold, warm-up and amplification.  Parity with the
//     reference's draws is unpinned (no test or fixture fixes them).
//   * LambdaRank has no reference counterpart; see lambdarank_gradients.
#include "bst_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <queue>
#include <string>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
int g_threads = 0;
// 0: glibc expf (what Rust's f32::exp calls on x86_64-linux-gnu);
// 1: exp evaluated in f64 and rounded once to f32 (the engine's device exp).
// glibc expf is off by one ulp from the correctly rounded value on ~7.6e-5 of
// inputs, so mode 1 is used where tests need the engine's exact arithmetic.
int g_exp_mode = 0;
inline float exp32(float x) { return g_exp_mode ? (float)std::exp((double)x) : expf(x); }

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Binning (data/binned/)
// ---------------------------------------------------------------------------
struct Mapper {
  bool categorical = false;
  std::vector<double> bounds;  // numeric upper bounds (incl. f64::MAX sentinels)
  std::vector<int32_t> cats;   // categorical: sorted categories
  uint32_t n_bins = 0;
  int missing = 0;  // 0 None, 2 NaN (MissingType, bin_mapper.rs:20-28)
  uint32_t default_bin = 0;
  bool trivial = false;
};

// BinMapper::value_to_bin + search_bin, bin_mapper.rs:210-258.
uint32_t value_to_bin(const Mapper& m, double v) {
  if (std::isnan(v)) return m.missing == 2 ? m.n_bins - 1 : m.default_bin;
  if (m.categorical) {
    // Rust `f64 as i32`: truncating, saturating; unknown category -> bin 0.
    int32_t c;
    if (v >= 2147483647.0)
      c = 2147483647;
    else if (v <= -2147483648.0)
      c = (int32_t)-2147483648LL;
    else
      c = (int32_t)v;
    auto it = std::lower_bound(m.cats.begin(), m.cats.end(), c);
    if (it != m.cats.end() && *it == c) return (uint32_t)(it - m.cats.begin());
    return 0;
  }
  size_t lo = 0, hi = m.bounds.empty() ? 0 : m.bounds.size() - 1;
  if (m.missing == 2 && hi > 0) hi -= 1;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (v <= m.bounds[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return (uint32_t)lo;
}

// bin_numeric, data/binned/dataset.rs:683-750.
Mapper bin_numeric(const std::vector<float>& data, uint32_t max_bins, bool has_nan) {
  float min_v = std::numeric_limits<float>::max();
  float max_v = -std::numeric_limits<float>::max();
  std::vector<float> vals;
  vals.reserve(data.size());
  for (float x : data)
    if (std::isfinite(x)) {
      min_v = std::min(min_v, x);
      max_v = std::max(max_v, x);
      vals.push_back(x);
    }
  Mapper m;
  if (vals.empty() || min_v >= max_v) {
    m.bounds = {std::numeric_limits<double>::max()};
    m.n_bins = 1;
    return m;
  }
  uint32_t nb = std::min<uint64_t>(max_bins, vals.size());
  std::stable_sort(vals.begin(), vals.end(), [](float a, float b) { return a < b; });
  vals.erase(std::unique(vals.begin(), vals.end(), [](float a, float b) { return a == b; }),
             vals.end());
  size_t nu = vals.size();
  if (nu <= nb) {
    for (size_t i = 0; i + 1 < nu; i++) m.bounds.push_back(((double)vals[i] + (double)vals[i + 1]) / 2.0);
  } else {
    for (uint32_t i = 1; i < nb; i++) {
      double q = (double)i / (double)nb;
      size_t idx = std::min<size_t>((size_t)std::floor(q * (double)(nu - 1)), nu - 2);
      double b = ((double)vals[idx] + (double)vals[idx + 1]) / 2.0;
      if (m.bounds.empty() || b > m.bounds.back()) m.bounds.push_back(b);
    }
  }
  m.bounds.push_back(std::numeric_limits<double>::max());
  if (has_nan) m.bounds.push_back(std::numeric_limits<double>::max());
  m.missing = has_nan ? 2 : 0;
  m.n_bins = (uint32_t)m.bounds.size();
  m.default_bin = has_nan ? m.n_bins - 1 : 0;
  return m;
}

// One step of the SplitMix64 generator (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014; the 64-bit
// finaliser constants of Vigna's splitmix64.c): advance the state by the
// golden gamma and return the mixed state.  dataset.rs:598-629 draws with it.
uint64_t splitmix64_next(uint64_t& state) {
  state += 0x9E3779B97F4A7C15ULL;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Deterministic sample of row ids for edge computation, dataset.rs:598-629:
// state 0xD1B54A32D192ED03 ^ n, draws z % n, duplicates rejected, sorted.
std::vector<uint32_t> binning_sample(int64_t n, uint64_t cnt) {
  std::vector<uint32_t> idx;
  if ((uint64_t)n <= cnt) return idx;
  uint64_t state = 0xD1B54A32D192ED03ULL ^ (uint64_t)n;
  std::unordered_set<uint32_t> seen;
  seen.reserve(cnt * 2);
  idx.reserve(cnt);
  while (idx.size() < cnt) {
    const uint64_t z = splitmix64_next(state);
    uint32_t r = (uint32_t)(z % (uint64_t)n);
    if (seen.insert(r).second) idx.push_back(r);
  }
  std::sort(idx.begin(), idx.end());
  return idx;
}

}  // namespace

struct ora_binned {
  int64_t n = 0;
  int32_t F = 0;
  std::vector<Mapper> mappers;         // per original feature
  std::vector<uint32_t> eff_to_orig;   // effective column -> original feature
  std::vector<uint32_t> eff_nbins, eff_offset;
  std::vector<uint8_t> eff_cat, eff_missing;
  std::vector<uint32_t> eff_default;        // missing bin of each effective column
  std::vector<std::vector<uint16_t>> bins;  // per effective column [n]
  uint32_t total_bins = 0;
  // Exclusive feature bundles: effective columns 0 .. bundles.size() - 1.
  struct Bundle {
    std::vector<uint32_t> features, offsets;  // original features, code of their bin 0
    uint32_t total_bins = 0;
  };
  std::vector<Bundle> bundles;
};

namespace {

// FeatureStats::update / into_analysis, feature_analysis.rs:313-432 (the
// fields that decide trivial / categorical / has-NaN).
struct Analysis {
  bool trivial, numeric;
  int64_t n_nan;
  int64_t n_nonzero;  // FeatureStats::non_zero_count (not NaN, != 0)
};

Analysis analyze(const float* col, int64_t n, bool user_cat, const bst_binning_config& cfg) {
  size_t max_unique = std::max<size_t>(cfg.max_categorical_cardinality, cfg.max_bins) + 10;
  std::vector<float> uniq;
  bool exceeded = false, all_int = true;
  int64_t valid = 0, nan = 0, nonzero = 0;
  for (int64_t i = 0; i < n; i++) {
    float v = col[i];
    if (std::isnan(v)) {
      nan++;
      continue;
    }
    valid++;
    if (v != 0.0f) nonzero++;
    if (all_int && (v - std::trunc(v)) != 0.0f) all_int = false;
    if (!exceeded && std::find(uniq.begin(), uniq.end(), v) == uniq.end()) {
      if (uniq.size() >= max_unique)
        exceeded = true;
      else
        uniq.push_back(v);
    }
  }
  size_t nu = uniq.size();
  Analysis a;
  a.trivial = valid == 0 || (nu == 1 && uniq[0] == 0.0f) || nu == 0;
  bool binary = nu == 2;
  bool auto_cat = !user_cat && all_int && nu <= cfg.max_categorical_cardinality;
  a.numeric = user_cat ? false : (binary ? true : !auto_cat);
  a.n_nan = nan;
  a.n_nonzero = nonzero;
  return a;
}

// ---------------------------------------------------------------------------
// Exclusive Feature Bundling (data/binned/bundling.rs)
// ---------------------------------------------------------------------------
// The conflict-graph row sample draws from two crates the reference pins in
// Cargo.lock and that are not in its tree: rand_xoshiro 0.6.0
// (Xoshiro256PlusPlus::seed_from_u64 fills the 256-bit state with four
// SplitMix64 outputs of the seed; next_u64 is Blackman and Vigna's
// xoshiro256++) and rand 0.8.5 (gen_range(0..=i) on usize is
// UniformInt::sample_single_inclusive: the high word of the 128-bit product of
// a draw and the range, accepted when the low word is at most the
// conservative zone (range << leading_zeros(range)) - 1).
// (splitmix64_next above is the same SplitMix64 step.)
struct Xoshiro256PlusPlus {
  uint64_t s[4];
  static Xoshiro256PlusPlus seed_from_u64(uint64_t seed) {
    Xoshiro256PlusPlus r;
    uint64_t sm = seed;
    for (int i = 0; i < 4; i++) r.s[i] = splitmix64_next(sm);  // little-endian words of the seed
    return r;
  }
  uint64_t next_u64() {
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
};

template <class Draw>
uint64_t sample_single_inclusive(uint64_t low, uint64_t high, Draw draw) {
  const uint64_t range = high - low + 1;
  if (range == 0) return draw();
  const uint64_t zone = (range << __builtin_clzll(range)) - 1;
  for (;;) {
    const unsigned __int128 m = (unsigned __int128)draw() * range;
    if ((uint64_t)m <= zone) return low + (uint64_t)(m >> 64);
  }
}

// sample_rows, bundling.rs:203-251.
std::vector<size_t> efb_sample_rows(size_t n_rows, size_t max_samples, uint64_t seed) {
  std::vector<size_t> result;
  if (n_rows <= max_samples) {
    for (size_t i = 0; i < n_rows; i++) result.push_back(i);
    return result;
  }
  const size_t n_first = max_samples / 10, n_last = max_samples / 10;
  const size_t n_random = max_samples - n_first - n_last;
  for (size_t i = 0; i < std::min(n_first, n_rows); i++) result.push_back(i);
  const size_t last_start = n_rows > n_last ? n_rows - n_last : 0;
  for (size_t i = last_start; i < n_rows; i++) result.push_back(i);
  const size_t middle_start = n_first, middle_end = last_start;
  if (middle_end > middle_start) {
    Xoshiro256PlusPlus rng = Xoshiro256PlusPlus::seed_from_u64(seed);
    std::vector<size_t> middle;
    for (size_t i = middle_start; i < middle_end; i++) middle.push_back(i);
    std::vector<size_t> selected(middle.begin(),
                                 middle.begin() + (ptrdiff_t)std::min(n_random, middle.size()));
    for (size_t i = selected.size(); i < middle.size(); i++) {
      const size_t j = (size_t)sample_single_inclusive(0, i, [&] { return rng.next_u64(); });
      if (j < selected.size()) selected[j] = middle[i];
    }
    result.insert(result.end(), selected.begin(), selected.end());
  }
  std::sort(result.begin(), result.end());
  result.erase(std::unique(result.begin(), result.end()), result.end());
  return result;
}

// create_bundle_plan, bundling.rs:393-485 (ConflictGraph::build :91-165,
// assign_bundles :277-339), BundlingConfig defaults (:55-65) with
// min_sparsity = the binning config's sparsity threshold (dataset.rs:229-236).
// Returns the bundles of two or more features, in creation order, each in
// insertion order (original feature indices).
std::vector<std::vector<uint32_t>> create_bundle_plan(const std::vector<Analysis>& an,
                                                      const std::vector<Mapper>& maps,
                                                      const float* X, int64_t n,
                                                      const bst_binning_config& cfg) {
  const float max_conflict_rate = 0.0001f;
  const size_t max_bundle_size = 256, max_sparse_features = 1000, max_sample_rows = 10000;
  auto density = [&](size_t f) { return (float)an[f].n_nonzero / (float)n; };
  const float sparse_threshold = 1.0f - cfg.sparsity_threshold;
  std::vector<size_t> sparse;  // original indices of sparse categorical features
  for (size_t f = 0; f < an.size(); f++)
    if (!maps[f].trivial && !an[f].numeric && density(f) <= sparse_threshold) sparse.push_back(f);
  if (sparse.size() < 2 || sparse.size() > max_sparse_features) return {};
  const std::vector<size_t> rows = efb_sample_rows((size_t)n, max_sample_rows, 42);
  // Bitset per feature over the sampled rows, conflicts = |intersection|.
  const size_t m = sparse.size();
  std::vector<std::vector<bool>> bits(m, std::vector<bool>(rows.size(), false));
  for (size_t i = 0; i < m; i++)
    for (size_t r = 0; r < rows.size(); r++) {
      const float v = X[(int64_t)sparse[i] * n + (int64_t)rows[r]];
      bits[i][r] = v != 0.0f && !std::isnan(v);
    }
  std::map<std::pair<size_t, size_t>, size_t> conflicts;
  for (size_t i = 0; i < m; i++)
    for (size_t j = i + 1; j < m; j++) {
      size_t c = 0;
      for (size_t r = 0; r < rows.size(); r++) c += bits[i][r] && bits[j][r];
      if (c > 0) conflicts[{i, j}] = c;
    }
  auto get_conflict = [&](size_t i, size_t j) {
    auto it = conflicts.find(i < j ? std::make_pair(i, j) : std::make_pair(j, i));
    return it == conflicts.end() ? (size_t)0 : it->second;
  };
  const size_t total_pairs = m * (m - 1) / 2;
  if ((float)conflicts.size() / (float)total_pairs > 0.5f) return {};
  const size_t max_conflicts = (size_t)std::ceil(max_conflict_rate * (float)rows.size());
  // Denser first; sort_by is stable.
  std::vector<size_t> sorted(m);
  for (size_t i = 0; i < m; i++) sorted[i] = i;
  std::stable_sort(sorted.begin(), sorted.end(), [&](size_t a, size_t b) {
    return density(sparse[b]) < density(sparse[a]);
  });
  std::vector<std::vector<size_t>> bundles;  // sparse positions
  std::vector<size_t> bundle_conflicts;
  for (size_t p : sorted) {
    size_t best_bundle = SIZE_MAX, best_conflict_increase = SIZE_MAX;
    for (size_t b = 0; b < bundles.size(); b++) {
      if (bundles[b].size() >= max_bundle_size) continue;
      size_t conflict_sum = 0;
      for (size_t q : bundles[b]) conflict_sum += get_conflict(p, q);
      if (bundle_conflicts[b] + conflict_sum <= max_conflicts &&
          conflict_sum < best_conflict_increase) {
        best_bundle = b;
        best_conflict_increase = conflict_sum;
      }
    }
    if (best_bundle != SIZE_MAX) {
      bundles[best_bundle].push_back(p);
      bundle_conflicts[best_bundle] += best_conflict_increase;
    } else {
      bundles.push_back({p});
      bundle_conflicts.push_back(0);
    }
  }
  std::vector<std::vector<uint32_t>> out;
  for (auto& b : bundles)
    if (b.size() >= 2) {
      std::vector<uint32_t> fs;
      for (size_t p : b) fs.push_back((uint32_t)sparse[p]);
      out.push_back(fs);
    }
  return out;
}

// ---------------------------------------------------------------------------
// Objectives (training/objectives/)
// ---------------------------------------------------------------------------
inline float wt(const float* w, int64_t i) { return w ? w[i] : 1.0f; }

// The enum payloads of Objective / Metric that the arithmetic needs
// (objectives/mod.rs:196-219; metrics/mod.rs:178-207).
struct Payload {
  float huber_delta = 1.0f;
  std::vector<float> alphas;
  float accuracy_threshold = 0.5f;
  double metric_huber_delta = 1.0;
};
Payload payload_of(const bst_gbdt_config* c) {
  Payload P;
  if (!c) return P;
  P.huber_delta = c->huber_delta;
  if (c->alphas && c->n_alphas > 0) P.alphas.assign(c->alphas, c->alphas + c->n_alphas);
  P.accuracy_threshold = c->accuracy_threshold;
  P.metric_huber_delta = c->metric_huber_delta;
  return P;
}

// Objective::n_outputs, objectives/mod.rs:221-233.
int n_outputs_of(const bst_gbdt_config& c) {
  if (c.objective == BST_OBJ_SOFTMAX) return (int)c.n_classes;
  if (c.objective == BST_OBJ_PINBALL) return (int)c.n_alphas;
  return 1;
}

// Rust f32::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives,
// NaN for NaN.
inline float signum32(float v) { return std::isnan(v) ? v : std::copysign(1.0f, v); }

void gradients(int obj, int K, const float* preds, const float* y, const float* w, int64_t n,
               float* g, float* h, const Payload& P = Payload()) {
  const float HESS_MIN = 1e-6f;
  if (obj == BST_OBJ_ABSOLUTE) {  // regression.rs:80-104
    for (int k = 0; k < K; k++)
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        float diff = preds[k * n + i] - y[i];
        g[k * n + i] = wi * signum32(diff);
        h[k * n + i] = wi;
      }
    return;
  }
  if (obj == BST_OBJ_PINBALL) {  // regression.rs:129-156
    for (int k = 0; k < K; k++) {
      const float alpha = P.alphas[k];
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        float diff = preds[k * n + i] - y[i];
        float gk = diff < 0.0f ? -alpha : 1.0f - alpha;
        g[k * n + i] = wi * gk;
        h[k * n + i] = wi;
      }
    }
    return;
  }
  if (obj == BST_OBJ_PSEUDO_HUBER) {  // regression.rs:185-216
    const float delta = P.huber_delta;
    const float inv_delta_sq = 1.0f / (delta * delta);
    for (int k = 0; k < K; k++)
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        float residual = preds[k * n + i] - y[i];
        float r_sq = residual * residual;
        float factor = 1.0f + r_sq * inv_delta_sq;
        float sqrt_factor = std::sqrt(factor);
        g[k * n + i] = wi * residual / sqrt_factor;
        h[k * n + i] = wi / (factor * sqrt_factor);
      }
    return;
  }
  if (obj == BST_OBJ_POISSON) {  // regression.rs:244-270
    const float MAX_EXP = 30.0f;
    for (int k = 0; k < K; k++)
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        float pc = preds[k * n + i];  // f32::clamp keeps NaN
        if (pc < -MAX_EXP) pc = -MAX_EXP;
        if (pc > MAX_EXP) pc = MAX_EXP;
        float e = exp32(pc);
        g[k * n + i] = wi * (e - y[i]);
        h[k * n + i] = fmaxf(wi * e, HESS_MIN);  // f32::max ignores NaN
      }
    return;
  }
  if (obj == BST_OBJ_HINGE) {  // classification.rs:87-116
    for (int k = 0; k < K; k++)
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        float yy = y[i] > 0.5f ? 1.0f : -1.0f;
        float margin = yy * preds[k * n + i];
        g[k * n + i] = margin < 1.0f ? -wi * yy : 0.0f;
        h[k * n + i] = wi;
      }
    return;
  }
  if (obj == BST_OBJ_SQUARED) {  // regression.rs:20-42
    for (int k = 0; k < K; k++)
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        g[k * n + i] = wi * (preds[k * n + i] - y[i]);
        h[k * n + i] = wi;
      }
  } else if (obj == BST_OBJ_LOGISTIC) {  // classification.rs:20-52
    for (int k = 0; k < K; k++)
      for (int64_t i = 0; i < n; i++) {
        float wi = wt(w, i);
        float p = 1.0f / (1.0f + exp32(-preds[k * n + i]));
        g[k * n + i] = wi * (p - y[i]);
        h[k * n + i] = fmaxf(wi * p * (1.0f - p), HESS_MIN);  // f32::max ignores NaN
      }
  } else if (obj == BST_OBJ_SOFTMAX) {  // classification.rs:141-187
    for (int64_t i = 0; i < n; i++) {
      float wi = wt(w, i);
      size_t label = (size_t)y[i];
      float mx = -std::numeric_limits<float>::infinity();
      for (int k = 0; k < K; k++) mx = fmaxf(mx, preds[k * n + i]);
      float s = 0.0f;
      for (int k = 0; k < K; k++) s += exp32(preds[k * n + i] - mx);
      for (int k = 0; k < K; k++) {
        float p = exp32(preds[k * n + i] - mx) / s;
        float ind = (size_t)k == label ? 1.0f : 0.0f;
        g[k * n + i] = wi * (p - ind);
        h[k * n + i] = fmaxf(wi * p * (1.0f - p), HESS_MIN);  // f32::max ignores NaN
      }
    }
  }
}

// LambdaRank gradients over ragged query groups.  The reference has no
// ranking objective (objectives/mod.rs:196-219), so this follows the
// published LambdaMART formulation (Burges, "From RankNet to LambdaRank to
// LambdaMART: An Overview", MSR-TR-2010-82) with sigma = 1 and the pair
// truncation of LightGBM's lambdarank objective, as this engine defines it:
//   - documents of a group ranked by score descending, ties by index;
//   - pair (a, b) of positions counts when labels differ and min(a, b) < T;
//   - rho = 1 / (1 + exp(s_hi - s_lo)) in f32 (s_hi: the higher label);
//   - dNDCG = (2^l_hi - 2^l_lo) |disc[a] - disc[b]| / maxDCG@T, disc[r] = 1/log2(2 + r);
//   - the higher-labelled document gets -rho dNDCG, the other +rho dNDCG,
//     both get rho (1 - rho) dNDCG; each document sums its pairs in f64 in
//     position order, then scales by its weight and rounds to f32.
// No lambda normalisation.  Parity with any reference: unpinned.
constexpr int kRankTrunc = 30;

std::vector<double> rank_discounts(int64_t max_group) {
  std::vector<double> d((size_t)std::max<int64_t>(max_group, 1));
  for (size_t r = 0; r < d.size(); r++) d[r] = 1.0 / std::log2(2.0 + (double)r);
  return d;
}

void lambdarank_gradients(const float* preds, const float* y, const float* w, int64_t n,
                          const int32_t* sizes, int64_t n_groups, float* g, float* h) {
  int64_t max_group = 1;
  for (int64_t q = 0; q < n_groups; q++) max_group = std::max<int64_t>(max_group, sizes[q]);
  const std::vector<double> disc = rank_discounts(max_group);
  std::vector<int64_t> off((size_t)n_groups + 1, 0);
  for (int64_t q = 0; q < n_groups; q++) off[q + 1] = off[q] + sizes[q];
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < n_groups; q++) {
    const int64_t b0 = off[q], m = sizes[q];
    std::vector<int64_t> order((size_t)m);
    for (int64_t i = 0; i < m; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return preds[b0 + a] > preds[b0 + b]; });
    auto label = [&](int64_t i) { return (int)y[b0 + i]; };
    std::vector<int> ls((size_t)m);
    for (int64_t i = 0; i < m; i++) ls[i] = label(i);
    std::sort(ls.begin(), ls.end(), std::greater<int>());
    double maxdcg = 0.0;
    for (int64_t r = 0; r < std::min<int64_t>(m, kRankTrunc); r++)
      maxdcg += (std::ldexp(1.0, ls[r]) - 1.0) * disc[r];
    for (int64_t a = 0; a < m; a++) {
      const int64_t i = order[a];
      double ga = 0.0, ha = 0.0;
      if (maxdcg > 0.0) {
        const double inv = 1.0 / maxdcg;
        for (int64_t b = 0; b < m; b++) {
          if (b == a || std::min(a, b) >= kRankTrunc) continue;
          const int64_t j = order[b];
          const int li = label(i), lj = label(j);
          if (li == lj) continue;
          const bool i_high = li > lj;
          const float shi = i_high ? preds[b0 + i] : preds[b0 + j];
          const float slo = i_high ? preds[b0 + j] : preds[b0 + i];
          const float rho = 1.0f / (1.0f + exp32(shi - slo));
          const double gap = std::ldexp(1.0, i_high ? li : lj) - std::ldexp(1.0, i_high ? lj : li);
          const double dn = gap * std::fabs(disc[a] - disc[b]) * inv;
          const double lam = (double)rho * dn;
          ga += i_high ? -lam : lam;
          ha += (double)rho * (1.0 - (double)rho) * dn;
        }
      }
      const double wi = (double)wt(w, b0 + i);
      g[b0 + i] = (float)(ga * wi);
      h[b0 + i] = (float)(ha * wi);
    }
  }
}

// compute_weighted_quantile, regression.rs:307-333: (value, weight) pairs
// sorted by value (a stable sort), total weight summed in f32 in sorted
// order, the first value whose running f32 weight reaches alpha * total.
float weighted_quantile(const float* v, const float* w, int64_t n, float alpha) {
  if (n == 0) return 0.0f;
  std::vector<std::pair<float, float>> sorted((size_t)n);
  for (int64_t i = 0; i < n; i++) sorted[i] = {v[i], wt(w, i)};
  std::stable_sort(sorted.begin(), sorted.end(),
                   [](const std::pair<float, float>& a, const std::pair<float, float>& b) {
                     return a.first < b.first;
                   });
  float total = 0.0f;
  for (auto& p : sorted) total += p.second;
  const float target = alpha * total;
  float cum = 0.0f;
  for (auto& p : sorted) {
    cum += p.second;
    if (cum >= target) return p.first;
  }
  return sorted.back().first;
}

// unweighted_quantile_indexed, utils.rs:17-53: the k-th smallest value,
// k = ceil(alpha * n as f32) - 1 clamped to [0, n - 1].
template <typename V>
float unweighted_quantile_indexed(const uint32_t* ind, uint32_t n, V value, float alpha) {
  if (n == 0) return std::numeric_limits<float>::quiet_NaN();
  if (n == 1) return value(ind[0]);
  int64_t k = (int64_t)std::ceil(alpha * (float)n) - 1;
  if (k < 0) k = 0;
  if (k > (int64_t)n - 1) k = (int64_t)n - 1;
  std::vector<float> vals(n);
  for (uint32_t i = 0; i < n; i++) vals[i] = value(ind[i]);
  std::nth_element(vals.begin(), vals.begin() + k, vals.end());
  return vals[(size_t)k];
}

// weighted_quantile_indexed, utils.rs:81-126: indices sorted by value (stable,
// ties keep the leaf's row order), total weight summed in f32 over the
// indices in the leaf's order, running f32 weight in sorted order.
template <typename V>
float weighted_quantile_indexed(const uint32_t* ind, uint32_t n, V value, const float* w,
                                float alpha) {
  if (n == 0) return std::numeric_limits<float>::quiet_NaN();
  if (n == 1) return value(ind[0]);
  std::vector<uint32_t> sc(ind, ind + n);
  std::stable_sort(sc.begin(), sc.end(), [&](uint32_t a, uint32_t b) { return value(a) < value(b); });
  float total = 0.0f;
  for (uint32_t i = 0; i < n; i++) total += w[ind[i]];
  const float threshold = total * alpha;
  float cum = 0.0f;
  for (uint32_t i = 0; i < n; i++) {
    cum += w[sc[i]];
    if (cum >= threshold) return value(sc[i]);
  }
  return value(sc[n - 1]);
}

std::vector<float> base_score(int obj, int K, const float* y, const float* w, int64_t n,
                              const Payload& P = Payload()) {
  if (obj == BST_OBJ_LAMBDARANK) return {0.0f};
  if (obj == BST_OBJ_HINGE) return {0.0f};  // classification.rs:118-123
  if (obj == BST_OBJ_ABSOLUTE || obj == BST_OBJ_PSEUDO_HUBER)  // regression.rs:106-117,218-230
    return {weighted_quantile(y, w, n, 0.5f)};
  if (obj == BST_OBJ_PINBALL) {  // regression.rs:158-172
    std::vector<float> out;
    for (int k = 0; k < K; k++) out.push_back(n == 0 ? 0.0f : weighted_quantile(y, w, n, P.alphas[k]));
    return out;
  }
  if (obj == BST_OBJ_POISSON) {  // regression.rs:272-296
    if (n == 0) return {0.0f};
    double sw = 0, swy = 0;
    for (int64_t i = 0; i < n; i++) {
      sw += (double)wt(w, i);
      swy += (double)wt(w, i) * (double)y[i];
    }
    double mean = sw > 0.0 ? std::max(swy / sw, 1e-7) : 1.0;
    return {(float)std::log(mean)};
  }
  if (obj == BST_OBJ_SQUARED) {  // regression.rs:44-68
    if (n == 0) return {0.0f};
    double sw = 0, swy = 0;
    for (int64_t i = 0; i < n; i++) {
      sw += (double)wt(w, i);
      swy += (double)wt(w, i) * (double)y[i];
    }
    return {sw > 0 ? (float)(swy / sw) : 0.0f};
  }
  if (obj == BST_OBJ_LOGISTIC) {  // classification.rs:54-76
    if (n == 0) return {0.0f};
    double pos = 0, tot = 0;
    for (int64_t i = 0; i < n; i++) {
      pos += (double)wt(w, i) * (double)y[i];
      tot += (double)wt(w, i);
    }
    double p = std::min(std::max(pos / tot, 1e-7), 1.0 - 1e-7);
    return {(float)std::log(p / (1.0 - p))};
  }
  // softmax, classification.rs:189-222
  std::vector<float> out(K, 0.0f);
  if (n == 0) return out;
  std::vector<double> cw(K, 0.0);
  double tot = 0;
  for (int64_t i = 0; i < n; i++) {
    size_t l = (size_t)y[i];
    if (l < (size_t)K) cw[l] += (double)wt(w, i);
    tot += (double)wt(w, i);
  }
  for (int k = 0; k < K; k++) {
    double p = std::min(std::max(cw[k] / tot, 1e-7), 1.0 - 1e-7);
    out[k] = (float)std::log(p);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Split finding (training/gbdt/split/)
// ---------------------------------------------------------------------------
struct Gain {  // NodeGainContext, gain.rs:16-52
  double lambda, offset;
  Gain(double pg, double ph, const bst_gbdt_config& c) {
    lambda = (double)c.lambda;
    double ps = pg * pg / (ph + lambda);
    offset = 0.5 * ps + (double)c.min_gain;
  }
  float gain(double gl, double hl, double gr, double hr) const {
    double sl = gl * gl / (hl + lambda);
    double sr = gr * gr / (hr + lambda);
    return (float)(0.5 * (sl + sr) - offset);
  }
};

// GainParams::is_valid_split, gain.rs:123-137.
inline bool valid_split(const bst_gbdt_config& c, double hl, double hr, uint32_t cl, uint32_t cr) {
  double mw = (double)c.min_child_weight;
  return hl >= mw && hr >= mw && cl >= c.min_samples_leaf && cr >= c.min_samples_leaf;
}

// GainParams::compute_leaf_weight, gain.rs:151-168.
float leaf_weight(const bst_gbdt_config& c, double g, double h) {
  double lam = c.lambda, al = c.alpha;
  if (al == 0.0) return (float)(-g / (h + lam));
  double ag = std::fabs(g);
  if (ag <= al) return 0.0f;
  double s = g > 0.0 ? -1.0 : 1.0;
  return (float)(s * (ag - al) / (h + lam));
}

inline uint32_t sat_sub(uint32_t a, uint32_t b) { return a > b ? a - b : 0; }

ora_split invalid_split() {
  ora_split s;
  memset(&s, 0, sizeof(s));
  s.gain = -std::numeric_limits<float>::infinity();
  return s;
}

// find_numerical_split, find.rs:286-407 (non-bundle columns).
ora_split numeric_split(const double* bins, uint32_t nb, uint32_t feat, double pg, double ph,
                        uint32_t pc, bool has_missing, const Gain& ctx,
                        const bst_gbdt_config& c) {
  ora_split best = invalid_split();
  if (nb < 2) return best;
  double lg = 0, lh = 0;
  uint32_t lc = 0;
  for (uint32_t b = 0; b + 1 < nb; b++) {
    lg += bins[2 * b];
    lh += bins[2 * b + 1];
    lc += 1;
    double rg = pg - lg, rh = ph - lh;
    uint32_t rc = sat_sub(pc, lc);
    if (valid_split(c, lh, rh, lc, rc)) {
      float g = ctx.gain(lg, lh, rg, rh);
      if (g > best.gain) {
        best.feature = feat;
        best.gain = g;
        best.default_left = 0;
        best.is_categorical = 0;
        best.bin = b;
      }
    }
  }
  if (has_missing) {
    double rg = 0, rh = 0;
    uint32_t rc = 0;
    uint32_t max_excl = nb - 1;  // the missing bin is the last one
    for (uint32_t b = max_excl; b-- > 1;) {
      rg += bins[2 * b];
      rh += bins[2 * b + 1];
      rc += 1;
      double lg2 = pg - rg, lh2 = ph - rh;
      uint32_t lc2 = sat_sub(pc, rc);
      if (valid_split(c, lh2, rh, lc2, rc)) {
        float g = ctx.gain(lg2, lh2, rg, rh);
        if (g > best.gain) {
          best.feature = feat;
          best.gain = g;
          best.default_left = 1;
          best.is_categorical = 0;
          best.bin = b - 1;
        }
      }
    }
  }
  return best;
}

inline void cat_insert(ora_split& s, uint32_t c) { s.left_cats[c >> 5] |= 1u << (c & 31); }

// find_categorical_split / find_onehot_split / find_sorted_split, find.rs:418-573.
ora_split categorical_split(const double* bins, uint32_t nc, uint32_t feat, double pg, double ph,
                            uint32_t pc, const Gain& ctx, const bst_gbdt_config& c) {
  ora_split best = invalid_split();
  if (nc < 2) return best;
  if (nc <= c.max_onehot_cats) {
    for (uint32_t k = 0; k < nc; k++) {
      double lg = bins[2 * k], lh = bins[2 * k + 1];
      uint32_t lc = 1;
      double rg = pg - lg, rh = ph - lh;
      uint32_t rc = sat_sub(pc, lc);
      if (valid_split(c, lh, rh, lc, rc)) {
        float g = ctx.gain(lg, lh, rg, rh);
        if (g > best.gain) {
          best = invalid_split();
          best.feature = feat;
          best.gain = g;
          best.is_categorical = 1;
          cat_insert(best, k);
        }
      }
    }
    return best;
  }
  std::vector<std::pair<uint32_t, double>> sc;
  for (uint32_t k = 0; k < nc; k++) {
    double g = bins[2 * k], h = bins[2 * k + 1];
    if (h > 1e-10) sc.push_back({k, -g / (h + 1e-6)});
  }
  if (sc.size() < 2) return best;
  std::stable_sort(sc.begin(), sc.end(),
                   [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& b) {
                     return a.second < b.second;
                   });
  ora_split cur = invalid_split();
  double lg = 0, lh = 0;
  uint32_t lc = 0;
  for (size_t i = 0; i + 1 < sc.size(); i++) {
    uint32_t k = sc[i].first;
    cat_insert(cur, k);
    lg += bins[2 * k];
    lh += bins[2 * k + 1];
    lc += 1;
    double rg = pg - lg, rh = ph - lh;
    uint32_t rc = sat_sub(pc, lc);
    if (valid_split(c, lh, rh, lc, rc)) {
      float g = ctx.gain(lg, lh, rg, rh);
      if (g > best.gain) {
        best = cur;
        best.feature = feat;
        best.gain = g;
        best.default_left = 0;
        best.is_categorical = 1;
      }
    }
  }
  return best;
}

// GreedySplitter::find_split, find.rs:113-270.  parallel = rayon reduce
// (a.gain > b.gain ? a : b), which returns the rightmost maximum; sequential
// keeps the first maximum.
ora_split find_split(const ora_binned& B, const double* hist, double pg, double ph, uint32_t pc,
                     const bst_gbdt_config& c, int n_threads, const std::vector<uint32_t>& feats) {
  Gain ctx(pg, ph, c);
  bool parallel = n_threads != 1 && feats.size() >= 16;
  std::vector<ora_split> per(feats.size());
#pragma omp parallel for schedule(dynamic, 4) if (parallel)
  for (size_t j = 0; j < feats.size(); j++) {
    uint32_t f = feats[j];
    const double* bins = hist + 2 * (size_t)B.eff_offset[f];
    per[j] = B.eff_cat[f] ? categorical_split(bins, B.eff_nbins[f], f, pg, ph, pc, ctx, c)
                          : numeric_split(bins, B.eff_nbins[f], f, pg, ph, pc,
                                          B.eff_missing[f] != 0, ctx, c);
  }
  ora_split best = invalid_split();
  for (auto& s : per) {
    if (parallel ? !(best.gain > s.gain) : (s.gain > best.gain)) best = s;
  }
  return best;
}

// ---------------------------------------------------------------------------
// Histograms (training/gbdt/histograms/ops.rs)
// ---------------------------------------------------------------------------
void build_hist(const ora_binned& B, const float* g, const float* h, const uint32_t* rows,
                int64_t nr, double* out) {
  int64_t total = (int64_t)B.total_bins * 2;
  std::fill(out, out + total, 0.0);
  int64_t F = (int64_t)B.eff_to_orig.size();
  // Feature-parallel, one feature per task, like the rayon path (ops.rs:92-99).
#pragma omp parallel for schedule(dynamic, 1) if (F >= 4 && nr * F >= 4096)
  for (int64_t f = 0; f < F; f++) {
    double* hs = out + 2 * (size_t)B.eff_offset[f];
    const uint16_t* col = B.bins[f].data();
    if (rows) {
      for (int64_t i = 0; i < nr; i++) {
        uint32_t r = rows[i];
        uint32_t b = col[r];
        hs[2 * b] += (double)g[r];
        hs[2 * b + 1] += (double)h[r];
      }
    } else {
      for (int64_t r = 0; r < nr; r++) {
        uint32_t b = col[r];
        hs[2 * b] += (double)g[r];
        hs[2 * b + 1] += (double)h[r];
      }
    }
  }
}

// Goes-left predicate, partition.rs:321-337.
inline bool goes_left(const ora_binned& B, const ora_split& s, uint32_t bin) {
  uint32_t f = s.feature;
  // Bundles: missing bin 0 (partition.rs:311-318); other columns: the mapper's.
  bool has_missing = B.eff_missing[f] != 0;
  if (has_missing && bin == B.eff_default[f]) return s.default_left != 0;
  if (s.is_categorical) return bin < 256 && ((s.left_cats[bin >> 5] >> (bin & 31)) & 1u);
  return bin <= s.bin;
}

// stable_partition_range, partition.rs:73-145.
int64_t stable_partition(const ora_binned& B, uint32_t* rows, int64_t n, const ora_split& s,
                         std::vector<uint32_t>& scratch) {
  scratch.resize(n);
  const uint16_t* col = B.bins[s.feature].data();
  int64_t lw = 0, rw = n;
  for (int64_t i = 0; i < n; i++) {
    uint32_t r = rows[i];
    if (goes_left(B, s, col[r]))
      scratch[lw++] = r;
    else
      scratch[--rw] = r;
  }
  std::reverse(scratch.begin() + lw, scratch.begin() + n);
  std::copy(scratch.begin(), scratch.begin() + n, rows);
  return lw;
}

}  // namespace

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------
struct OTree {
  std::vector<uint32_t> split_index, left, right;
  std::vector<float> thr, leaf, gain, cover;
  std::vector<uint8_t> dleft, is_leaf, stype;
  std::vector<std::pair<uint32_t, std::vector<uint32_t>>> cat_nodes;
  uint32_t alloc() {
    split_index.push_back(0);
    left.push_back(0);
    right.push_back(0);
    thr.push_back(0.0f);
    leaf.push_back(0.0f);
    gain.push_back(0.0f);
    cover.push_back(0.0f);
    dleft.push_back(0);
    is_leaf.push_back(0);
    stype.push_back(0);
    return (uint32_t)split_index.size() - 1;
  }
};

struct ora_model {
  int32_t objective = 0, n_groups = 1, n_features = 0, transform = 0;
  std::vector<float> base;
  std::vector<OTree> trees;
  std::vector<uint32_t> groups;
  std::vector<double> metric;
  int64_t hist_rows = 0;
  double ms_total = 0;
};

namespace {

// next_up_f32, grower.rs:543-556.
float next_up(float x) {
  if (std::isnan(x) || x == std::numeric_limits<float>::infinity()) return x;
  if (x == 0.0f) {
    uint32_t one = 1;
    float r;
    memcpy(&r, &one, 4);
    return r;
  }
  uint32_t b;
  memcpy(&b, &x, 4);
  b = x > 0 ? b + 1 : b - 1;
  float r;
  memcpy(&r, &b, 4);
  return r;
}

struct Cand {  // NodeCandidate, expansion.rs:242-320
  uint32_t node, tree_node;
  uint32_t depth;
  ora_split split;
  double g, h;
  uint32_t count;
};
struct CandLess {  // BinaryHeap max by (gain, partition node id)
  bool operator()(const Cand& a, const Cand& b) const {
    float ga = a.split.gain, gb = b.split.gain;
    if (ga < gb) return true;
    if (ga > gb) return false;
    return a.node < b.node;
  }
};

// Counter-based sampling keys.  The reference draws row subsamples, GOSS
// small-gradient rows and column subsets from rand 0.8.5 streams (row.rs:106,
// 185,256; column.rs:138-142), which no test pins and which column sampling
// seeds from OS entropy.  This engine defines each draw as "the k candidates
// with the smallest (key, index)", key = a 31-bit splitmix64 hash of (seed,
// stream, round, output, candidate index): the same distributions (uniform
// k-subsets) with draws that parallel hardware reproduces exactly.
constexpr uint32_t kStreamRows = 1, kStreamGoss = 2, kStreamCols = 3, kStreamLevel = 4;

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t key_base(uint64_t seed, uint32_t stream, uint32_t round, uint32_t output) {
  return mix64(seed ^ mix64(((uint64_t)round << 32) | ((uint64_t)output << 8) | stream));
}
inline uint32_t key_of(uint64_t base, uint32_t i) {
  return (uint32_t)(mix64(base ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull)) >> 33);
}
// The k candidates with the smallest (key, index), returned in index order.
std::vector<uint32_t> smallest_keys(const std::vector<uint32_t>& cand, uint64_t base, size_t k) {
  std::vector<std::pair<uint32_t, uint32_t>> kv(cand.size());
  for (size_t j = 0; j < cand.size(); j++) kv[j] = {key_of(base, cand[j]), cand[j]};
  k = std::min(k, kv.size());
  std::nth_element(kv.begin(), kv.begin() + (k ? k - 1 : 0), kv.end());
  std::vector<uint32_t> out;
  if (k) {
    // nth_element leaves the k smallest in [0, k) (unordered) only up to the
    // pivot; take everything <= the k-th smallest pair.
    const auto piv = kv[k - 1];
    for (auto& p : kv)
      if (p <= piv) out.push_back(p.second);
  }
  std::sort(out.begin(), out.end());
  return out;
}

// Row and column sampling rules (training/sampling/{row,column}.rs); the
// draws are the smallest keys of the stream, the counts and the gradient
// edits are the reference's.

// sample_uniform, row.rs:170-208: target_count = (n as f32 * subsample) as
// usize, clamped to [1, n]; every row not drawn gets (g, h) = (0, 0).
// Returns the drawn rows in index order.
std::vector<uint32_t> sample_uniform(float* g, float* h, int64_t n, float subsample, uint64_t base) {
  const size_t tc = std::min<size_t>(std::max<size_t>((size_t)((float)n * subsample), 1), (size_t)n);
  std::vector<uint32_t> all((size_t)n);
  for (int64_t i = 0; i < n; i++) all[i] = (uint32_t)i;
  std::vector<uint32_t> sampled = smallest_keys(all, base, tc);
  std::vector<uint8_t> mask((size_t)n, 0);
  for (auto r : sampled) mask[r] = 1;
  for (int64_t i = 0; i < n; i++)
    if (!mask[i]) g[i] = h[i] = 0.0f;
  return sampled;
}

// sample_goss, row.rs:211-270.  top_k = (n as f32 * top_rate) as usize
// clamped to [1, n - 1]; other_k = max((n as f32 * other_rate) as usize, 1);
// the threshold is the (n - top_k)-th smallest |g h| (select_nth_unstable);
// rows with |g h| >= threshold are kept as they are; other_k rows of the rest
// (all of them if fewer) are drawn and their g and h multiplied by
// (n - top_k) / other_k in f32.  Returns the kept rows in index order.
std::vector<uint32_t> sample_goss(float* g, float* h, int64_t n, float top_rate, float other_rate,
                                  uint64_t base) {
  std::vector<float> mag((size_t)n);
  for (int64_t i = 0; i < n; i++) mag[i] = std::fabs(g[i] * h[i]);
  const size_t top_k = std::min<size_t>(std::max<size_t>((size_t)((float)n * top_rate), 1), (size_t)n - 1);
  const size_t other_k = std::max<size_t>((size_t)((float)n * other_rate), 1);
  std::vector<float> sm = mag;
  std::nth_element(sm.begin(), sm.begin() + (n - top_k), sm.end());
  const float thr = sm[n - top_k];
  const float mult = (float)(n - top_k) / (float)other_k;
  std::vector<uint32_t> sampled, small;
  for (int64_t i = 0; i < n; i++)
    if (mag[i] >= thr)
      sampled.push_back((uint32_t)i);
    else
      small.push_back((uint32_t)i);
  std::vector<uint32_t> pick = smallest_keys(small, base, other_k);
  for (auto r : pick) {
    g[r] *= mult;
    h[r] *= mult;
  }
  sampled.insert(sampled.end(), pick.begin(), pick.end());
  std::sort(sampled.begin(), sampled.end());
  return sampled;
}

// GOSS warm-up, row.rs:109-112: no sampling before round (1.0 / lr) as usize
// (an f32 division).
size_t goss_warmup(float learning_rate) { return (size_t)(1.0f / learning_rate); }

// ColSampler::sample_tree, column.rs:180-186: k = ceil(F as f32 * rate),
// clamped to [1, F]; the product is an f32 product.
uint32_t colsample_k(uint32_t F, float rate) {
  const float prod = (float)F * rate;
  uint64_t k = (uint64_t)std::ceil(prod);
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(k, 1), F);
}

struct Grower {
  const ora_binned& B;
  const bst_gbdt_config& c;
  int n_threads;
  std::vector<uint32_t> idx, scratch;
  std::vector<uint32_t> leaf_begin, leaf_count;
  uint32_t n_part = 0;
  std::vector<std::vector<double>> hist;  // by partition node id
  std::vector<float> last_leaf;           // by partition node id (lr applied)
  std::vector<std::pair<uint32_t, uint32_t>> leaf_map;  // (tree node, partition node) per leaf
  int64_t hist_rows = 0;
  std::vector<uint32_t> feats, tree_feats;
  uint32_t round = 0, output = 0;  // identify the tree for column sampling

  Grower(const ora_binned& b, const bst_gbdt_config& cfg, int nt)
      : B(b), c(cfg), n_threads(nt) {}

  void build(uint32_t node, const float* g, const float* h) {
    auto& hs = hist[node];
    hs.assign((size_t)B.total_bins * 2, 0.0);
    uint32_t nr = leaf_count[node];
    if (nr == 0) return;
    build_hist(B, g, h, idx.data() + leaf_begin[node], nr, hs.data());
    hist_rows += nr;
  }
  std::pair<double, double> sums(uint32_t node) {  // pool.rs:536-549 (feature 0)
    double sg = 0, sh = 0;
    const auto& hs = hist[node];
    for (uint32_t b = 0; b < B.eff_nbins[0]; b++) {
      sg += hs[2 * b];
      sh += hs[2 * b + 1];
    }
    return {sg, sh};
  }
  ora_split split_of(uint32_t node, double g, double h, uint32_t cnt) {
    return find_split(B, hist[node].data(), g, h, cnt, c, n_threads, feats);
  }

  // TreeGrower::grow, grower.rs:265-342 (+ process_candidate :346-481).
  OTree grow(const float* g, const float* h, const std::vector<uint32_t>* sampled) {
    int64_t n = B.n;
    uint32_t max_leaves = c.growth == BST_GROW_DEPTHWISE ? (1u << c.max_depth) : c.max_leaves;
    size_t max_nodes = std::max<size_t>(2 * (size_t)max_leaves - 1, 63);
    if (sampled)
      idx = *sampled;
    else {
      idx.resize(n);
      for (int64_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
    }
    leaf_begin.assign(max_nodes + 2, 0);
    leaf_count.assign(max_nodes + 2, 0);
    leaf_count[0] = (uint32_t)idx.size();
    n_part = 1;
    hist.assign(max_nodes + 2, {});
    last_leaf.assign(max_nodes + 2, std::numeric_limits<float>::quiet_NaN());
    leaf_map.clear();
    sample_columns();

    OTree t;
    uint32_t root = t.alloc();
    build(0, g, h);
    auto [tg, th] = sums(0);
    ora_split rs = split_of(0, tg, th, (uint32_t)n);
    Cand rootc{0, root, 0, rs, tg, th, (uint32_t)n};

    auto make_leaf = [&](const Cand& cd) {
      float w = leaf_weight(c, cd.g, cd.h);
      t.is_leaf[cd.tree_node] = 1;
      t.leaf[cd.tree_node] = w;
      t.cover[cd.tree_node] = (float)cd.h;
      last_leaf[cd.node] = w * c.learning_rate;
      leaf_map.push_back({cd.tree_node, cd.node});
      hist[cd.node].clear();
    };
    auto expand = [&](const Cand& cd, std::vector<Cand>& out) {
      if (!(cd.split.gain > -std::numeric_limits<float>::infinity()) ||
          cd.split.gain <= c.min_gain) {
        make_leaf(cd);
        return;
      }
      t.gain[cd.tree_node] = cd.split.gain;
      t.cover[cd.tree_node] = (float)cd.h;
      uint32_t lt, rt;
      apply_split(t, cd.tree_node, cd.split, lt, rt);
      // partitioner.split: left keeps the id, right gets a new id.
      uint32_t begin = leaf_begin[cd.node], cnt = leaf_count[cd.node];
      int64_t lc = stable_partition(B, idx.data() + begin, cnt, cd.split, scratch);
      uint32_t rnode = n_part++;
      leaf_count[cd.node] = (uint32_t)lc;
      leaf_begin[rnode] = begin + (uint32_t)lc;
      leaf_count[rnode] = cnt - (uint32_t)lc;
      uint32_t lnode = cd.node;
      uint32_t lcount = (uint32_t)lc, rcount = cnt - (uint32_t)lc;
      bool left_small = lcount <= rcount;
      uint32_t small = left_small ? lnode : rnode, large = left_small ? rnode : lnode;
      uint32_t scount = left_small ? lcount : rcount, gcount = left_small ? rcount : lcount;
      std::vector<double> parent = std::move(hist[cd.node]);
      build(small, g, h);
      hist[large] = std::move(parent);
      for (size_t k = 0; k < hist[large].size(); k++) hist[large][k] -= hist[small][k];
      auto [sg, sh] = sums(small);
      double lg_ = cd.g - sg, lh_ = cd.h - sh;
      set_level(cd.depth + 1);
      ora_split ss = split_of(small, sg, sh, scount);
      ora_split ls = split_of(large, lg_, lh_, gcount);
      Cand cl, cr;
      if (left_small) {
        cl = {lnode, lt, cd.depth + 1, ss, sg, sh, scount};
        cr = {rnode, rt, cd.depth + 1, ls, lg_, lh_, gcount};
      } else {
        cl = {lnode, lt, cd.depth + 1, ls, lg_, lh_, gcount};
        cr = {rnode, rt, cd.depth + 1, ss, sg, sh, scount};
      }
      out.push_back(cl);
      out.push_back(cr);
    };

    if (c.growth == BST_GROW_DEPTHWISE) {
      std::vector<Cand> level{rootc}, next;
      uint32_t depth = 0;
      while (depth < c.max_depth && !level.empty()) {
        next.clear();
        for (auto& cd : level) expand(cd, next);
        level.swap(next);
        depth++;
      }
      for (auto& cd : level) make_leaf(cd);
    } else {
      std::priority_queue<Cand, std::vector<Cand>, CandLess> heap;
      heap.push(rootc);
      uint32_t n_leaves = 1;
      while (n_leaves < c.max_leaves && !heap.empty()) {
        Cand cd = heap.top();
        heap.pop();
        n_leaves++;
        std::vector<Cand> kids;
        expand(cd, kids);
        for (auto& k : kids) heap.push(k);
      }
      while (!heap.empty()) {
        make_leaf(heap.top());
        heap.pop();
      }
    }
    for (size_t i = 0; i < t.leaf.size(); i++)  // apply_learning_rate, mutable_tree.rs:267-273
      if (t.is_leaf[i]) t.leaf[i] = t.leaf[i] * c.learning_rate;
    return t;
  }

  // apply_split_to_builder, grower.rs:576-650.
  void apply_split(OTree& t, uint32_t node, const ora_split& s, uint32_t& lt, uint32_t& rt) {
    uint32_t of = B.eff_to_orig[s.feature];
    const Mapper& m = B.mappers[of];
    uint32_t a = t.alloc(), b = t.alloc();
    t.split_index[node] = of;
    t.left[node] = a;
    t.right[node] = b;
    if (!s.is_categorical) {
      float up = (float)m.bounds[s.bin];
      t.thr[node] = next_up(up);
      t.dleft[node] = (uint8_t)s.default_left;
      t.stype[node] = 0;
      lt = a;
      rt = b;
    } else {
      std::vector<uint32_t> words;
      int maxc = -1;
      for (int k = 0; k < 256; k++)
        if ((s.left_cats[k >> 5] >> (k & 31)) & 1u) maxc = k;
      if (maxc >= 0) {
        words.assign((size_t)(maxc >> 5) + 1, 0u);
        for (int k = 0; k <= maxc; k++)
          if ((s.left_cats[k >> 5] >> (k & 31)) & 1u) words[k >> 5] |= 1u << (k & 31);
      }
      t.thr[node] = 0.0f;
      t.dleft[node] = s.default_left ? 0 : 1;
      t.stype[node] = 1;
      t.cat_nodes.push_back({node, words});
      lt = b;  // training-left rows take the inference-right child
      rt = a;
    }
  }

  // ColSampler per tree (column.rs:180-186,266-283): colsample_k effective
  // columns, drawn as the k smallest sampling keys of this tree.
  void sample_columns() {
    uint32_t F = (uint32_t)B.eff_to_orig.size();
    feats.resize(F);
    for (uint32_t i = 0; i < F; i++) feats[i] = i;
    float rate = c.colsample_bytree;
    if (!(std::fabs(rate - 1.0f) < 1e-6f))
      feats = smallest_keys(feats, key_base(c.seed, kStreamCols, round, output), colsample_k(F, rate));
    tree_feats = feats;
    set_level(0);
  }

  // ColSampler::sample_level (column.rs:199-224), called for the root (depth
  // 0, grower.rs:279-280) and for each expansion's children
  // (grower.rs:443-444): ceil(|tree features| x colsample_bylevel) of the
  // tree's features, the smallest sampling keys of (round, output, depth).
  void set_level(uint32_t depth) {
    float rate = c.colsample_bylevel;
    if (std::fabs(rate - 1.0f) < 1e-6f) {
      feats = tree_feats;
      return;
    }
    feats = smallest_keys(tree_feats, key_base(c.seed, kStreamLevel, round, output | (depth << 12)),
                          colsample_k((uint32_t)tree_feats.size(), rate));
  }

  // update_predictions_from_last_tree, grower.rs:208-219.
  void update_preds(float* preds) {
    for (uint32_t node = 0; node < n_part; node++) {
      float v = last_leaf[node];
      if (std::isnan(v)) continue;
      for (uint32_t i = leaf_begin[node]; i < leaf_begin[node] + leaf_count[node]; i++)
        preds[idx[i]] += v;
    }
  }
};

// Tree traversal, tree_view.rs:119-160.
uint32_t traverse(const bst_forest_arrays* f, int64_t base, const float* X, int64_t n, int64_t row) {
  uint32_t node = 0;
  while (!f->is_leaf[base + node]) {
    int64_t k = base + node;
    float x = X[(int64_t)f->split_index[k] * n + row];
    if (std::isnan(x)) {
      node = f->default_left[k] ? f->left[k] : f->right[k];
    } else if (f->split_type[k] == 0) {
      node = x < f->split_threshold[k] ? f->left[k] : f->right[k];
    } else {
      // float_to_category: Rust `f32 as u32` saturates (negatives -> 0).
      uint32_t cat = x <= 0.0f ? 0u : (x >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)x);
      uint32_t size = f->cat_size[k], w = cat >> 5;
      bool right = size != 0 && w < size && ((f->cat_words[f->cat_start[k] + w] >> (cat & 31)) & 1u);
      node = right ? f->right[k] : f->left[k];
    }
  }
  return node;
}

float sigmoid_t(float x) {  // transform.rs:106-115
  float c = x;  // f32::clamp keeps NaN
  if (c < -500.0f) c = -500.0f;
  if (c > 500.0f) c = 500.0f;
  if (c >= 0.0f) return 1.0f / (1.0f + exp32(-c));
  float e = exp32(c);
  return e / (1.0f + e);
}

void transform(int tr, float* out, int K, int64_t n) {
  if (tr == BST_TRANSFORM_SIGMOID) {
    for (int64_t i = 0; i < (int64_t)K * n; i++) out[i] = sigmoid_t(out[i]);
  } else if (tr == BST_TRANSFORM_SOFTMAX) {  // transform.rs:83-103
    for (int64_t i = 0; i < n; i++) {
      float mx = -std::numeric_limits<float>::infinity();
      for (int k = 0; k < K; k++) mx = fmaxf(mx, out[k * n + i]);
      float s = 0.0f;
      for (int k = 0; k < K; k++) {
        float e = exp32(out[k * n + i] - mx);
        out[k * n + i] = e;
        s += e;
      }
      if (s > 0.0f)
        for (int k = 0; k < K; k++) out[k * n + i] /= s;
    }
  }
}

double metric_value(int metric, const float* p, int K, const float* y, const float* w, int64_t n,
                    const Payload& P = Payload()) {
  if (n == 0) return 0.0;
  const double EPS = 1e-15;
  if (metric == BST_METRIC_MAE) {  // regression.rs:55-87
    double s = 0;
    for (int k = 0; k < K; k++) {
      double sa = 0, sw = 0;
      for (int64_t i = 0; i < n; i++) {
        double ae = std::fabs((double)p[k * n + i] - (double)y[i]);
        sa += (double)wt(w, i) * ae;
        sw += (double)wt(w, i);
      }
      s += sw > 0.0 ? sa / sw : 0.0;
    }
    return s / K;
  }
  if (metric == BST_METRIC_MAPE) {  // regression.rs:89-124
    double sa = 0, sw = 0;
    for (int64_t i = 0; i < n; i++) {
      double pp = p[i], l = y[i];
      double ape = std::fabs(pp - l) / std::max(std::fabs(l), EPS);
      sa += (double)wt(w, i) * ape;
      sw += (double)wt(w, i);
    }
    return sw > 0.0 ? (sa / sw) * 100.0 : 0.0;
  }
  if (metric == BST_METRIC_QUANTILE) {  // regression.rs:126-174
    double tl = 0, tw = 0;
    for (int q = 0; q < K && q < (int)P.alphas.size(); q++) {
      double a = (double)P.alphas[q];
      double l = 0, ww = 0;
      for (int64_t i = 0; i < n; i++) {
        double r = (double)y[i] - (double)p[q * n + i];
        double wi = (double)wt(w, i);
        double loss = r >= 0.0 ? a * r : (1.0 - a) * (-r);
        l += wi * loss;
        ww += wi;
      }
      tl += l;
      tw += ww;
    }
    return tw > 0.0 ? tl / tw : 0.0;
  }
  if (metric == BST_METRIC_HUBER) {  // regression.rs:216-250
    double sl = 0, sw = 0, d = P.metric_huber_delta;
    for (int64_t i = 0; i < n; i++) {
      double r = std::fabs((double)p[i] - (double)y[i]);
      double wi = (double)wt(w, i);
      double loss = r <= d ? 0.5 * r * r : d * (r - 0.5 * d);
      sl += wi * loss;
      sw += wi;
    }
    return sw > 0.0 ? sl / sw : 0.0;
  }
  if (metric == BST_METRIC_POISSON_DEVIANCE) {  // regression.rs:176-214
    const double E9 = 1e-9;
    double sd = 0, sw = 0;
    for (int64_t i = 0; i < n; i++) {
      double yy = y[i], mu = std::max((double)p[i], E9), wi = (double)wt(w, i);
      double dev = yy > E9 ? 2.0 * (yy * std::log(yy / mu) - (yy - mu)) : 2.0 * mu;
      sd += wi * dev;
      sw += wi;
    }
    return sw > 0.0 ? sd / sw : 0.0;
  }
  if (metric == BST_METRIC_ACCURACY || metric == BST_METRIC_MARGIN_ACCURACY) {
    // classification.rs:46-122 (margin accuracy: threshold 0 on raw scores)
    const float t = metric == BST_METRIC_ACCURACY ? P.accuracy_threshold : 0.0f;
    double sc = 0, sw = 0;
    for (int64_t i = 0; i < n; i++) {
      float pc = p[i] >= t ? 1.0f : 0.0f;
      double correct = std::fabs(pc - y[i]) < 0.5f ? 1.0 : 0.0;
      sc += (double)wt(w, i) * correct;
      sw += (double)wt(w, i);
    }
    return sw > 0.0 ? sc / sw : 0.0;
  }
  if (metric == BST_METRIC_MULTICLASS_ACCURACY) {  // classification.rs:124-192
    double sc = 0, sw = 0;
    for (int64_t i = 0; i < n; i++) {
      float pc;
      if (K == 1) {
        pc = std::round(p[i]);  // f32::round: half away from zero
      } else {
        // Iterator::max_by keeps the last of equal maxima; NaN compares Less.
        int best = 0;
        for (int k = 1; k < K; k++) {
          float va = p[(int64_t)best * n + i], vb = p[(int64_t)k * n + i];
          bool less = std::isnan(va) || std::isnan(vb) ? true : !(vb < va);
          if (less) best = k;  // cmp(best, k) != Greater -> k becomes the max
        }
        pc = (float)best;
      }
      double correct = std::fabs(pc - y[i]) < 0.5f ? 1.0 : 0.0;
      sc += (double)wt(w, i) * correct;
      sw += (double)wt(w, i);
    }
    return sw > 0.0 ? sc / sw : 0.0;
  }
  if (metric == BST_METRIC_RMSE) {  // regression.rs:17-40
    double s = 0;
    for (int k = 0; k < K; k++) {
      double ss = 0, sw = 0;
      for (int64_t i = 0; i < n; i++) {
        double d = (double)p[k * n + i] - (double)y[i];
        ss += (double)wt(w, i) * d * d;
        sw += (double)wt(w, i);
      }
      s += sw > 0.0 ? std::sqrt(ss / sw) : 0.0;
    }
    return s / K;
  }
  if (metric == BST_METRIC_LOGLOSS) {  // classification.rs:13-40
    double sl = 0, sw = 0;
    for (int64_t i = 0; i < n; i++) {
      double q = std::min(std::max((double)p[i], EPS), 1.0 - EPS), l = y[i];
      sl += (double)wt(w, i) * (-(l * std::log(q) + (1.0 - l) * std::log(1.0 - q)));
      sw += (double)wt(w, i);
    }
    return sw > 0 ? sl / sw : 0.0;
  }
  if (metric == BST_METRIC_MLOGLOSS) {  // classification.rs:332-360
    double sl = 0, sw = 0;
    for (int64_t i = 0; i < n; i++) {
      size_t c = (size_t)std::llround((double)y[i]);
      double q = std::min(std::max((double)p[c * n + i], EPS), 1.0 - EPS);
      sl += (double)wt(w, i) * (-std::log(q));
      sw += (double)wt(w, i);
    }
    return sw > 0 ? sl / sw : 0.0;
  }
  return std::numeric_limits<double>::quiet_NaN();
}

int default_metric(int obj) {  // default_metric_for_objective, metrics/mod.rs:213-232
  if (obj == BST_OBJ_LAMBDARANK) return BST_METRIC_NONE;  // no ranking metric in the reference
  if (obj == BST_OBJ_PINBALL) return BST_METRIC_QUANTILE;
  if (obj == BST_OBJ_HINGE) return BST_METRIC_MARGIN_ACCURACY;
  return obj == BST_OBJ_LOGISTIC ? BST_METRIC_LOGLOSS
                                 : (obj == BST_OBJ_SOFTMAX ? BST_METRIC_MLOGLOSS : BST_METRIC_RMSE);
}

// Metric::higher_is_better, metrics/mod.rs:286-301.
bool higher_is_better(int metric) {
  return metric == BST_METRIC_ACCURACY || metric == BST_METRIC_MARGIN_ACCURACY ||
         metric == BST_METRIC_MULTICLASS_ACCURACY;
}
// PredictionKind::Margin metrics read untransformed scores, eval.rs:150-184.
bool metric_on_margins(int metric) { return metric == BST_METRIC_MARGIN_ACCURACY; }

int transform_of(int obj) {
  return obj == BST_OBJ_LOGISTIC ? BST_TRANSFORM_SIGMOID
                                 : (obj == BST_OBJ_SOFTMAX ? BST_TRANSFORM_SOFTMAX
                                                           : BST_TRANSFORM_IDENTITY);
}

void export_sizes(const ora_model& m, bst_model_info* info) {
  int64_t nodes = 0, words = 0;
  for (auto& t : m.trees) {
    nodes += (int64_t)t.split_index.size();
    for (auto& cn : t.cat_nodes) words += (int64_t)cn.second.size();
  }
  info->n_trees = (int64_t)m.trees.size();
  info->n_groups = m.n_groups;
  info->n_features = m.n_features;
  info->total_nodes = nodes;
  info->total_cat_words = words;
  info->output_transform = m.transform;
  info->objective = m.objective;
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

const char* ora_last_error(void) { return g_err.c_str(); }

int ora_num_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

void ora_set_exp_mode(int mode) { g_exp_mode = mode; }

void ora_set_num_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

int ora_bin(const float* X, int64_t n, int32_t F, const uint8_t* is_cat,
            const bst_binning_config* cfgp, ora_binned** out) {
  if (!X || !out) return fail(BST_E_INVALID, "null argument");
  if (n <= 0 || F <= 0) return fail(BST_E_SHAPE, "empty dataset");
  bst_binning_config cfg = cfgp ? *cfgp
                                : bst_binning_config{256, 0.9f, 1, 0, 200000};
  auto* B = new ora_binned();
  B->n = n;
  B->F = F;
  B->mappers.resize(F);
  std::vector<Analysis> an(F);
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t f = 0; f < F; f++) an[f] = analyze(X + (int64_t)f * n, n, is_cat && is_cat[f], cfg);
  std::vector<uint32_t> sample = binning_sample(n, cfg.sample_cnt);
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t f = 0; f < F; f++) {
    const float* col = X + (int64_t)f * n;
    Mapper m;
    if (an[f].numeric) {
      std::vector<float> vals;
      if (sample.empty())
        vals.assign(col, col + n);
      else {
        vals.resize(sample.size());
        for (size_t i = 0; i < sample.size(); i++) vals[i] = col[sample[i]];
      }
      m = bin_numeric(vals, cfg.max_bins, an[f].n_nan > 0);
    } else {
      std::unordered_set<int32_t> seen;
      for (int64_t i = 0; i < n; i++)
        if (std::isfinite(col[i])) {
          double v = col[i];
          int32_t c = v >= 2147483647.0 ? 2147483647
                                        : (v <= -2147483648.0 ? (int32_t)-2147483648LL : (int32_t)v);
          seen.insert(c);
        }
      m.categorical = true;
      m.cats.assign(seen.begin(), seen.end());
      std::sort(m.cats.begin(), m.cats.end());
      if (m.cats.empty()) m.cats.push_back(0);
      m.n_bins = (uint32_t)m.cats.size();
    }
    m.trivial = an[f].trivial;
    B->mappers[f] = m;
  }
  // Optional EFB (dataset.rs:229-236); the effective columns are the bundles
  // first, then the other non-trivial features (feature_views, dataset.rs:460-520).
  std::vector<uint8_t> bundled((size_t)F, 0);
  uint32_t off = 0;
  if (cfg.enable_bundling) {
    for (const auto& fs : create_bundle_plan(an, B->mappers, X, n, cfg)) {
      // build_bundle, dataset.rs:898-952: codes offset + bin, offsets from 1.
      ora_binned::Bundle bd;
      bd.features = fs;
      uint32_t o = 1;
      for (uint32_t f : fs) {
        bd.offsets.push_back(o);
        o += B->mappers[f].n_bins;
        bundled[f] = 1;
      }
      bd.total_bins = o;
      B->bundles.push_back(bd);
      B->eff_to_orig.push_back(fs[0]);  // decode_split_to_original: the bundle's first feature
      B->eff_nbins.push_back(o);
      B->eff_offset.push_back(off);
      B->eff_cat.push_back(1);      // bundles are categorical columns ...
      B->eff_missing.push_back(1);  // ... whose code 0 (all defaults) is the missing bin
      B->eff_default.push_back(0);
      off += o;
    }
  }
  const size_t n_bundles = B->bundles.size();
  for (int32_t f = 0; f < F; f++) {
    if (B->mappers[f].trivial || bundled[f]) continue;
    B->eff_to_orig.push_back((uint32_t)f);
    B->eff_nbins.push_back(B->mappers[f].n_bins);
    B->eff_offset.push_back(off);
    B->eff_cat.push_back(B->mappers[f].categorical);
    B->eff_missing.push_back(B->mappers[f].missing != 0);
    B->eff_default.push_back(B->mappers[f].default_bin);
    off += B->mappers[f].n_bins;
  }
  B->total_bins = off;
  size_t E = B->eff_to_orig.size();
  B->bins.resize(E);
  for (size_t e = 0; e < n_bundles; e++) {
    const auto& bd = B->bundles[e];
    auto& v = B->bins[e];
    v.assign((size_t)n, 0);
    for (size_t k = 0; k < bd.features.size(); k++) {
      const float* col = X + (int64_t)bd.features[k] * n;
      for (int64_t i = 0; i < n; i++)
        if (v[i] == 0 && col[i] != 0.0f && !std::isnan(col[i]))
          v[i] = (uint16_t)(bd.offsets[k] + value_to_bin(B->mappers[bd.features[k]], (double)col[i]));
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t e = n_bundles; e < E; e++) {
    const Mapper& m = B->mappers[B->eff_to_orig[e]];
    const float* col = X + (int64_t)B->eff_to_orig[e] * n;
    auto& v = B->bins[e];
    v.resize(n);
    for (int64_t i = 0; i < n; i++) v[i] = (uint16_t)value_to_bin(m, (double)col[i]);
  }
  if (E == 0) {
    delete B;
    return fail(BST_E_SHAPE, "all features are trivial");
  }
  *out = B;
  return BST_OK;
}

void ora_binned_free(ora_binned* b) { delete b; }

int ora_feature_info(const ora_binned* b, int32_t f, uint32_t* nb, int32_t* mt, uint32_t* db,
                     int32_t* cat, int32_t* triv) {
  if (!b || f < 0 || f >= b->F) return fail(BST_E_INVALID, "bad feature");
  const Mapper& m = b->mappers[f];
  if (nb) *nb = m.n_bins;
  if (mt) *mt = m.missing;
  if (db) *db = m.default_bin;
  if (cat) *cat = m.categorical;
  if (triv) *triv = m.trivial;
  return BST_OK;
}

int ora_bin_bounds(const ora_binned* b, int32_t f, double* out, int64_t cap) {
  if (!b || f < 0 || f >= b->F) return fail(BST_E_INVALID, "bad feature");
  const Mapper& m = b->mappers[f];
  int64_t k = 0;
  if (m.categorical)
    for (int32_t c : m.cats) {
      if (k < cap) out[k] = c;
      k++;
    }
  else
    for (double v : m.bounds) {
      if (k < cap) out[k] = v;
      k++;
    }
  return (int)k;
}

int ora_get_bins(const ora_binned* b, uint16_t* out) {
  if (!b || !out) return fail(BST_E_INVALID, "null");
  std::fill(out, out + (size_t)b->F * b->n, 0);
  // A bundled feature: its own codes decoded (BundleStorage::decode), any
  // other code stands for the value 0.0.
  for (size_t e = 0; e < b->bundles.size(); e++) {
    const auto& bd = b->bundles[e];
    for (size_t k = 0; k < bd.features.size(); k++) {
      const Mapper& m = b->mappers[bd.features[k]];
      const uint32_t lo = bd.offsets[k], hi = lo + m.n_bins;
      const uint16_t zero_bin = (uint16_t)value_to_bin(m, 0.0);
      uint16_t* o = out + (size_t)bd.features[k] * b->n;
      for (int64_t r = 0; r < b->n; r++) {
        const uint32_t c = b->bins[e][r];
        o[r] = c >= lo && c < hi ? (uint16_t)(c - lo) : zero_bin;
      }
    }
  }
  for (size_t e = b->bundles.size(); e < b->eff_to_orig.size(); e++)
    std::copy(b->bins[e].begin(), b->bins[e].end(), out + (size_t)b->eff_to_orig[e] * b->n);
  return BST_OK;
}

int ora_hist_layout(const ora_binned* b, int32_t* ne, uint32_t* e2o, uint32_t* enb, uint32_t* eoff) {
  if (!b) return fail(BST_E_INVALID, "null");
  size_t E = b->eff_to_orig.size();
  if (ne) *ne = (int32_t)E;
  for (size_t e = 0; e < E; e++) {
    if (e2o) e2o[e] = b->eff_to_orig[e];
    if (enb) enb[e] = b->eff_nbins[e];
    if (eoff) eoff[e] = b->eff_offset[e];
  }
  return (int)b->total_bins;
}

uint32_t ora_value_to_bin(const ora_binned* b, int32_t f, double v) {
  if (!b || f < 0 || f >= b->F) return 0xFFFFFFFFu;
  return value_to_bin(b->mappers[f], v);
}

float ora_leaf_weight(const bst_gbdt_config* cfg, double g, double h) {
  return leaf_weight(*cfg, g, h);
}

int ora_gradients(int32_t obj, int32_t K, const float* p, const float* y, const float* w, int64_t n,
                  float* g, float* h) {
  if (!p || !y || !g || !h) return fail(BST_E_INVALID, "null");
  if (obj == BST_OBJ_LAMBDARANK) return fail(BST_E_CONFIG, "use ora_lambdarank_gradients");
  gradients(obj, K, p, y, w, n, g, h);
  return BST_OK;
}

int ora_lambdarank_gradients(const float* p, const float* y, const float* w, int64_t n,
                             const int32_t* sizes, int64_t n_groups, float* g, float* h) {
  if (!p || !y || !g || !h || !sizes) return fail(BST_E_INVALID, "null");
  int64_t tot = 0;
  for (int64_t q = 0; q < n_groups; q++) {
    if (sizes[q] < 1) return fail(BST_E_CONFIG, "empty query group");
    tot += sizes[q];
  }
  if (tot != n) return fail(BST_E_SHAPE, "query group sizes must sum to the row count");
  lambdarank_gradients(p, y, w, n, sizes, n_groups, g, h);
  return BST_OK;
}

int ora_base_score(int32_t obj, int32_t K, const float* y, const float* w, int64_t n, float* out) {
  auto b = base_score(obj, K, y, w, n);
  for (size_t i = 0; i < b.size(); i++) out[i] = b[i];
  return BST_OK;
}

int ora_gradients_cfg(const bst_gbdt_config* cfg, const float* p, const float* y, const float* w,
                      int64_t n, float* g, float* h) {
  if (!cfg || !p || !y || !g || !h) return fail(BST_E_INVALID, "null");
  if (cfg->objective == BST_OBJ_LAMBDARANK) return fail(BST_E_CONFIG, "use ora_lambdarank_gradients");
  gradients(cfg->objective, n_outputs_of(*cfg), p, y, w, n, g, h, payload_of(cfg));
  return BST_OK;
}

int ora_base_score_cfg(const bst_gbdt_config* cfg, const float* y, const float* w, int64_t n,
                       float* out) {
  if (!cfg || !out) return fail(BST_E_INVALID, "null");
  auto b = base_score(cfg->objective, n_outputs_of(*cfg), y, w, n, payload_of(cfg));
  for (size_t i = 0; i < b.size(); i++) out[i] = b[i];
  return BST_OK;
}

float ora_weighted_quantile(const float* v, const float* w, int64_t n, float alpha) {
  return weighted_quantile(v, w, n, alpha);
}

float ora_quantile_indexed(const uint32_t* ind, uint32_t n, const float* values, const float* w,
                           float alpha) {
  auto value = [&](uint32_t i) { return values[i]; };
  return w ? weighted_quantile_indexed(ind, n, value, w, alpha)
           : unweighted_quantile_indexed(ind, n, value, alpha);
}

int ora_histogram(const ora_binned* b, const float* g, const float* h, const uint32_t* rows,
                  int64_t nr, double* out) {
  if (!b || !g || !h || !out) return fail(BST_E_INVALID, "null");
  build_hist(*b, g, h, rows, rows ? nr : b->n, out);
  return BST_OK;
}

int ora_find_split(const ora_binned* b, const double* hist, double pg, double ph, uint32_t pc,
                   const bst_gbdt_config* cfg, int32_t nt, const uint32_t* feats, int32_t nf,
                   ora_split* out) {
  if (!b || !hist || !cfg || !out) return fail(BST_E_INVALID, "null");
  std::vector<uint32_t> fs;
  if (feats)
    fs.assign(feats, feats + nf);
  else
    for (uint32_t i = 0; i < b->eff_to_orig.size(); i++) fs.push_back(i);
  *out = find_split(*b, hist, pg, ph, pc, *cfg, nt, fs);
  return BST_OK;
}

int ora_partition(const ora_binned* b, uint32_t* rows, int64_t n, const ora_split* s, int64_t* lc) {
  if (!b || !rows || !s) return fail(BST_E_INVALID, "null");
  std::vector<uint32_t> scratch;
  int64_t l = stable_partition(*b, rows, n, *s, scratch);
  if (lc) *lc = l;
  return BST_OK;
}

int ora_train(const ora_binned* b, const float* X, const float* y, int64_t ny, const float* w,
              const float* VX, const float* vy, int64_t nv, const bst_gbdt_config* cfgp,
              const bst_ext_config* extp, ora_model** out) {
  if (!b || !y || !cfgp || !out) return fail(BST_E_INVALID, "null");
  const bst_gbdt_config& c = *cfgp;
  bst_ext_config ext;
  ext.schedule = 0;
  ext.n_threads = 0;
  ext.goss_top_rate = 0;
  ext.goss_other_rate = 0;
  ext.query_group_sizes = nullptr;
  ext.n_query_groups = 0;
  ext.autotune = 0;
  if (extp) ext = *extp;
  int64_t n = b->n;
  if (ny < n) return fail(BST_E_TARGETS, "fewer targets than rows");
  const int K = n_outputs_of(c);
  const Payload P = payload_of(&c);
  if (K < 1) return fail(BST_E_CONFIG, "the objective has no outputs");
  if (c.objective == BST_OBJ_LAMBDARANK) {
    if (!ext.query_group_sizes || ext.n_query_groups < 1)
      return fail(BST_E_CONFIG, "LambdaRank needs query groups");
    int64_t tot = 0;
    for (int64_t q = 0; q < ext.n_query_groups; q++) {
      if (ext.query_group_sizes[q] < 1) return fail(BST_E_CONFIG, "empty query group");
      tot += ext.query_group_sizes[q];
    }
    if (tot != n) return fail(BST_E_SHAPE, "query group sizes must sum to the row count");
  }
  auto t0 = std::chrono::steady_clock::now();
  auto* M = new ora_model();
  M->objective = c.objective;
  M->n_groups = K;
  M->n_features = b->F;
  M->transform = transform_of(c.objective);
  M->base = base_score(c.objective, K, y, w, n, P);
  std::vector<float> preds((size_t)K * n), g((size_t)K * n), h((size_t)K * n);
  for (int k = 0; k < K; k++) std::fill(preds.begin() + (size_t)k * n, preds.begin() + (size_t)(k + 1) * n, M->base[k]);
  int metric = c.metric == BST_METRIC_DEFAULT ? default_metric(c.objective) : c.metric;
  std::vector<float> vpreds;
  if (VX && vy && nv > 0) {
    vpreds.resize((size_t)K * nv);
    for (int k = 0; k < K; k++) std::fill(vpreds.begin() + (size_t)k * nv, vpreds.begin() + (size_t)(k + 1) * nv, M->base[k]);
  }
  Grower gr(*b, c, ext.n_threads);
  bool goss = ext.goss_top_rate > 0.0f;
  bool uniform = !goss && !(std::fabs(c.subsample - 1.0f) < 1e-6f);
  size_t warmup = goss ? goss_warmup(c.learning_rate) : 0;
  int es_patience = (int)c.early_stopping_rounds;
  const bool hib = higher_is_better(metric);
  // Leaf renewal (trainer.rs:265-291): quantile objectives replace each
  // leaf's Newton value by the alpha-quantile of its rows' residuals.
  const bool renew = c.objective == BST_OBJ_ABSOLUTE || c.objective == BST_OBJ_PINBALL;
  double best_val = 0;
  bool have_best = false;
  size_t best_round = 0, cur_round = 0, best_n_trees = 0;
  for (uint32_t round = 0; round < c.n_trees; round++) {
    if (c.objective == BST_OBJ_LAMBDARANK)
      lambdarank_gradients(preds.data(), y, w, n, ext.query_group_sizes, ext.n_query_groups,
                           g.data(), h.data());
    else
      gradients(c.objective, K, preds.data(), y, w, n, g.data(), h.data(), P);
    for (int k = 0; k < K; k++) {
      float* gk = g.data() + (size_t)k * n;
      float* hk = h.data() + (size_t)k * n;
      std::vector<uint32_t> sampled;
      bool use_sample = false;
      if (uniform) {  // sample_uniform, row.rs:170-208
        sampled = sample_uniform(gk, hk, n, c.subsample, key_base(c.seed, kStreamRows, round, (uint32_t)k));
        use_sample = true;
      } else if (goss && round >= warmup) {  // sample_goss, row.rs:211-270
        sampled = sample_goss(gk, hk, n, ext.goss_top_rate, ext.goss_other_rate,
                              key_base(c.seed, kStreamGoss, round, (uint32_t)k));
        use_sample = true;
      }
      gr.round = round;
      gr.output = (uint32_t)k;
      OTree t = gr.grow(gk, hk, use_sample ? &sampled : nullptr);
      M->trees.push_back(std::move(t));
      M->groups.push_back((uint32_t)k);
      OTree& tt = M->trees.back();
      if (renew) {
        const float alpha = c.objective == BST_OBJ_PINBALL ? P.alphas[k] : 0.5f;
        const float* pk = preds.data() + (size_t)k * n;
        auto residual = [&](uint32_t i) { return y[i] - pk[i]; };
        for (auto& lm : gr.leaf_map) {
          const uint32_t tn = lm.first, pn = lm.second;
          const uint32_t cnt = gr.leaf_count[pn];
          if (cnt == 0) continue;
          const uint32_t* ind = gr.idx.data() + gr.leaf_begin[pn];
          float v = w ? weighted_quantile_indexed(ind, cnt, residual, w, alpha)
                      : unweighted_quantile_indexed(ind, cnt, residual, alpha);
          if (std::isnan(v)) continue;
          const float sv = v * c.learning_rate;
          tt.leaf[tn] = sv;
          gr.last_leaf[pn] = sv;
        }
      }
      // Prediction updates: partition path unless sampled (trainer.rs:331-339).
      if (!use_sample) {
        gr.update_preds(preds.data() + (size_t)k * n);
      } else {
        // Traverse raw features of the training rows.
        bst_model_info info;
        std::vector<int64_t> off{0, (int64_t)tt.split_index.size()};
        std::vector<uint32_t> cs(tt.split_index.size(), 0), csz(tt.split_index.size(), 0), words;
        for (auto& cn : tt.cat_nodes) {
          cs[cn.first] = (uint32_t)words.size();
          csz[cn.first] = (uint32_t)cn.second.size();
          words.insert(words.end(), cn.second.begin(), cn.second.end());
        }
        bst_forest_arrays fa{};
        uint32_t grp = 0;
        fa.tree_offset = off.data();
        fa.tree_group = &grp;
        fa.split_index = tt.split_index.data();
        fa.split_threshold = tt.thr.data();
        fa.left = tt.left.data();
        fa.right = tt.right.data();
        fa.default_left = tt.dleft.data();
        fa.is_leaf = tt.is_leaf.data();
        fa.split_type = tt.stype.data();
        fa.leaf_value = tt.leaf.data();
        fa.cat_start = cs.data();
        fa.cat_size = csz.data();
        fa.cat_words = words.data();
        (void)info;
        if (!X) {
          delete M;
          return fail(BST_E_INVALID, "row sampling needs raw features");
        }
        float* pk = preds.data() + (size_t)k * n;
        for (int64_t i = 0; i < n; i++) pk[i] += tt.leaf[traverse(&fa, 0, X, n, i)];
      }
      if (!vpreds.empty()) {
        std::vector<int64_t> off{0, (int64_t)tt.split_index.size()};
        std::vector<uint32_t> cs(tt.split_index.size(), 0), csz(tt.split_index.size(), 0), words;
        for (auto& cn : tt.cat_nodes) {
          cs[cn.first] = (uint32_t)words.size();
          csz[cn.first] = (uint32_t)cn.second.size();
          words.insert(words.end(), cn.second.begin(), cn.second.end());
        }
        bst_forest_arrays fa{};
        uint32_t grp = 0;
        fa.tree_offset = off.data();
        fa.tree_group = &grp;
        fa.split_index = tt.split_index.data();
        fa.split_threshold = tt.thr.data();
        fa.left = tt.left.data();
        fa.right = tt.right.data();
        fa.default_left = tt.dleft.data();
        fa.is_leaf = tt.is_leaf.data();
        fa.split_type = tt.stype.data();
        fa.leaf_value = tt.leaf.data();
        fa.cat_start = cs.data();
        fa.cat_size = csz.data();
        fa.cat_words = words.data();
        float* vk = vpreds.data() + (size_t)k * nv;
        for (int64_t i = 0; i < nv; i++) vk[i] += tt.leaf[traverse(&fa, 0, VX, nv, i)];
      }
    }
    // evaluate_round (eval.rs:218-260) on transformed predictions.
    double value = std::numeric_limits<double>::quiet_NaN();
    if (metric != BST_METRIC_NONE) {
      const int tr = metric_on_margins(metric) ? BST_TRANSFORM_IDENTITY : M->transform;
      if (!vpreds.empty()) {
        std::vector<float> tp = vpreds;
        transform(tr, tp.data(), K, nv);
        value = metric_value(metric, tp.data(), K, vy, nullptr, nv, P);
      } else {
        std::vector<float> tp = preds;
        transform(tr, tp.data(), K, n);
        value = metric_value(metric, tp.data(), K, y, w, n, P);
      }
    }
    M->metric.push_back(value);
    if (es_patience > 0) {  // EarlyStopping::update, callback.rs:78-118
      bool imp = !have_best || (hib ? value > best_val : value < best_val);
      if (imp) {
        best_val = value;
        have_best = true;
        best_round = cur_round;
        best_n_trees = M->trees.size();
      }
      cur_round++;
      if (cur_round - best_round > (size_t)es_patience) break;
    }
  }
  if (es_patience > 0 && best_n_trees > 0 && best_n_trees < M->trees.size()) {
    M->trees.resize(best_n_trees);
    M->groups.resize(best_n_trees);
  }
  M->hist_rows = gr.hist_rows;
  M->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = M;
  return BST_OK;
}

int ora_model_info(const ora_model* m, bst_model_info* info) {
  if (!m || !info) return fail(BST_E_INVALID, "null");
  export_sizes(*m, info);
  return BST_OK;
}

int ora_model_export(const ora_model* m, bst_forest_arrays* o) {
  if (!m || !o) return fail(BST_E_INVALID, "null");
  for (int k = 0; k < m->n_groups; k++) o->base_score[k] = m->base[k];
  int64_t off = 0, woff = 0;
  for (size_t t = 0; t < m->trees.size(); t++) {
    const OTree& tr = m->trees[t];
    o->tree_offset[t] = off;
    o->tree_group[t] = m->groups[t];
    size_t nn = tr.split_index.size();
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> cn = tr.cat_nodes;
    std::sort(cn.begin(), cn.end(), [](auto& a, auto& b) { return a.first < b.first; });
    for (size_t i = 0; i < nn; i++) {
      o->split_index[off + i] = tr.split_index[i];
      o->split_threshold[off + i] = tr.thr[i];
      o->left[off + i] = tr.left[i];
      o->right[off + i] = tr.right[i];
      o->default_left[off + i] = tr.dleft[i];
      o->is_leaf[off + i] = tr.is_leaf[i];
      o->split_type[off + i] = tr.stype[i];
      o->leaf_value[off + i] = tr.is_leaf[i] ? tr.leaf[i] : 0.0f;
      o->gain[off + i] = tr.gain[i];
      o->cover[off + i] = tr.cover[i];
      o->cat_start[off + i] = 0;
      o->cat_size[off + i] = 0;
    }
    for (auto& c : cn) {
      o->cat_start[off + c.first] = (uint32_t)woff;
      o->cat_size[off + c.first] = (uint32_t)c.second.size();
      for (uint32_t wv : c.second) o->cat_words[woff++] = wv;
    }
    off += (int64_t)nn;
  }
  o->tree_offset[m->trees.size()] = off;
  return BST_OK;
}

int ora_model_metric(const ora_model* m, double* out, int64_t cap, int64_t* nr) {
  if (!m) return fail(BST_E_INVALID, "null");
  for (int64_t i = 0; i < (int64_t)m->metric.size() && i < cap; i++) out[i] = m->metric[i];
  if (nr) *nr = (int64_t)m->metric.size();
  return BST_OK;
}

int ora_model_stats(const ora_model* m, int64_t* hr, double* ms) {
  if (!m) return fail(BST_E_INVALID, "null");
  if (hr) *hr = m->hist_rows;
  if (ms) *ms = m->ms_total;
  return BST_OK;
}

void ora_model_free(ora_model* m) { delete m; }

int ora_predict(const bst_model_info* info, const bst_forest_arrays* f, const float* X, int64_t n,
                int32_t F, int32_t raw, float* out, uint32_t* leaf_idx) {
  if (!info || !f || !X || !out) return fail(BST_E_INVALID, "null");
  if (F < info->n_features) return fail(BST_E_SHAPE, "too few features");
  int K = (int)info->n_groups;
  // predict_into, predictor.rs:197-302: base score, then f32 += leaf in
  // forest order per row.
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    for (int k = 0; k < K; k++) out[k * n + i] = f->base_score[k];
    for (int64_t t = 0; t < info->n_trees; t++) {
      int64_t base = f->tree_offset[t];
      uint32_t leaf = traverse(f, base, X, n, i);
      out[(int64_t)f->tree_group[t] * n + i] += f->leaf_value[base + leaf];
      if (leaf_idx) leaf_idx[t * n + i] = leaf;
    }
  }
  if (!raw) transform(info->output_transform, out, K, n);
  return BST_OK;
}

int ora_partition_predict(const ora_binned* b, const bst_model_info* info,
                          const bst_forest_arrays* f, const uint8_t* by_threshold,
                          const float* X, float* out) {
  if (!b || !info || !f || !out) return fail(BST_E_INVALID, "null");
  const int64_t n = b->n;
  const int K = (int)info->n_groups;
  std::vector<int32_t> orig_to_eff(b->F, -1);
  for (size_t e = 0; e < b->eff_to_orig.size(); e++) orig_to_eff[b->eff_to_orig[e]] = (int32_t)e;
  // Per node: the split bin a numeric threshold came from (the bin b with
  // next_up((f32)bound_b) == threshold, grower.rs:543-556,603-604).
  std::vector<int32_t> split_bin((size_t)info->total_nodes, -1);
  for (int64_t t = 0; t < info->n_trees; t++) {
    if (by_threshold && by_threshold[t]) continue;
    for (int64_t k = f->tree_offset[t]; k < f->tree_offset[t + 1]; k++) {
      if (f->is_leaf[k] || f->split_type[k]) continue;
      const uint32_t of = f->split_index[k];
      if (of >= (uint32_t)b->F || orig_to_eff[of] < 0) return fail(BST_E_SHAPE, "split on a feature without bins");
      const Mapper& m = b->mappers[of];
      const size_t nb = m.bounds.size() - (m.missing == 2 ? 1 : 0);
      int32_t found = -1;
      for (size_t i = 0; i < nb; i++)
        if (next_up((float)m.bounds[i]) == f->split_threshold[k]) {
          if (found >= 0) return fail(BST_E_SHAPE, "two bins give the same threshold");
          found = (int32_t)i;
        }
      if (found < 0) return fail(BST_E_SHAPE, "threshold is not a bin boundary");
      split_bin[k] = found;
    }
  }
  const bool any_thr = by_threshold && std::any_of(by_threshold, by_threshold + info->n_trees,
                                                   [](uint8_t v) { return v != 0; });
  if (any_thr && !X) return fail(BST_E_INVALID, "threshold traversal needs the raw features");
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    for (int k = 0; k < K; k++) out[k * n + r] = f->base_score[k];
    for (int64_t t = 0; t < info->n_trees; t++) {
      const int64_t base = f->tree_offset[t];
      uint32_t node = 0;
      if (by_threshold && by_threshold[t]) {
        node = traverse(f, base, X, n, r);
      } else {
        // RowPartitioner::split's rule (partition.rs:321-337) on the row's bins.
        while (!f->is_leaf[base + node]) {
          const int64_t k = base + node;
          const uint32_t of = f->split_index[k];
          const int32_t e = orig_to_eff[of];
          const uint32_t bin = b->bins[e][r];
          const bool missing = b->eff_missing[e] != 0 && bin == b->eff_default[e];
          bool go_left;
          if (!f->split_type[k]) {
            go_left = missing ? f->default_left[k] != 0 : bin <= (uint32_t)split_bin[k];
          } else {
            // Categorical: training-left rows take the inference-right child
            // and the stored default direction is inverted (grower.rs:612-648).
            const uint32_t size = f->cat_size[k], w = bin >> 5;
            const bool in_set = size != 0 && w < size &&
                                ((f->cat_words[f->cat_start[k] + w] >> (bin & 31)) & 1u);
            const bool train_left = missing ? f->default_left[k] == 0 : in_set;
            go_left = !train_left;
          }
          node = go_left ? f->left[k] : f->right[k];
        }
      }
      out[(int64_t)f->tree_group[t] * n + r] += f->leaf_value[base + node];
    }
  }
  return BST_OK;
}

void ora_splitmix64_stream(uint64_t state, int64_t count, uint64_t* out) {
  for (int64_t i = 0; i < count; i++) out[i] = splitmix64_next(state);
}

void ora_xoshiro256pp(const uint64_t* state, uint64_t seed, int64_t count, uint64_t* out) {
  Xoshiro256PlusPlus r = Xoshiro256PlusPlus::seed_from_u64(seed);
  if (state) std::copy(state, state + 4, r.s);
  for (int64_t i = 0; i < count; i++) out[i] = r.next_u64();
}

int ora_sample_inclusive(const uint64_t* draws, int64_t n_draws, uint64_t low, uint64_t high,
                         uint64_t* value, int64_t* used) {
  if (!draws || !value || low > high) return fail(BST_E_INVALID, "bad argument");
  int64_t k = 0;
  bool ran_out = false;
  const uint64_t v = sample_single_inclusive(low, high, [&] {
    if (k >= n_draws) {
      ran_out = true;
      return (uint64_t)0;  // the low word of 0 * range is 0: accepted
    }
    return draws[k++];
  });
  if (ran_out) return fail(BST_E_INVALID, "every draw was rejected");
  *value = v;
  if (used) *used = k;
  return BST_OK;
}

int64_t ora_efb_sample_rows(int64_t n, int64_t max_samples, uint64_t seed, int64_t* out,
                            int64_t cap) {
  const std::vector<size_t> r = efb_sample_rows((size_t)n, (size_t)max_samples, seed);
  for (size_t i = 0; i < r.size() && (int64_t)i < cap; i++) out[i] = (int64_t)r[i];
  return (int64_t)r.size();
}

int32_t ora_bundles(const ora_binned* b, int32_t* sizes, uint32_t* features, int64_t cap) {
  if (!b) return fail(BST_E_INVALID, "null");
  int64_t k = 0;
  for (size_t i = 0; i < b->bundles.size(); i++) {
    if (sizes) sizes[i] = (int32_t)b->bundles[i].features.size();
    for (uint32_t f : b->bundles[i].features) {
      if (features && k < cap) features[k] = f;
      k++;
    }
  }
  return (int32_t)b->bundles.size();
}

int64_t ora_binning_sample(int64_t n, uint64_t cnt, uint32_t* out) {
  std::vector<uint32_t> idx = binning_sample(n, cnt);
  if (out) std::copy(idx.begin(), idx.end(), out);
  return (int64_t)idx.size();
}

int ora_sample_rows(const float* g, const float* h, int64_t n, const bst_gbdt_config* cfg,
                    const bst_ext_config* ext, uint32_t round, uint32_t output, float* g_out,
                    float* h_out, uint32_t* rows_out, int64_t* n_rows_out) {
  if (!g || !h || !cfg || !g_out || !h_out || !n_rows_out || n < 1) return fail(BST_E_INVALID, "bad argument");
  std::copy(g, g + n, g_out);
  std::copy(h, h + n, h_out);
  const bool goss = ext && ext->goss_top_rate > 0.0f;
  const bool uniform = !goss && !(std::fabs(cfg->subsample - 1.0f) < 1e-6f);
  std::vector<uint32_t> rows;
  if (uniform) {
    rows = sample_uniform(g_out, h_out, n, cfg->subsample, key_base(cfg->seed, kStreamRows, round, output));
  } else if (goss && round >= goss_warmup(cfg->learning_rate)) {
    if (n < 2) return fail(BST_E_CONFIG, "GOSS needs at least two rows");
    rows = sample_goss(g_out, h_out, n, ext->goss_top_rate, ext->goss_other_rate,
                       key_base(cfg->seed, kStreamGoss, round, output));
  } else {
    *n_rows_out = -1;  // this round trains on every row
    return BST_OK;
  }
  if (rows_out) std::copy(rows.begin(), rows.end(), rows_out);
  *n_rows_out = (int64_t)rows.size();
  return BST_OK;
}

int ora_sample_cols(uint32_t n_feat, const bst_gbdt_config* cfg, uint32_t round, uint32_t output,
                    uint32_t* out, int32_t* k) {
  if (!cfg || !out || !k || n_feat < 1) return fail(BST_E_INVALID, "bad argument");
  std::vector<uint32_t> feats(n_feat);
  for (uint32_t i = 0; i < n_feat; i++) feats[i] = i;
  if (!(std::fabs(cfg->colsample_bytree - 1.0f) < 1e-6f))
    feats = smallest_keys(feats, key_base(cfg->seed, kStreamCols, round, output),
                          colsample_k(n_feat, cfg->colsample_bytree));
  std::copy(feats.begin(), feats.end(), out);
  *k = (int32_t)feats.size();
  return BST_OK;
}

int ora_sample_level(const uint32_t* tree_feats, uint32_t n_tree_feats, const bst_gbdt_config* cfg,
                     uint32_t round, uint32_t output, uint32_t depth, uint32_t* out, int32_t* k) {
  if (!tree_feats || !cfg || !out || !k) return fail(BST_E_INVALID, "null");
  std::vector<uint32_t> t(tree_feats, tree_feats + n_tree_feats), f;
  if (std::fabs(cfg->colsample_bylevel - 1.0f) < 1e-6f)
    f = t;
  else
    f = smallest_keys(t, key_base(cfg->seed, kStreamLevel, round, output | (depth << 12)),
                      colsample_k(n_tree_feats, cfg->colsample_bylevel));
  std::copy(f.begin(), f.end(), out);
  *k = (int32_t)f.size();
  return BST_OK;
}

double ora_metric(int32_t metric, const float* p, int32_t K, const float* y, const float* w,
                  int64_t n) {
  return metric_value(metric, p, K, y, w, n);
}

double ora_metric_cfg(const bst_gbdt_config* cfg, int32_t metric, const float* p, int32_t K,
                      const float* y, const float* w, int64_t n) {
  return metric_value(metric, p, K, y, w, n, payload_of(cfg));
}

int ora_default_metric(int32_t objective) { return default_metric(objective); }

}  // extern "C"
