// binning.cpp — bin mapper construction on the host.
//
// The per-element quantisation (value_to_bin over all n x F values) runs on
// the GPU (k_value_to_bin); this file derives the bin edges exactly as the
// reference does:
//   analysis      FeatureStats::update/into_analysis, data/binned/feature_analysis.rs:313-432
//   sampling      sample_indices_for_binning (splitmix64, 200k rows), dataset.rs:598-629
//   numeric edges bin_numeric, dataset.rs:683-750
//   categories    create_bin_mappers categorical branch, dataset.rs:660-677
//   bundling      Exclusive Feature Bundling of sparse categorical features,
//                 bundling.rs:203-485, encoded by build_bundle, dataset.rs:898-952
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>
#include <unordered_set>

#include "host_common.h"

namespace bst {

namespace {

struct Stats {
  bool trivial = false, numeric = true;
  int64_t n_nan = 0;
  int64_t n_nonzero = 0;  // finite or infinite values != 0 (FeatureStats::non_zero_count)
};

Stats analyze(const float* col, int64_t n, bool user_cat, const bst_binning_config& cfg) {
  const size_t cap = std::max<size_t>(cfg.max_categorical_cardinality, cfg.max_bins) + 10;
  std::vector<float> uniq;
  uniq.reserve(cap);
  bool exceeded = false, all_int = true;
  int64_t valid = 0, nan = 0, nonzero = 0;
  for (int64_t i = 0; i < n; i++) {
    const float v = col[i];
    if (std::isnan(v)) {
      nan++;
      continue;
    }
    valid++;
    if (v != 0.0f) nonzero++;
    if (all_int && (v - std::trunc(v)) != 0.0f) all_int = false;
    if (!exceeded) {
      bool seen = false;
      for (float u : uniq)
        if (u == v) {
          seen = true;
          break;
        }
      if (!seen) {
        if (uniq.size() >= cap)
          exceeded = true;
        else
          uniq.push_back(v);
      }
    }
  }
  Stats s;
  const size_t nu = uniq.size();
  s.trivial = valid == 0 || nu == 0 || (nu == 1 && uniq[0] == 0.0f);
  const bool auto_cat = !user_cat && all_int && nu <= cfg.max_categorical_cardinality;
  s.numeric = user_cat ? false : (nu == 2 ? true : !auto_cat);
  s.n_nan = nan;
  s.n_nonzero = nonzero;
  return s;
}

Mapper numeric_mapper(std::vector<float> vals, uint32_t max_bins, bool has_nan) {
  // Keep finite values; min/max decide the single-bin case.
  float lo = std::numeric_limits<float>::max(), hi = -std::numeric_limits<float>::max();
  size_t k = 0;
  for (float v : vals)
    if (std::isfinite(v)) {
      lo = std::min(lo, v);
      hi = std::max(hi, v);
      vals[k++] = v;
    }
  vals.resize(k);
  Mapper m;
  if (vals.empty() || lo >= hi) {
    m.bounds = {std::numeric_limits<double>::max()};
    m.n_bins = 1;
    return m;
  }
  const uint64_t nb = std::min<uint64_t>(max_bins, vals.size());
  std::stable_sort(vals.begin(), vals.end(), [](float a, float b) { return a < b; });
  vals.erase(std::unique(vals.begin(), vals.end(), [](float a, float b) { return a == b; }),
             vals.end());
  const size_t nu = vals.size();
  if (nu <= nb) {
    for (size_t i = 0; i + 1 < nu; i++)
      m.bounds.push_back(((double)vals[i] + (double)vals[i + 1]) / 2.0);
  } else {
    for (uint64_t i = 1; i < nb; i++) {
      const double q = (double)i / (double)nb;
      const size_t j = std::min<size_t>((size_t)std::floor(q * (double)(nu - 1)), nu - 2);
      const double b = ((double)vals[j] + (double)vals[j + 1]) / 2.0;
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

std::vector<uint32_t> edge_sample(int64_t n, uint64_t cnt) {
  std::vector<uint32_t> idx;
  if ((uint64_t)n <= cnt) return idx;
  uint64_t state = 0xD1B54A32D192ED03ULL ^ (uint64_t)n;
  std::unordered_set<uint32_t> seen;
  seen.reserve(cnt * 2);
  idx.reserve(cnt);
  while (idx.size() < cnt) {
    state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const uint32_t r = (uint32_t)(z % (uint64_t)n);
    if (seen.insert(r).second) idx.push_back(r);
  }
  std::sort(idx.begin(), idx.end());
  return idx;
}

int32_t to_i32_sat(double v) {
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (int32_t)0x80000000;
  return (int32_t)v;
}

template <class Fn>
void parallel_for(int32_t count, Fn fn) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(hw ? hw : 4, 16);
  nt = std::max(1, std::min(nt, (int)count));
  std::vector<std::thread> th;
  std::atomic<int32_t> next{0};
  for (int t = 0; t < nt; t++)
    th.emplace_back([&] {
      for (int32_t i = next++; i < count; i = next++) fn(i);
    });
  for (auto& t : th) t.join();
}

}  // namespace

std::vector<Mapper> build_mappers(const float* X, int64_t n, int32_t F, const uint8_t* is_cat,
                                  const bst_binning_config& cfg) {
  std::vector<Mapper> out(F);
  std::vector<Stats> st(F);
  parallel_for(F, [&](int32_t f) { st[f] = analyze(X + (int64_t)f * n, n, is_cat && is_cat[f], cfg); });
  const std::vector<uint32_t> sample = edge_sample(n, cfg.sample_cnt);
  parallel_for(F, [&](int32_t f) {
    const float* col = X + (int64_t)f * n;
    Mapper m;
    if (st[f].numeric) {
      std::vector<float> vals;
      if (sample.empty())
        vals.assign(col, col + n);
      else {
        vals.resize(sample.size());
        for (size_t i = 0; i < sample.size(); i++) vals[i] = col[sample[i]];
      }
      m = numeric_mapper(std::move(vals), cfg.max_bins, st[f].n_nan > 0);
    } else {
      std::unordered_set<int32_t> seen;
      for (int64_t i = 0; i < n; i++)
        if (std::isfinite(col[i])) seen.insert(to_i32_sat((double)col[i]));
      m.categorical = true;
      m.cats.assign(seen.begin(), seen.end());
      std::sort(m.cats.begin(), m.cats.end());
      if (m.cats.empty()) m.cats.push_back(0);
      m.n_bins = (uint32_t)m.cats.size();
    }
    m.trivial = st[f].trivial;
    m.n_nonzero = st[f].n_nonzero;
    out[f] = std::move(m);
  });
  return out;
}

namespace {

// The row sample of the conflict graph draws from rand_xoshiro 0.6's
// Xoshiro256PlusPlus::seed_from_u64 (the state is the first four outputs of
// SplitMix64 from the seed; xoshiro256++ of Blackman and Vigna) through rand
// 0.8.5's UniformInt::sample_single_inclusive for 64-bit integers (widening
// multiply, the high word accepted when the low word is within the
// conservative zone (range << leading_zeros(range)) - 1).
struct Xoshiro256pp {
  uint64_t s[4];
  explicit Xoshiro256pp(uint64_t seed) {
    uint64_t x = seed;
    for (uint64_t& w : s) {
      x += 0x9E3779B97F4A7C15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Uniform in [0, hi].
  uint64_t below_incl(uint64_t hi) {
    const uint64_t range = hi + 1;
    if (range == 0) return next();
    const uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
      const unsigned __int128 m = (unsigned __int128)next() * range;
      if ((uint64_t)m <= zone) return (uint64_t)(m >> 64);
    }
  }
};

// sample_rows, bundling.rs:203-251: the first and last tenth of the sample
// are the first and last rows, the rest a reservoir sample of the middle.
std::vector<uint32_t> bundle_rows(int64_t n, uint64_t max_samples, uint64_t seed) {
  std::vector<uint32_t> out;
  if ((uint64_t)n <= max_samples) {
    out.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) out[(size_t)i] = (uint32_t)i;
    return out;
  }
  const uint64_t n_first = max_samples / 10, n_last = max_samples / 10;
  const uint64_t n_random = max_samples - n_first - n_last;
  for (uint64_t i = 0; i < std::min<uint64_t>(n_first, (uint64_t)n); i++) out.push_back((uint32_t)i);
  const uint64_t last_start = (uint64_t)n > n_last ? (uint64_t)n - n_last : 0;
  for (uint64_t i = last_start; i < (uint64_t)n; i++) out.push_back((uint32_t)i);
  if (last_start > n_first) {
    const uint64_t len = last_start - n_first;
    std::vector<uint32_t> sel;
    for (uint64_t i = 0; i < std::min(n_random, len); i++) sel.push_back((uint32_t)(n_first + i));
    Xoshiro256pp rng(seed);
    for (uint64_t i = sel.size(); i < len; i++) {
      const uint64_t j = rng.below_incl(i);
      if (j < sel.size()) sel[j] = (uint32_t)(n_first + i);
    }
    out.insert(out.end(), sel.begin(), sel.end());
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

}  // namespace

// create_bundle_plan, bundling.rs:393-485, with BundlingConfig's defaults
// (max conflict rate 1e-4, 256 features per bundle, 1000 sparse features,
// 10000 sampled rows, seed 42) and min_sparsity = the binning config's
// sparsity threshold (dataset.rs:229-236).
std::vector<Bundle> plan_bundles(const float* X, int64_t n, int32_t F,
                                 const std::vector<Mapper>& maps, const bst_binning_config& cfg) {
  // Sparse categorical features: density (non-zero share, in f32) at most
  // 1 - min_sparsity (feature_analysis.rs:250-261).
  const float limit = 1.0f - cfg.sparsity_threshold;
  std::vector<uint32_t> sparse;
  std::vector<float> density;
  for (int32_t f = 0; f < F; f++) {
    const Mapper& m = maps[(size_t)f];
    const float d = (float)m.n_nonzero / (float)n;
    if (!m.trivial && m.categorical && d <= limit) {
      sparse.push_back((uint32_t)f);
      density.push_back(d);
    }
  }
  const size_t S = sparse.size();
  if (S < 2 || S > 1000) return {};
  const std::vector<uint32_t> rows = bundle_rows(n, 10000, 42);
  const size_t R = rows.size(), words = (R + 63) / 64;
  // Non-zero, non-NaN rows of each feature among the sampled rows.
  std::vector<uint64_t> bits(S * words, 0ull);
  for (size_t i = 0; i < S; i++) {
    const float* col = X + (int64_t)sparse[i] * n;
    for (size_t k = 0; k < R; k++) {
      const float v = col[rows[k]];
      if (v != 0.0f && !std::isnan(v)) bits[i * words + k / 64] |= 1ull << (k % 64);
    }
  }
  std::vector<uint64_t> conflict(S * S, 0);
  size_t pairs = 0;
  for (size_t i = 0; i < S; i++)
    for (size_t j = i + 1; j < S; j++) {
      uint64_t c = 0;
      for (size_t w = 0; w < words; w++) c += __builtin_popcountll(bits[i * words + w] & bits[j * words + w]);
      conflict[i * S + j] = conflict[j * S + i] = c;
      pairs += c > 0;
    }
  const size_t total_pairs = S * (S - 1) / 2;
  if ((float)pairs / (float)total_pairs > 0.5f) return {};
  const uint64_t max_conflicts = (uint64_t)std::ceil(0.0001f * (float)R);
  // Denser features first (a stable sort: ties keep feature order), each into
  // the earliest bundle with the smallest conflict increase that stays within
  // the budget, or a new bundle (assign_bundles, bundling.rs:277-339).
  std::vector<size_t> order(S);
  for (size_t i = 0; i < S; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return density[a] > density[b]; });
  std::vector<std::vector<size_t>> members;
  std::vector<uint64_t> used;
  for (size_t p : order) {
    size_t best = SIZE_MAX;
    uint64_t best_inc = UINT64_MAX;
    for (size_t b = 0; b < members.size(); b++) {
      if (members[b].size() >= 256) continue;
      uint64_t inc = 0;
      for (size_t q : members[b]) inc += conflict[p * S + q];
      if (used[b] + inc <= max_conflicts && inc < best_inc) {
        best = b;
        best_inc = inc;
      }
    }
    if (best != SIZE_MAX) {
      members[best].push_back(p);
      used[best] += best_inc;
    } else {
      members.push_back({p});
      used.push_back(0);
    }
  }
  std::vector<Bundle> out;
  for (const auto& mb : members) {
    if (mb.size() < 2) continue;
    Bundle b;
    uint32_t off = 1;  // encoded bin 0: every feature of the bundle at its default
    for (size_t p : mb) {
      b.features.push_back(sparse[p]);
      b.offsets.push_back(off);
      off += maps[sparse[p]].n_bins;
    }
    b.total_bins = off;
    out.push_back(std::move(b));
  }
  return out;
}

uint32_t cat_bin(const Mapper& m, float v) {
  const int32_t c = to_i32_sat((double)v);
  const auto it = std::lower_bound(m.cats.begin(), m.cats.end(), c);
  return it != m.cats.end() && *it == c ? (uint32_t)(it - m.cats.begin()) : 0u;
}

// build_bundle, dataset.rs:898-952: a row takes the code offset + bin of the
// first feature of the bundle (in bundle order) whose value is non-zero and
// not NaN; 0 when there is none.
void encode_bundle(const Bundle& b, const float* X, int64_t n, const std::vector<Mapper>& maps,
                   uint8_t* out) {
  std::fill(out, out + n, (uint8_t)0);
  for (size_t j = 0; j < b.features.size(); j++) {
    const Mapper& m = maps[b.features[j]];
    const float* col = X + (int64_t)b.features[j] * n;
    for (int64_t i = 0; i < n; i++) {
      const float v = col[i];
      if (out[i] == 0 && v != 0.0f && !std::isnan(v))
        out[i] = (uint8_t)(b.offsets[j] + cat_bin(m, v));
    }
  }
}

// next_up_f32, training/gbdt/grower.rs:543-556.
float next_up_f32(float x) {
  if (std::isnan(x) || x == std::numeric_limits<float>::infinity()) return x;
  if (x == 0.0f) {
    uint32_t one = 1;
    float r;
    std::memcpy(&r, &one, 4);
    return r;
  }
  uint32_t b;
  std::memcpy(&b, &x, 4);
  b = x > 0 ? b + 1 : b - 1;
  float r;
  std::memcpy(&r, &b, 4);
  return r;
}

// Largest f32 that is <= b, so that for every f32 v:  v <= b  <=>  v <= round_down_f32(b).
float round_down_f32(double b) {
  float f = (float)b;
  if ((double)f > b) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}

}  // namespace bst
