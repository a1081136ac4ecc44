// binning.cpp — bin mapper construction on the host.
//
// The per-element quantisation (value_to_bin over all n x F values) runs on
// the GPU (k_value_to_bin); this file derives the bin edges exactly as the
// reference does:
//   analysis      FeatureStats::update/into_analysis, data/binned/feature_analysis.rs:313-432
//   sampling      sample_indices_for_binning (splitmix64, 200k rows), dataset.rs:598-629
//   numeric edges bin_numeric, dataset.rs:683-750
//   categories    create_bin_mappers categorical branch, dataset.rs:660-677
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
};

Stats analyze(const float* col, int64_t n, bool user_cat, const bst_binning_config& cfg) {
  const size_t cap = std::max<size_t>(cfg.max_categorical_cardinality, cfg.max_bins) + 10;
  std::vector<float> uniq;
  uniq.reserve(cap);
  bool exceeded = false, all_int = true;
  int64_t valid = 0, nan = 0;
  for (int64_t i = 0; i < n; i++) {
    const float v = col[i];
    if (std::isnan(v)) {
      nan++;
      continue;
    }
    valid++;
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
    out[f] = std::move(m);
  });
  return out;
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
