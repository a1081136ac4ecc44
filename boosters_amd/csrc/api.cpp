// api.cpp — the extern "C" boundary (include/boosters_amd.h).
//
// Every entry point converts C++ exceptions into status codes; nothing throws
// or aborts across the boundary.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <mutex>

#include "host_common.h"
#include "trainer.h"

using namespace bst;

// ---- static scratch of the allocation-free predict entry point ------------
// bst_predict_device_noalloc's two-phase launches keep every tree's leaf value
// per row here instead of in an allocated buffer.  One copy per GPU, for the
// life of the process; its device address is looked up once per device
// (bst_ctx_create, device current) and only that address is used from host
// code.  Calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocScratchFloats = (size_t)8 << 20;  // 32 MiB
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current).
float* scratch_on_device(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) raise(BST_E_DEVICE, "device index out of range");
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      raise(BST_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<float*>(p);
  }
  return g_scratch_dev[device];
}
}  // namespace

namespace bst {
std::string& last_error_ref() {
  static thread_local std::string err;
  return err;
}
}  // namespace bst

namespace {

#define g_err (::bst::last_error_ref())

bst_binning_config default_binning() {
  bst_binning_config b;
  b.max_bins = 256;
  b.sparsity_threshold = 0.9f;
  b.enable_bundling = 1;
  b.max_categorical_cardinality = 0;
  b.sample_cnt = 200000;
  return b;
}

void set_device(bst_ctx* ctx) { BST_HIP(hipSetDevice(ctx->device)); }

// Builds the device dataset from a host [F][n] matrix (raw_dev: device copy
// of the same matrix, owned by the dataset).
bst_dataset* make_dataset(bst_ctx* ctx, const float* X, int64_t n, int32_t F,
                          const uint8_t* is_cat, const bst_binning_config& cfg,
                          DevBuf<float>&& raw_dev) {
  need(n > 0 && F > 0, BST_E_SHAPE, "empty dataset");
  need(n < (int64_t)UINT32_MAX / 2, BST_E_SHAPE, "too many rows for 32-bit row ids");
  need(cfg.max_bins >= 2 && cfg.max_bins <= 256, BST_E_CONFIG,
       "max_bins must be in [2, 256] (u8 bins + NaN bin)");
  const auto t0 = std::chrono::steady_clock::now();
  auto* ds = new bst_dataset();
  std::unique_ptr<bst_dataset> guard_ds(ds);
  ds->ctx = ctx;
  ds->mappers = build_mappers(X, n, F, is_cat, cfg);
  // Effective columns (BinnedDataset::feature_views, dataset.rs:460-520):
  // the bundles of sparse categorical features first, each one categorical
  // column of total_bins codes whose code 0 (every feature at its default)
  // is the missing bin; then every other non-trivial feature in feature order.
  if (cfg.enable_bundling) ds->bundles = plan_bundles(X, n, F, ds->mappers, cfg);
  std::vector<uint8_t> bundled((size_t)F, 0);
  for (const Bundle& b : ds->bundles) {
    if (b.total_bins > 256)
      raise(BST_E_CONFIG, "a bundle of sparse categorical features needs " +
                              std::to_string(b.total_bins) +
                              " codes; at most 256 are supported (disable bundling)");
    for (uint32_t f : b.features) bundled[f] = 1;
    ds->eff_to_orig.push_back(b.features[0]);
    FeatMeta fm{};
    fm.n_bins = b.total_bins;
    fm.default_bin = 0;
    fm.is_cat = 1;
    fm.has_missing = 1;
    ds->fmeta.push_back(fm);
    ds->d.max_cat_bins = std::max<int32_t>(ds->d.max_cat_bins, (int32_t)b.total_bins);
  }
  const int32_t n_bundles = (int32_t)ds->bundles.size();
  bool hi = false;
  for (int32_t f = 0; f < F; f++) {
    const Mapper& m = ds->mappers[f];
    if (m.trivial || bundled[f]) continue;
    if (m.n_bins > 257 || (m.categorical && m.n_bins > 256))
      raise(BST_E_CONFIG, "feature " + std::to_string(f) + " has " + std::to_string(m.n_bins) +
                              " bins; at most 256 categories / 256 bins + NaN are supported");
    ds->eff_to_orig.push_back((uint32_t)f);
    FeatMeta fm{};
    fm.n_bins = m.n_bins;
    fm.default_bin = m.default_bin;
    fm.is_cat = m.categorical;
    fm.has_missing = m.missing != 0;
    ds->fmeta.push_back(fm);
    hi = hi || m.n_bins > 256;
    if (m.categorical) ds->d.max_cat_bins = std::max<int32_t>(ds->d.max_cat_bins, (int32_t)m.n_bins);
  }
  const int32_t E = (int32_t)ds->eff_to_orig.size();
  need(E > 0, BST_E_SHAPE, "all features are trivial");
  need(E <= 1024, BST_E_SHAPE, "more than 1024 non-trivial features");
  DatasetDev& d = ds->d;
  d.n = n;
  d.F = F;
  d.E = E;
  d.hi_off = hi ? E : -1;
  {
    // Row: E bin bytes, then the NaN bits (one per feature) when any feature
    // has 257 bins, padded to 16 bytes and so that the bytes of the padding
    // features of the last 32-feature group (bins and NaN bits, all zero)
    // lie inside the row: histogram lanes of padding features read them
    // instead of testing their feature index.
    const int32_t ng32 = (E + 31) / 32;
    const int32_t used =
        std::max(E + (hi ? (E + 7) / 8 : 0), std::max(32 * ng32, hi ? E + 4 * ng32 : 0));
    d.S = (used + 15) / 16 * 16;
  }
  d.FG = 32;
  d.NG = (E + d.FG - 1) / d.FG;
  d.hi_words = (n + 31) / 32;
  // Tables for value_to_bin and thresholds.
  std::vector<BinMeta> bm(E);
  std::vector<float> thr;
  std::vector<int32_t> cats;
  ds->thr_up.assign((size_t)E * 256, 0.0f);
  for (int32_t e = 0; e < E; e++) {
    const Mapper& m = ds->mappers[ds->eff_to_orig[e]];
    BinMeta b{};
    b.orig = ds->eff_to_orig[e];
    if (e < n_bundles) {
      // Coded on the host below; the device pass writes zeros here.
      b.is_cat = 1;
      bm[e] = b;
      continue;
    }
    b.n_bins = m.n_bins;
    b.default_bin = m.default_bin;
    b.is_cat = m.categorical;
    b.missing_nan = m.missing == 2;
    if (m.categorical) {
      b.table_off = (uint32_t)cats.size();
      cats.insert(cats.end(), m.cats.begin(), m.cats.end());
    } else {
      int64_t ns = (int64_t)m.n_bins - 1 - (m.missing == 2 ? 1 : 0);
      b.n_search = (uint32_t)std::max<int64_t>(ns, 0);
      b.table_off = (uint32_t)thr.size();
      for (uint32_t i = 0; i < b.n_search; i++) thr.push_back(round_down_f32(m.bounds[i]));
      for (uint32_t i = 0; i < m.n_bins && i < 256; i++)
        ds->thr_up[(size_t)e * 256 + i] = next_up_f32((float)m.bounds[i]);
    }
    bm[e] = b;
  }
  if (thr.empty()) thr.push_back(0.0f);
  if (cats.empty()) cats.push_back(0);
  hipStream_t s = ctx->stream;
  DevBuf<BinMeta> d_bm(E);
  DevBuf<float> d_thr(thr.size());
  DevBuf<int32_t> d_cats(cats.size());
  h2d(s, d_bm.p, bm.data(), bm.size());
  h2d(s, d_thr.p, thr.data(), thr.size());
  h2d(s, d_cats.p, cats.data(), cats.size());
  ds->raw = std::move(raw_dev);
  ds->bins_cm.alloc((size_t)E * n);
  ds->bins_rm.alloc((size_t)n * d.S + 64);
  if (hi) ds->hi_cm.alloc((size_t)E * d.hi_words);
  ds->meta.alloc(E);
  h2d(s, ds->meta.p, ds->fmeta.data(), ds->fmeta.size());
  d.bins_cm = ds->bins_cm.p;
  d.bins_rm = ds->bins_rm.p;
  d.hi_cm = hi ? ds->hi_cm.p : nullptr;
  d.meta = ds->meta.p;
  d.raw = ds->raw.p;
  launch_value_to_bin(s, ds->raw.p, n, E, d_bm.p, d_thr.p, d_cats.p, d.bins_cm, d.hi_cm,
                      d.hi_words);
  BST_HIP(hipGetLastError());
  if (n_bundles > 0) {
    std::vector<uint8_t> codes((size_t)n_bundles * n);
    for (int32_t e = 0; e < n_bundles; e++)
      encode_bundle(ds->bundles[e], X, n, ds->mappers, codes.data() + (size_t)e * n);
    h2d(s, d.bins_cm, codes.data(), codes.size());  // columns 0 .. n_bundles - 1
    BST_HIP(hipStreamSynchronize(s));               // `codes` is released on return
  }
  launch_to_rowmajor(s, d);
  BST_HIP(hipGetLastError());
  BST_HIP(hipStreamSynchronize(s));
  ds->ms_binning =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return guard_ds.release();
}

int default_metric_of(int obj) {
  if (obj == BST_OBJ_LAMBDARANK) return BST_METRIC_NONE;  // no ranking metric in the reference
  return obj == BST_OBJ_LOGISTIC ? BST_METRIC_LOGLOSS
                                 : (obj == BST_OBJ_SOFTMAX ? BST_METRIC_MLOGLOSS : BST_METRIC_RMSE);
}

// Objective::n_outputs, objectives/mod.rs:221-233.
int n_outputs_of(const bst_gbdt_config& c) {
  if (c.objective == BST_OBJ_SOFTMAX) return (int)c.n_classes;
  if (c.objective == BST_OBJ_PINBALL) return (int)c.n_alphas;
  return 1;
}

ObjParams obj_params_of(const bst_gbdt_config& c) {
  ObjParams op;
  op.huber_delta = c.huber_delta;
  for (uint32_t i = 0; i < c.n_alphas && i < 16 && c.alphas; i++) op.alphas[i] = c.alphas[i];
  return op;
}

int validate(const bst_gbdt_config* c) {
  if (!c) raise(BST_E_INVALID, "null config");
  // GBDTConfig::validate, model/gbdt/config.rs:252-310.
  if (!(c->learning_rate > 0.0f)) raise(BST_E_CONFIG, "learning_rate must be positive");
  if (c->n_trees == 0) raise(BST_E_CONFIG, "n_trees must be at least 1");
  if (!(c->subsample > 0.0f && c->subsample <= 1.0f)) raise(BST_E_CONFIG, "subsample must be in (0, 1]");
  if (!(c->colsample_bytree > 0.0f && c->colsample_bytree <= 1.0f))
    raise(BST_E_CONFIG, "colsample_bytree must be in (0, 1]");
  if (!(c->colsample_bylevel > 0.0f && c->colsample_bylevel <= 1.0f))
    raise(BST_E_CONFIG, "colsample_bylevel must be in (0, 1]");
  if (c->lambda < 0.0f) raise(BST_E_CONFIG, "lambda must be non-negative");
  if (c->alpha < 0.0f) raise(BST_E_CONFIG, "alpha must be non-negative");
  if (c->min_child_weight < 0.0f) raise(BST_E_CONFIG, "min_child_weight must be non-negative");
  if (c->min_gain < 0.0f) raise(BST_E_CONFIG, "min_gain must be non-negative");
  if (c->objective < 0 || c->objective > BST_OBJ_POISSON)
    raise(BST_E_CONFIG, "objective not supported by this engine build");
  // Objective payloads (the bindings' Objective.huber / Objective.pinball checks).
  if (c->objective == BST_OBJ_PSEUDO_HUBER && !(c->huber_delta > 0.0f))
    raise(BST_E_CONFIG, "pseudo-Huber delta must be positive");
  const bool quantile_metric = c->metric == BST_METRIC_QUANTILE ||
                               (c->metric == BST_METRIC_DEFAULT && c->objective == BST_OBJ_PINBALL);
  if (c->objective == BST_OBJ_PINBALL || quantile_metric) {
    if (c->n_alphas < 1 || !c->alphas) raise(BST_E_CONFIG, "quantiles need at least one alpha");
    for (uint32_t i = 0; i < c->n_alphas; i++)
      if (!(c->alphas[i] > 0.0f && c->alphas[i] < 1.0f))
        raise(BST_E_CONFIG, "each quantile alpha must be in (0, 1)");
  }
  if (c->objective == BST_OBJ_PINBALL && c->n_alphas > kPredMaxGroups)
    raise(BST_E_CONFIG, "pinball supports at most 16 quantiles");
  if (c->metric < BST_METRIC_DEFAULT || c->metric > BST_METRIC_MULTICLASS_ACCURACY)
    raise(BST_E_CONFIG, "metric not supported by this engine build");
  if (c->metric == BST_METRIC_HUBER && !(c->metric_huber_delta > 0.0))
    raise(BST_E_CONFIG, "Huber metric delta must be positive");
  if (c->objective == BST_OBJ_SOFTMAX && (c->n_classes < 2 || c->n_classes > kMaxOutputs))
    raise(BST_E_CONFIG, "softmax needs 2..1024 classes");
  if (c->growth == BST_GROW_DEPTHWISE && c->max_depth > 16) raise(BST_E_CONFIG, "max_depth > 16");
  if (c->growth == BST_GROW_LEAFWISE && (c->max_leaves < 1 || c->max_leaves > 65536))
    raise(BST_E_CONFIG, "max_leaves must be in [1, 65536]");
  return BST_OK;
}

}  // namespace

void bst_model::upload(hipStream_t s) {
  const size_t nn = split_index.size();
  std::vector<DevNode> nodes(nn);
  for (size_t i = 0; i < nn; i++) {
    uint32_t meta = split_index[i] & 0xFFFFFFu;
    if (is_leaf[i]) meta |= kNodeLeaf;
    if (default_left[i]) meta |= kNodeDefLeft;
    if (split_type[i]) meta |= kNodeCat;
    nodes[i] = {meta, is_leaf[i] ? leaf_value[i] : split_threshold[i], left[i], right[i]};
  }
  d_nodes.alloc(std::max<size_t>(nn, 1));
  d_tree_off.alloc(tree_offset.size());
  // Padded by 8 entries: a prediction block loads 8 trees' groups at once.
  std::vector<uint32_t> group_pad(tree_group);
  group_pad.resize(tree_group.size() + 8, 0u);
  d_tree_group.alloc(group_pad.size());
  d_cat_start.alloc(std::max<size_t>(nn, 1));
  d_cat_size.alloc(std::max<size_t>(nn, 1));
  d_cat_words.alloc(std::max<size_t>(cat_words.size(), 1));
  // Base scores padded by kPredMaxGroups zeros: a prediction block loads a
  // whole window of them unconditionally.
  std::vector<float> base_pad(base_score);
  base_pad.resize(base_score.size() + kPredMaxGroups, 0.0f);
  d_base.alloc(base_pad.size());
  h2d(s, d_nodes.p, nodes.data(), nn);
  h2d(s, d_tree_off.p, tree_offset.data(), tree_offset.size());
  h2d(s, d_tree_group.p, group_pad.data(), group_pad.size());
  h2d(s, d_cat_start.p, cat_start.data(), nn);
  h2d(s, d_cat_size.p, cat_size.data(), nn);
  h2d(s, d_cat_words.p, cat_words.data(), cat_words.size());
  h2d(s, d_base.p, base_pad.data(), base_pad.size());
  // Predictor layout: trees of at most kPredChunkNodes nodes testing features
  // below 2^16 - 1 get a compact copy (engine.h), renumbered breadth first so
  // that a node's children are adjacent (left, left + 1), and are grouped
  // into LDS chunks of consecutive trees; any other tree is a chunk of its
  // own, traversed in the 16-byte layout above.
  const size_t T = tree_offset.empty() ? 0 : tree_offset.size() - 1;
  need(T == 0 || tree_offset.back() < (int64_t)UINT32_MAX, BST_E_SHAPE,
       "forest has more than 2^32 nodes");
  auto fbits = [](float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
  };
  std::vector<uint2> cn;
  std::vector<uint32_t> ctoff(std::max<size_t>(T, 1), 0), cdep(std::max<size_t>(T, 1), 0), cref;
  std::vector<uint32_t> corig;  // per compact node
  std::vector<PredChunk> ch;
  std::vector<uint32_t> new_id, order;
  auto compactable = [&](size_t t) {
    const int64_t b = tree_offset[t], e = tree_offset[t + 1];
    if (e - b + 1 > kPredChunkNodes) return false;  // with the chunk's sentinel leaf
    for (int64_t k = b; k < e; k++)
      if (!is_leaf[k] && split_index[k] >= 65535u) return false;
    return true;
  };
  // Tree t appended to the chunk that starts at compact node `begin`.
  std::vector<uint32_t> depth;
  auto append_compact = [&](size_t t, uint32_t begin) {
    const int64_t b = tree_offset[t];
    const uint32_t nn = (uint32_t)(tree_offset[t + 1] - b);
    new_id.assign(nn, UINT32_MAX);
    depth.assign(nn, 0u);
    order.clear();
    new_id[0] = 0;
    order.push_back(0);
    uint32_t next = 1, levels = 0;
    for (size_t i = 0; i < order.size(); i++) {
      const uint32_t o = order[i];
      levels = std::max(levels, depth[o]);
      if (is_leaf[b + o]) continue;
      const uint32_t l = left[b + o], r = right[b + o];
      new_id[l] = next;
      new_id[r] = next + 1;
      depth[l] = depth[r] = depth[o] + 1u;
      next += 2;
      order.push_back(l);
      order.push_back(r);
    }
    const size_t c0 = cn.size();
    const uint32_t rel = (uint32_t)c0 - begin;  // the tree's root index in the chunk
    ctoff[t] = (uint32_t)c0;
    cdep[t] = levels;
    cn.resize(c0 + next, make_uint2(0u, 0u));
    corig.resize(c0 + next, 0u);
    for (uint32_t o : order) {
      const int64_t k = b + o;
      uint2 w;
      if (is_leaf[k]) {
        const float v = leaf_value[k];
        w.x = (rel + new_id[o]) << kCShift;
        w.y = ~(std::isnan(v) ? 0x7FC00000u : fbits(v));
      } else {
        const uint32_t dl = default_left[k] ? 1u : 0u;
        w.x = (split_index[k] + 1u) | (dl ? kCDefLeft : 0u) |
              ((rel + new_id[left[k]]) << kCShift);
        if (split_type[k]) {
          w.x |= kCCat;
          w.y = (uint32_t)(cref.size() / 2) + 1u;
          cref.push_back(cat_start[k]);
          cref.push_back(cat_size[k]);
        } else {
          // A NaN threshold sends every value right (x < NaN is false) and
          // NaN by default_left: key 0.
          const float thr = split_threshold[k];
          w.y = (std::isnan(thr) ? 0u : pred_order_key(fbits(thr))) + dl;
        }
      }
      cn[c0 + new_id[o]] = w;
      corig[c0 + new_id[o]] = o;
    }
  };
  for (size_t t = 0; t < T;) {
    if (!compactable(t)) {
      ch.push_back({(uint32_t)t, (uint32_t)(t + 1), 0u, 0u});
      t++;
      continue;
    }
    if (cn.size() & 1) {  // chunks start 16-byte aligned
      cn.push_back(make_uint2(0u, 0u));
      corig.push_back(0u);
    }
    const uint32_t begin = (uint32_t)cn.size();
    // The trees that fit the buffer with the chunk's sentinel leaf; a few
    // more than a multiple of kPredILP are cut back to it when later trees
    // follow (they fill the next chunk's groups instead).
    size_t e = t, fill = 0;
    while (e < T && compactable(e) &&
           fill + (size_t)(tree_offset[e + 1] - tree_offset[e]) + 1 <= (size_t)kPredChunkNodes) {
      fill += (size_t)(tree_offset[e + 1] - tree_offset[e]);
      e++;
    }
    const size_t nt = e - t, idle = (kPredILP - nt % kPredILP) % kPredILP;
    if (e < T && nt > (size_t)kPredILP && idle * 10 > nt) e = t + nt / kPredILP * kPredILP;
    for (size_t u = t; u < e; u++) append_compact(u, begin);
    // Sentinel leaf (value 0) at its own index.
    cn.push_back(make_uint2(((uint32_t)cn.size() - begin) << kCShift, 0xFFFFFFFFu));
    corig.push_back(0u);
    ch.push_back({(uint32_t)t, (uint32_t)e, begin, (uint32_t)(cn.size() - begin)});
    t = e;
  }
  if (cn.size() & 1) cn.push_back(make_uint2(0u, 0u));
  if (cn.empty()) cn.assign(2, make_uint2(0u, 0u));
  corig.resize(cn.size(), 0u);
  pred_has_cat = cref.empty() ? 0 : 1;
  if (cref.empty()) cref.assign(2, 0u);
  n_chunks = (int)ch.size();
  d_cnodes.alloc(cn.size());
  // Tree offsets and levels padded by 8 entries: a block loads a group's
  // (at most 8) in one unconditional load.
  ctoff.resize(ctoff.size() + 8, 0u);
  cdep.resize(cdep.size() + 8, 0u);
  d_ctree_off.alloc(ctoff.size());
  d_cdepth.alloc(cdep.size());
  d_corig.alloc(corig.size());
  d_cat_ref.alloc(cref.size());
  d_chunks.alloc(std::max<size_t>(ch.size(), 1));
  h2d(s, d_cnodes.p, cn.data(), cn.size());
  h2d(s, d_ctree_off.p, ctoff.data(), ctoff.size());
  h2d(s, d_cdepth.p, cdep.data(), cdep.size());
  h2d(s, d_corig.p, corig.data(), corig.size());
  h2d(s, d_cat_ref.p, cref.data(), cref.size());
  h2d(s, d_chunks.p, ch.data(), ch.size());
  BST_HIP(hipStreamSynchronize(s));
}

extern "C" {

const char* bst_last_error(void) { return g_err.c_str(); }
const char* bst_version(void) { return "boosters_amd 0.1.0 (gfx950)"; }

void bst_gbdt_config_default(bst_gbdt_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->objective = BST_OBJ_SQUARED;
  c->n_classes = 0;
  c->metric = BST_METRIC_DEFAULT;
  c->n_trees = 100;
  c->learning_rate = 0.3f;
  c->growth = BST_GROW_DEPTHWISE;
  c->max_depth = 6;
  c->max_leaves = 31;
  c->max_onehot_cats = 4;
  c->lambda = 1.0f;
  c->alpha = 0.0f;
  c->min_child_weight = 1.0f;
  c->min_gain = 0.0f;
  c->min_samples_leaf = 1;
  c->subsample = 1.0f;
  c->colsample_bytree = 1.0f;
  c->colsample_bylevel = 1.0f;
  c->binning = default_binning();
  c->early_stopping_rounds = 0;
  c->cache_size = 8;
  c->seed = 42;
  c->verbosity = 0;
  c->huber_delta = 1.0f;
  c->n_alphas = 0;
  c->alphas = nullptr;
  c->accuracy_threshold = 0.5f;
  c->metric_huber_delta = 1.0;
}

void bst_ext_config_default(bst_ext_config* e) {
  if (!e) return;
  std::memset(e, 0, sizeof(*e));
  e->schedule = BST_SCHED_GRAPH;
  e->n_threads = 0;
  e->autotune = 1;
}

int bst_gbdt_config_validate(const bst_gbdt_config* c) {
  return guard([&] { return validate(c); });
}

int bst_ctx_create(int device, bst_ctx** out) {
  return guard([&] {
    need(out != nullptr, BST_E_INVALID, "null out");
    int n = 0;
    BST_HIP(hipGetDeviceCount(&n));
    need(device >= 0 && device < n, BST_E_DEVICE, "no such HIP device");
    hipDeviceProp_t prop;
    BST_HIP(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
      raise(BST_E_DEVICE, std::string("engine is built for gfx950, device is ") + prop.gcnArchName);
    auto* c = new bst_ctx();
    c->device = device;
    c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (prop.sharedMemPerBlock > 0) c->lds_per_block = prop.sharedMemPerBlock;
    BST_HIP(hipSetDevice(device));
    std::unique_ptr<bst_ctx> guard_c(c);
    c->noalloc_scratch = scratch_on_device(device);
    BST_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = guard_c.release();
    return BST_OK;
  });
}

int bst_ctx_destroy(bst_ctx* ctx) {
  return guard([&] {
    if (!ctx) return BST_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->pool.release();
    ctx->pred_leaf.reset();
    if (ctx->staging) (void)hipHostFree(ctx->staging);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BST_OK;
  });
}

int bst_ctx_sync(bst_ctx* ctx) {
  return guard([&] {
    need(ctx, BST_E_INVALID, "null ctx");
    BST_HIP(hipStreamSynchronize(ctx->stream));
    return BST_OK;
  });
}

void* bst_ctx_stream(bst_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int bst_ctx_set_profiling(bst_ctx* ctx, int32_t enable) {
  return guard([&] {
    need(ctx != nullptr, BST_E_INVALID, "null ctx");
    ctx->profile = enable != 0;
    for (int k = 0; k < 3; k++) {
      ctx->k_launches[k] = 0;
      ctx->k_ms[k] = 0;
      ctx->k_bytes[k] = 0;
    }
    return BST_OK;
  });
}

int bst_ctx_kernel_stats(bst_ctx* ctx, int32_t kind, int64_t* launches, double* ms,
                         double* bytes) {
  return guard([&] {
    need(ctx && kind >= 0 && kind < 3, BST_E_INVALID, "bad argument");
    if (launches) *launches = ctx->k_launches[kind];
    if (ms) *ms = ctx->k_ms[kind];
    if (bytes) *bytes = ctx->k_bytes[kind];
    return BST_OK;
  });
}

int bst_kernel_copy(bst_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return guard([&] {
    need(ctx && dst && src && bytes >= 0 && bytes % 16 == 0, BST_E_INVALID, "bad argument");
    set_device(ctx);
    launch_copy16(ctx->stream, dst, src, bytes);
    BST_HIP(hipGetLastError());
    BST_HIP(hipStreamSynchronize(ctx->stream));
    return BST_OK;
  });
}

int bst_device_alloc(bst_ctx* ctx, int64_t bytes, void** out) {
  return guard([&] {
    need(ctx && out && bytes > 0, BST_E_INVALID, "bad argument");
    set_device(ctx);
    hipError_t e = hipMalloc(out, (size_t)bytes);
    if (e != hipSuccess) raise(BST_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return BST_OK;
  });
}

int bst_device_free(bst_ctx* ctx, void* p) {
  return guard([&] {
    need(ctx != nullptr, BST_E_INVALID, "null ctx");
    set_device(ctx);
    if (p) BST_HIP(hipFree(p));
    return BST_OK;
  });
}

int bst_memcpy_h2d(bst_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return guard([&] {
    need(ctx && dst && src, BST_E_INVALID, "null argument");
    set_device(ctx);
    BST_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    BST_HIP(hipStreamSynchronize(ctx->stream));
    return BST_OK;
  });
}

int bst_memcpy_d2h(bst_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return guard([&] {
    need(ctx && dst && src, BST_E_INVALID, "null argument");
    set_device(ctx);
    BST_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    BST_HIP(hipStreamSynchronize(ctx->stream));
    return BST_OK;
  });
}

int bst_dataset_create(bst_ctx* ctx, const float* X, int64_t n, int32_t F, const uint8_t* is_cat,
                       const bst_binning_config* binning, bst_dataset** out) {
  return guard([&] {
    need(ctx && X && out, BST_E_INVALID, "null argument");
    set_device(ctx);
    DevBuf<float> raw((size_t)n * F);
    h2d(ctx->stream, raw.p, X, (size_t)n * F);
    *out = make_dataset(ctx, X, n, F, is_cat, binning ? *binning : default_binning(),
                        std::move(raw));
    return BST_OK;
  });
}

int bst_dataset_create_device(bst_ctx* ctx, const float* Xd, int64_t n, int32_t F,
                              const uint8_t* is_cat, const bst_binning_config* binning,
                              bst_dataset** out) {
  return guard([&] {
    need(ctx && Xd && out, BST_E_INVALID, "null argument");
    set_device(ctx);
    // Edges are derived on the host from a copy; quantisation runs on device.
    std::vector<float> X((size_t)n * F);
    d2h(ctx->stream, X.data(), Xd, X.size());
    DevBuf<float> raw((size_t)n * F);
    BST_HIP(hipMemcpyAsync(raw.p, Xd, raw.bytes(), hipMemcpyDeviceToDevice, ctx->stream));
    BST_HIP(hipStreamSynchronize(ctx->stream));
    *out = make_dataset(ctx, X.data(), n, F, is_cat, binning ? *binning : default_binning(),
                        std::move(raw));
    return BST_OK;
  });
}

int bst_dataset_destroy(bst_dataset* ds) {
  return guard([&] {
    if (ds) {
      (void)hipSetDevice(ds->ctx->device);
      delete ds;
    }
    return BST_OK;
  });
}

int bst_dataset_feature_info(const bst_dataset* ds, int32_t f, uint32_t* nb, int32_t* mt,
                             uint32_t* db, int32_t* cat, int32_t* triv) {
  return guard([&] {
    need(ds && f >= 0 && f < ds->d.F, BST_E_INVALID, "bad feature");
    const Mapper& m = ds->mappers[f];
    if (nb) *nb = m.n_bins;
    if (mt) *mt = m.missing;
    if (db) *db = m.default_bin;
    if (cat) *cat = m.categorical;
    if (triv) *triv = m.trivial;
    return BST_OK;
  });
}

int bst_dataset_bin_bounds(const bst_dataset* ds, int32_t f, double* out, int64_t cap) {
  int r = guard([&] {
    need(ds && f >= 0 && f < ds->d.F, BST_E_INVALID, "bad feature");
    const Mapper& m = ds->mappers[f];
    int64_t k = 0;
    if (m.categorical)
      for (int32_t c : m.cats) {
        if (out && k < cap) out[k] = c;
        k++;
      }
    else
      for (double v : m.bounds) {
        if (out && k < cap) out[k] = v;
        k++;
      }
    return (int)k;
  });
  return r;
}

int bst_dataset_get_bins(const bst_dataset* ds, uint16_t* out) {
  return guard([&] {
    need(ds && out, BST_E_INVALID, "null argument");
    const auto& d = ds->d;
    std::vector<uint8_t> lo((size_t)d.E * d.n);
    std::vector<uint32_t> hi;
    d2h(ds->ctx->stream, lo.data(), d.bins_cm, lo.size());
    if (d.hi_cm) {
      hi.resize((size_t)d.E * d.hi_words);
      d2h(ds->ctx->stream, hi.data(), d.hi_cm, hi.size());
    }
    BST_HIP(hipStreamSynchronize(ds->ctx->stream));
    std::fill(out, out + (size_t)d.F * d.n, (uint16_t)0);
    // A bundled feature's bin is decoded from its bundle's code: its own
    // range of codes, else the bin of 0.0 (the value the code stands for).
    for (size_t e = 0; e < ds->bundles.size(); e++) {
      const Bundle& b = ds->bundles[e];
      for (size_t j = 0; j < b.features.size(); j++) {
        const Mapper& m = ds->mappers[b.features[j]];
        const uint32_t lo_c = b.offsets[j], hi_c = b.offsets[j] + m.n_bins;
        const uint16_t zero_bin = (uint16_t)cat_bin(m, 0.0f);
        uint16_t* o = out + (size_t)b.features[j] * d.n;
        for (int64_t r = 0; r < d.n; r++) {
          const uint32_t c = lo[e * d.n + r];
          o[r] = c >= lo_c && c < hi_c ? (uint16_t)(c - lo_c) : zero_bin;
        }
      }
    }
    for (int32_t e = (int32_t)ds->bundles.size(); e < d.E; e++) {
      uint16_t* o = out + (size_t)ds->eff_to_orig[e] * d.n;
      for (int64_t r = 0; r < d.n; r++) {
        uint16_t b = lo[(size_t)e * d.n + r];
        if (d.hi_cm && ((hi[(size_t)e * d.hi_words + (r >> 5)] >> (r & 31)) & 1u)) b |= 256;
        o[r] = b;
      }
    }
    return BST_OK;
  });
}

int bst_dataset_hist_layout(const bst_dataset* ds, int32_t* ne, uint32_t* e2o, uint32_t* enb,
                            uint32_t* eoff) {
  return guard([&] {
    need(ds != nullptr, BST_E_INVALID, "null dataset");
    uint32_t off = 0;
    if (ne) *ne = ds->d.E;
    for (int32_t e = 0; e < ds->d.E; e++) {
      if (e2o) e2o[e] = ds->eff_to_orig[e];
      if (enb) enb[e] = ds->fmeta[e].n_bins;
      if (eoff) eoff[e] = off;
      off += ds->fmeta[e].n_bins;
    }
    return (int)off;
  });
}

static int train_impl(bst_ctx* ctx, bst_dataset* ds, const float* y_dev, int64_t n_targets,
                      const float* w_dev, const float* vX_dev, const float* vy_dev, int64_t nv,
                      const bst_gbdt_config* cfg, const bst_ext_config* ext, bst_model** out) {
  need(ctx && ds && y_dev && cfg && out, BST_E_INVALID, "null argument");
  validate(cfg);
  if (n_targets < ds->d.n) raise(BST_E_TARGETS, "fewer targets than rows");
  TrainOptions opt;
  opt.cfg = *cfg;
  bst_ext_config_default(&opt.ext);
  if (ext) opt.ext = *ext;
  opt.K = n_outputs_of(*cfg);
  opt.obj = obj_params_of(*cfg);
  if (cfg->alphas && cfg->n_alphas > 0) opt.alphas.assign(cfg->alphas, cfg->alphas + cfg->n_alphas);
  opt.cfg.alphas = nullptr;  // borrowed for the call only: the trainer keeps opt.alphas
  need(opt.ext.schedule >= BST_SCHED_PLAIN && opt.ext.schedule <= BST_SCHED_GRAPH, BST_E_CONFIG,
       "unknown schedule");
  need(opt.ext.hist_fg == 0 || opt.ext.hist_fg == 16 || opt.ext.hist_fg == 32, BST_E_CONFIG,
       "hist_fg must be 0, 16 or 32");
  {
    const int32_t r = opt.ext.hist_rows_min;
    need(r == 0 || (r >= 512 && r <= 8192 && (r & (r - 1)) == 0), BST_E_CONFIG,
         "hist_rows_min must be 0 or a power of two in 512..8192");
  }
  need(opt.ext.hist_blocks_per_cu >= 0 && opt.ext.hist_blocks_per_cu <= 8, BST_E_CONFIG,
       "hist_blocks_per_cu must be in 0..8");
  need(opt.ext.hist_priv >= 0 && opt.ext.hist_priv <= 2, BST_E_CONFIG,
       "hist_priv must be 0 (tuned), 1 (LDS) or 2 (global memory)");
  if (opt.ext.goss_top_rate > 0.0f) {  // RowSamplingParams::Goss
    need(opt.ext.goss_top_rate < 1.0f && opt.ext.goss_other_rate > 0.0f &&
             opt.ext.goss_other_rate <= 1.0f &&
             opt.ext.goss_top_rate + opt.ext.goss_other_rate <= 1.0f,
         BST_E_CONFIG, "GOSS rates: 0 < top < 1, 0 < other <= 1, top + other <= 1");
    need(ds->d.n >= 2, BST_E_CONFIG, "GOSS needs at least two rows");
  }
  if (cfg->objective == BST_OBJ_LAMBDARANK) {
    // Query groups: ragged, in row order, summing to the row count.
    need(opt.ext.query_group_sizes && opt.ext.n_query_groups >= 1, BST_E_CONFIG,
         "LambdaRank needs query_group_sizes");
    opt.group_off.assign((size_t)opt.ext.n_query_groups + 1, 0);
    int64_t max_group = 1;
    for (int64_t q = 0; q < opt.ext.n_query_groups; q++) {
      const int32_t sz = opt.ext.query_group_sizes[q];
      need(sz >= 1, BST_E_CONFIG, "empty query group");
      need(sz <= kRankMaxGroupRows, BST_E_CONFIG, "query group larger than 4096 rows");
      opt.group_off[q + 1] = opt.group_off[q] + sz;
      max_group = std::max<int64_t>(max_group, sz);
    }
    need(opt.group_off.back() == ds->d.n, BST_E_SHAPE,
         "query group sizes must sum to the row count");
    opt.rank_disc.resize((size_t)max_group);
    for (size_t r = 0; r < opt.rank_disc.size(); r++)
      opt.rank_disc[r] = 1.0 / std::log2(2.0 + (double)r);
    opt.ext.query_group_sizes = nullptr;  // not kept past this call
  }
  const bool trace = std::getenv("BST_TRACE_SETUP") != nullptr;
  auto ms_since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
  };
  const auto tc0 = std::chrono::steady_clock::now();
  bst_model* m = nullptr;
  std::chrono::steady_clock::time_point td;
  {
    // The trainer's buffers come from (and go back to) the context's pool.
    PoolScope pool_scope(&ctx->pool);
    Trainer tr(ctx, ds, opt);
    if (trace) {
      BST_HIP(hipStreamSynchronize(ctx->stream));
      std::fprintf(stderr, "[bst] trainer buffers %.3f ms\n", ms_since(tc0));
    }
    const auto tt = std::chrono::steady_clock::now();
    m = tr.train(y_dev, w_dev, vX_dev, vy_dev, nv);
    if (trace) std::fprintf(stderr, "[bst] train %.3f ms\n", ms_since(tt));
    ctx->last_ms_total = tr.stats.ms_total;
    ctx->last_ms_per_tree = tr.stats.n_trees ? tr.stats.ms_total / (double)tr.stats.n_trees : 0.0;
    ctx->last_hist_rows = tr.stats.hist_rows;
    ctx->last_metric = tr.metric_history;
    td = std::chrono::steady_clock::now();
  }
  if (trace) std::fprintf(stderr, "[bst] trainer teardown %.3f ms\n", ms_since(td));
  *out = m;
  return BST_OK;
}

int bst_train(bst_ctx* ctx, bst_dataset* ds, const float* targets, int64_t n_targets,
              const float* weights, const float* vX, const float* vy, int64_t nv,
              const bst_gbdt_config* cfg, const bst_ext_config* ext, bst_model** out) {
  return guard([&] {
    need(ctx && ds && targets, BST_E_INVALID, "null argument");
    set_device(ctx);
    const int64_t n = ds->d.n;
    if (n_targets < n) raise(BST_E_TARGETS, "fewer targets than rows");
    DevBuf<float> y(n), w, dvX, dvy;
    h2d(ctx->stream, y.p, targets, n);
    if (weights) {
      w.alloc(n);
      h2d(ctx->stream, w.p, weights, n);
    }
    if (vX && vy && nv > 0) {
      dvX.alloc((size_t)nv * ds->d.F);
      dvy.alloc(nv);
      h2d(ctx->stream, dvX.p, vX, (size_t)nv * ds->d.F);
      h2d(ctx->stream, dvy.p, vy, nv);
    }
    BST_HIP(hipStreamSynchronize(ctx->stream));
    return train_impl(ctx, ds, y.p, n_targets, w.p, dvX.p, dvy.p, (vX && vy) ? nv : 0, cfg, ext,
                      out);
  });
}

int bst_train_device(bst_ctx* ctx, bst_dataset* ds, const float* y, int64_t n_targets,
                     const float* w, const bst_gbdt_config* cfg, const bst_ext_config* ext,
                     bst_model** out) {
  return guard([&] {
    need(ctx && ds, BST_E_INVALID, "null argument");
    set_device(ctx);
    return train_impl(ctx, ds, y, n_targets, w, nullptr, nullptr, 0, cfg, ext, out);
  });
}

int bst_ctx_set_keep_predictions(bst_ctx* ctx, int32_t enable) {
  return guard([&] {
    need(ctx != nullptr, BST_E_INVALID, "null ctx");
    ctx->keep_preds = enable != 0;
    if (!ctx->keep_preds) ctx->last_preds.clear();
    return BST_OK;
  });
}

int bst_last_train_predictions(bst_ctx* ctx, float* out, int64_t cap, int64_t* count) {
  return guard([&] {
    need(ctx != nullptr, BST_E_INVALID, "null ctx");
    const int64_t k = (int64_t)ctx->last_preds.size();
    if (count) *count = k;
    if (out) {
      need(cap >= k, BST_E_SHAPE, "output buffer smaller than the kept predictions");
      std::memcpy(out, ctx->last_preds.data(), (size_t)k * sizeof(float));
    }
    return BST_OK;
  });
}

int bst_last_train_stats(bst_ctx* ctx, double* ms_total, double* ms_per_tree, int64_t* hr) {
  return guard([&] {
    need(ctx != nullptr, BST_E_INVALID, "null ctx");
    if (ms_total) *ms_total = ctx->last_ms_total;
    if (ms_per_tree) *ms_per_tree = ctx->last_ms_per_tree;
    if (hr) *hr = ctx->last_hist_rows;
    return BST_OK;
  });
}

int bst_last_train_tuning(bst_ctx* ctx, int32_t* hist_fg, int32_t* hist_rows_min,
                          int32_t* hist_blocks_per_cu, int32_t* hist_priv, int32_t* tuned,
                          double* tune_ms) {
  return guard([&] {
    need(ctx, BST_E_INVALID, "null context");
    const HistTuning& t = ctx->last_tuning;
    if (hist_fg) *hist_fg = t.fg;
    if (hist_rows_min) *hist_rows_min = (int32_t)t.tile_min;
    if (hist_blocks_per_cu) *hist_blocks_per_cu = (int32_t)t.blocks_per_cu;
    if (hist_priv) *hist_priv = (int32_t)t.priv;
    if (tuned) *tuned = t.tuned;
    if (tune_ms) *tune_ms = t.tune_ms;
    return BST_OK;
  });
}

int bst_last_train_metric(bst_ctx* ctx, double* out, int64_t cap, int64_t* nr) {
  return guard([&] {
    need(ctx != nullptr, BST_E_INVALID, "null ctx");
    for (int64_t i = 0; i < (int64_t)ctx->last_metric.size() && i < cap; i++)
      out[i] = ctx->last_metric[i];
    if (nr) *nr = (int64_t)ctx->last_metric.size();
    return BST_OK;
  });
}

int bst_model_info_get(const bst_model* m, bst_model_info* info) {
  return guard([&] {
    need(m && info, BST_E_INVALID, "null argument");
    *info = m->info;
    return BST_OK;
  });
}

int bst_model_export(const bst_model* m, bst_forest_arrays* o) {
  return guard([&] {
    need(m && o, BST_E_INVALID, "null argument");
    std::copy(m->base_score.begin(), m->base_score.end(), o->base_score);
    std::copy(m->tree_offset.begin(), m->tree_offset.end(), o->tree_offset);
    std::copy(m->tree_group.begin(), m->tree_group.end(), o->tree_group);
    std::copy(m->split_index.begin(), m->split_index.end(), o->split_index);
    std::copy(m->split_threshold.begin(), m->split_threshold.end(), o->split_threshold);
    std::copy(m->left.begin(), m->left.end(), o->left);
    std::copy(m->right.begin(), m->right.end(), o->right);
    std::copy(m->default_left.begin(), m->default_left.end(), o->default_left);
    std::copy(m->is_leaf.begin(), m->is_leaf.end(), o->is_leaf);
    std::copy(m->split_type.begin(), m->split_type.end(), o->split_type);
    std::copy(m->leaf_value.begin(), m->leaf_value.end(), o->leaf_value);
    std::copy(m->gain.begin(), m->gain.end(), o->gain);
    std::copy(m->cover.begin(), m->cover.end(), o->cover);
    std::copy(m->cat_start.begin(), m->cat_start.end(), o->cat_start);
    std::copy(m->cat_size.begin(), m->cat_size.end(), o->cat_size);
    if (o->cat_words) std::copy(m->cat_words.begin(), m->cat_words.end(), o->cat_words);
    return BST_OK;
  });
}

int bst_model_import(bst_ctx* ctx, const bst_model_info* info, const bst_forest_arrays* in,
                     bst_model** out) {
  return guard([&] {
    need(info && in && out, BST_E_INVALID, "null argument");
    need(info->n_trees >= 0 && info->n_groups >= 1 && info->n_groups <= kMaxOutputs &&
             info->total_nodes >= 0 && info->total_cat_words >= 0 && info->n_features >= 0,
         BST_E_FORMAT, "bad model shape");
    std::unique_ptr<bst_model> m(new bst_model());
    m->info = *info;
    const int64_t T = info->n_trees, N = info->total_nodes, W = info->total_cat_words;
    m->base_score.assign(in->base_score, in->base_score + info->n_groups);
    m->tree_offset.assign(in->tree_offset, in->tree_offset + T + 1);
    m->tree_group.assign(in->tree_group, in->tree_group + T);
    m->split_index.assign(in->split_index, in->split_index + N);
    m->split_threshold.assign(in->split_threshold, in->split_threshold + N);
    m->left.assign(in->left, in->left + N);
    m->right.assign(in->right, in->right + N);
    m->default_left.assign(in->default_left, in->default_left + N);
    m->is_leaf.assign(in->is_leaf, in->is_leaf + N);
    m->split_type.assign(in->split_type, in->split_type + N);
    m->leaf_value.assign(in->leaf_value, in->leaf_value + N);
    // Gains and covers are optional (tree.rs:146-161): absent means zeros.
    if (in->gain) m->gain.assign(in->gain, in->gain + N); else m->gain.assign(N, 0.0f);
    if (in->cover) m->cover.assign(in->cover, in->cover + N); else m->cover.assign(N, 0.0f);
    m->cat_start.assign(in->cat_start, in->cat_start + N);
    m->cat_size.assign(in->cat_size, in->cat_size + N);
    if (W > 0) m->cat_words.assign(in->cat_words, in->cat_words + W);
    check_forest(*m);
    (void)ctx;
    *out = m.release();
    return BST_OK;
  });
}

int bst_model_destroy(bst_model* m) {
  return guard([&] {
    delete m;
    return BST_OK;
  });
}

static void predict_dev(bst_ctx* ctx, const bst_model* mc, const float* Xd, int64_t n, int32_t F,
                        int32_t raw, float* outd, uint32_t* leafd) {
  auto* m = const_cast<bst_model*>(mc);
  need(F >= m->info.n_features, BST_E_SHAPE, "fewer features than the model was trained on");
  if (m->device_for != ctx->device) {
    m->upload(ctx->stream);
    m->device_for = ctx->device;
  }
  const PredModel pm = m->view();
  int64_t scratch_floats = 0;
  float* scratch_p = nullptr;
  if (pred_two_phase(pm, n, ctx->n_cu, ctx->lds_per_block, &scratch_floats)) {
    // The context's grow-only leaf buffer; earlier work on the stream may
    // still read the old one, so the stream drains before it is replaced.
    if ((size_t)scratch_floats > ctx->pred_leaf.n) {
      BST_HIP(hipStreamSynchronize(ctx->stream));
      ctx->pred_leaf.alloc((size_t)scratch_floats);
    }
    scratch_p = ctx->pred_leaf.p;
  }
  launch_predict(ctx->stream, pm, Xd, n, n, F, raw ? 0 : m->info.output_transform, outd, n, leafd,
                 scratch_p, ctx->lds_per_block, ctx->n_cu);
  BST_HIP(hipGetLastError());
}

int bst_predict(bst_ctx* ctx, const bst_model* m, const float* X, int64_t n, int32_t F,
                int32_t raw, float* out, uint32_t* leaf_idx) {
  return guard([&] {
    need(ctx && m && X && out, BST_E_INVALID, "null argument");
    set_device(ctx);
    if (n == 0) return BST_OK;
    DevBuf<float> dX((size_t)n * F), dout((size_t)n * m->info.n_groups);
    DevBuf<uint32_t> dleaf;
    if (leaf_idx) dleaf.alloc((size_t)n * std::max<int64_t>(m->info.n_trees, 1));
    h2d(ctx->stream, dX.p, X, (size_t)n * F);
    predict_dev(ctx, m, dX.p, n, F, raw, dout.p, leaf_idx ? dleaf.p : nullptr);
    d2h(ctx->stream, out, dout.p, dout.n);
    if (leaf_idx) d2h(ctx->stream, leaf_idx, dleaf.p, (size_t)n * m->info.n_trees);
    BST_HIP(hipStreamSynchronize(ctx->stream));
    return BST_OK;
  });
}

int bst_predict_device(bst_ctx* ctx, const bst_model* m, const float* Xd, int64_t n, int32_t F,
                       int32_t raw, float* outd, uint32_t* leafd) {
  return guard([&] {
    need(ctx && m && Xd && outd, BST_E_INVALID, "null argument");
    set_device(ctx);
    if (n == 0) return BST_OK;
    predict_dev(ctx, m, Xd, n, F, raw, outd, leafd);
    return BST_OK;
  });
}

int bst_predict_device_noalloc(bst_ctx* ctx, const bst_model* m, const float* Xd, int64_t n,
                               int32_t F, int32_t raw, float* outd, uint32_t* leafd) {
  return guard([&] {
    need(ctx && m && Xd && outd, BST_E_INVALID, "null argument");
    need(n >= 0, BST_E_SHAPE, "negative row count");
    need(m->device_for == ctx->device && m->d_nodes.p, BST_E_INVALID,
         "model is not resident on this device: predict with it once through bst_predict or "
         "bst_predict_device (which uploads it) before the allocation-free call");
    need(F >= m->info.n_features, BST_E_SHAPE, "fewer features than the model was trained on");
    need(ctx->noalloc_scratch != nullptr, BST_E_DEVICE, "scratch not set up on this device");
    // A row of the two-phase launch needs one float per tree.
    need(m->info.n_trees <= (int64_t)kNoAllocScratchFloats, BST_E_OOM,
         "forest has more trees than the allocation-free scratch holds leaf values for "
         "(8388608 per row)");
    BST_HIP(hipSetDevice(ctx->device));
    if (n == 0) return BST_OK;
    const PredModel pm = m->view();
    const int transform = raw ? 0 : m->info.output_transform;
    if (!pred_two_phase(pm, n, ctx->n_cu, ctx->lds_per_block, nullptr)) {
      launch_predict(ctx->stream, pm, Xd, n, n, F, transform, outd, n, leafd, nullptr,
                     ctx->lds_per_block, ctx->n_cu);
    } else {
      // Row batches whose leaf values fit the scratch, one after another on
      // the ctx stream (each reuses the scratch after the previous one's
      // reads); every row's sum is the same as in one launch.
      const int64_t per = std::max<int64_t>(
          1, std::min<int64_t>(n, (int64_t)kNoAllocScratchFloats / std::max<int64_t>(pm.n_trees, 1)));
      for (int64_t r0 = 0; r0 < n; r0 += per) {
        const int64_t nb = std::min(per, n - r0);
        launch_predict_rows(ctx->stream, pm, Xd + r0, nb, n, F, transform, outd + r0, n,
                            leafd ? leafd + r0 : nullptr, n, ctx->noalloc_scratch,
                            ctx->lds_per_block, ctx->n_cu);
      }
    }
    BST_HIP(hipGetLastError());
    return BST_OK;
  });
}

// ---- component entry points ------------------------------------------------
int bst_kernel_lambdarank(bst_ctx* ctx, const float* preds, const float* y, const float* w,
                          int64_t n, const int32_t* sizes, int64_t n_groups, float* g, float* h) {
  return guard([&] {
    need(ctx && preds && y && g && h && sizes, BST_E_INVALID, "null argument");
    need(n_groups >= 1, BST_E_CONFIG, "no query groups");
    std::vector<int64_t> off((size_t)n_groups + 1, 0);
    int64_t max_group = 1;
    for (int64_t q = 0; q < n_groups; q++) {
      need(sizes[q] >= 1 && sizes[q] <= kRankMaxGroupRows, BST_E_CONFIG,
           "query group sizes must be 1..4096");
      off[q + 1] = off[q] + sizes[q];
      max_group = std::max<int64_t>(max_group, sizes[q]);
    }
    need(off.back() == n, BST_E_SHAPE, "query group sizes must sum to n");
    std::vector<double> disc((size_t)max_group);
    for (size_t r = 0; r < disc.size(); r++) disc[r] = 1.0 / std::log2(2.0 + (double)r);
    set_device(ctx);
    hipStream_t s = ctx->stream;
    DevBuf<float> dp(n), dy(n), dw, dg(n), dh(n);
    DevBuf<int64_t> doff(off.size());
    DevBuf<double> ddisc(disc.size());
    DevBuf<uint32_t> am(2);
    h2d(s, dp.p, preds, n);
    h2d(s, dy.p, y, n);
    if (w) {
      dw.alloc(n);
      h2d(s, dw.p, w, n);
    }
    h2d(s, doff.p, off.data(), off.size());
    h2d(s, ddisc.p, disc.data(), disc.size());
    BST_HIP(hipMemsetAsync(am.p, 0, am.bytes(), s));
    launch_lambdarank(s, dp.p, dy.p, dw.p, doff.p, n_groups, ddisc.p, dg.p, dh.p, am.p);
    d2h(s, g, dg.p, n);
    d2h(s, h, dh.p, n);
    BST_HIP(hipStreamSynchronize(s));
    return BST_OK;
  });
}

namespace {
// One gradient launch over host buffers (the per-kernel parity entry points).
int gradients_impl(bst_ctx* ctx, int32_t obj, int32_t K, const ObjParams& op, const float* preds,
                   const float* y, const float* w, int64_t n, float* g, float* h) {
  need(ctx && preds && y && g && h, BST_E_INVALID, "null argument");
  need(n >= 0, BST_E_INVALID, "negative row count");
  set_device(ctx);
  hipStream_t s = ctx->stream;
  DevBuf<float> dp((size_t)K * n), dy(n), dw, dg((size_t)K * n), dh((size_t)K * n);
  DevBuf<uint32_t> am(2 * K);
  h2d(s, dp.p, preds, (size_t)K * n);
  h2d(s, dy.p, y, n);
  if (w) {
    dw.alloc(n);
    h2d(s, dw.p, w, n);
  }
  BST_HIP(hipMemsetAsync(am.p, 0, am.bytes(), s));
  launch_gradients(s, obj, op, dp.p, dy.p, dw.p, n, K, dg.p, dh.p, am.p);
  d2h(s, g, dg.p, dg.n);
  d2h(s, h, dh.p, dh.n);
  BST_HIP(hipStreamSynchronize(s));
  return BST_OK;
}
}  // namespace

int bst_kernel_gradients(bst_ctx* ctx, int32_t obj, int32_t K, const float* preds,
                         const float* y, const float* w, int64_t n, float* g, float* h) {
  return guard([&] {
    need(obj >= 0 && obj <= BST_OBJ_POISSON && obj != BST_OBJ_LAMBDARANK, BST_E_CONFIG,
         "objective");
    need(obj != BST_OBJ_PINBALL, BST_E_CONFIG,
         "pinball needs its alphas: use bst_kernel_gradients_cfg");
    need(K >= 1 && K <= kMaxOutputs && (obj == BST_OBJ_SOFTMAX || K == 1), BST_E_CONFIG,
         "outputs must be 1, or 2..1024 for softmax");
    return gradients_impl(ctx, obj, K, ObjParams(), preds, y, w, n, g, h);
  });
}

int bst_kernel_gradients_cfg(bst_ctx* ctx, const bst_gbdt_config* cfg, const float* preds,
                             const float* y, const float* w, int64_t n, float* g, float* h) {
  return guard([&] {
    need(cfg != nullptr, BST_E_INVALID, "null config");
    validate(cfg);
    need(cfg->objective != BST_OBJ_LAMBDARANK, BST_E_CONFIG,
         "LambdaRank gradients: use bst_kernel_lambdarank");
    const int K = n_outputs_of(*cfg);
    return gradients_impl(ctx, cfg->objective, K, obj_params_of(*cfg), preds, y, w, n, g, h);
  });
}

namespace {
int scale_exp(double maxabs, int64_t n) {
  if (!(maxabs > 0.0) || !std::isfinite(maxabs)) return 0;
  int lg = 0;
  while ((1LL << lg) < n && lg < 62) lg++;
  int x;
  std::frexp(maxabs, &x);
  return std::min(62 - lg - x, 1000);
}
}  // namespace

int bst_kernel_histogram(bst_ctx* ctx, const bst_dataset* ds, const float* grad,
                         const float* hess, const uint32_t* rows, int64_t n_rows, double* out,
                         int64_t out_len) {
  return guard([&] {
    need(ctx && ds && grad && hess && out, BST_E_INVALID, "null argument");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    const auto& d = ds->d;
    const int64_t m = rows ? n_rows : d.n;
    std::vector<uint32_t> idx(std::max<int64_t>(m, 1));
    std::vector<float2> ogh(std::max<int64_t>(m, 1));
    double mg = 0, mh = 0;
    for (int64_t i = 0; i < m; i++) {
      const uint32_t r = rows ? rows[i] : (uint32_t)i;
      need(r < d.n, BST_E_SHAPE, "row id out of range");
      idx[i] = r;
      ogh[i] = make_float2(grad[r], hess[r]);
      mg = std::max(mg, (double)std::fabs(grad[r]));
      mh = std::max(mh, (double)std::fabs(hess[r]));
    }
    const int32_t qe[2] = {scale_exp(mg, std::max<int64_t>(m, 1)),
                           scale_exp(mh, std::max<int64_t>(m, 1))};
    const int64_t slot_elems = (int64_t)d.NG * 256 * d.FG * 2;
    DevBuf<uint32_t> didx(idx.size());
    DevBuf<float2> dogh(ogh.size());
    DevBuf<int32_t> dq(2);
    DevBuf<long long> dhist(slot_elems), dsum(2);
    h2d(s, didx.p, idx.data(), idx.size());
    h2d(s, dogh.p, ogh.data(), ogh.size());
    h2d(s, dq.p, qe, 2);
    BST_HIP(hipMemsetAsync(dhist.p, 0, dhist.bytes(), s));
    BST_HIP(hipMemsetAsync(dsum.p, 0, dsum.bytes(), s));
    std::vector<HistItem> items;
    const uint32_t tile = 8192;
    for (int64_t b = 0; b < m; b += tile)
      for (int g = 0; g < d.NG; g++)
        items.push_back({0u, 0u, (uint32_t)b, (uint32_t)std::min<int64_t>(tile, m - b), (uint32_t)g, 0u});
    DevBuf<HistItem> ditems(std::max<size_t>(items.size(), 1));
    h2d(s, ditems.p, items.data(), items.size());
    launch_hist(s, d, ditems.p, (int)items.size(), didx.p, dogh.p, 0, dq.p, dhist.p, slot_elems,
                dsum.p, true, 0);
    BST_HIP(hipGetLastError());
    std::vector<long long> hh(slot_elems), ss(2);
    d2h(s, hh.data(), dhist.p, hh.size());
    d2h(s, ss.data(), dsum.p, 2);
    BST_HIP(hipStreamSynchronize(s));
    int64_t off = 0;
    for (int32_t e = 0; e < d.E; e++) {
      const uint32_t nb = ds->fmeta[e].n_bins;
      const int grp = e / d.FG, lf = e % d.FG;
      long long sg = 0, sh = 0;
      for (uint32_t b = 0; b < nb; b++) {
        long long vg, vh;
        if (b < 256) {
          const int64_t k = ((int64_t)(grp * 256 + (int)b) * d.FG + lf) * 2;
          vg = hh[k];
          vh = hh[k + 1];
          sg += vg;
          sh += vh;
        } else {
          vg = ss[0] - sg;
          vh = ss[1] - sh;
        }
        if (2 * (off + b) + 1 < out_len) {
          out[2 * (off + b)] = std::ldexp((double)vg, -qe[0]);
          out[2 * (off + b) + 1] = std::ldexp((double)vh, -qe[1]);
        }
      }
      off += nb;
    }
    return BST_OK;
  });
}

int bst_kernel_find_split(bst_ctx* ctx, const bst_dataset* ds, const double* hist, double pg,
                          double ph, uint32_t pc, const bst_gbdt_config* cfg, int32_t n_threads,
                          bst_split_info* out) {
  return guard([&] {
    need(ctx && ds && hist && cfg && out, BST_E_INVALID, "null argument");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    const auto& d = ds->d;
    // Fixed-point conversion of the given f64 histogram.
    double mg = std::fabs(pg), mh = std::fabs(ph);
    int64_t off = 0, nbt = 0;
    for (int32_t e = 0; e < d.E; e++) {
      const uint32_t nb = ds->fmeta[e].n_bins;
      double sg = 0, sh = 0;
      for (uint32_t b = 0; b < nb; b++) {
        sg += std::fabs(hist[2 * (off + b)]);
        sh += std::fabs(hist[2 * (off + b) + 1]);
      }
      mg = std::max(mg, sg);
      mh = std::max(mh, sh);
      off += nb;
      nbt += nb;
    }
    const int32_t qe[2] = {scale_exp(mg, 2), scale_exp(mh, 2)};
    const int64_t slot_elems = (int64_t)d.NG * 256 * d.FG * 2;
    std::vector<long long> hh(slot_elems, 0);
    off = 0;
    for (int32_t e = 0; e < d.E; e++) {
      const uint32_t nb = ds->fmeta[e].n_bins;
      const int grp = e / d.FG, lf = e % d.FG;
      for (uint32_t b = 0; b < nb && b < 256; b++) {
        const int64_t k = ((int64_t)(grp * 256 + (int)b) * d.FG + lf) * 2;
        hh[k] = std::llrint(std::ldexp(hist[2 * (off + b)], qe[0]));
        hh[k + 1] = std::llrint(std::ldexp(hist[2 * (off + b) + 1], qe[1]));
      }
      off += nb;
    }
    const long long sums[2] = {std::llrint(std::ldexp(pg, qe[0])), std::llrint(std::ldexp(ph, qe[1]))};
    DevBuf<long long> dhist(slot_elems), dsum(2);
    DevBuf<int32_t> dq(2);
    DevBuf<uint32_t> dcount(1), dnodes(2), dfeats(d.E);
    DevBuf<unsigned long long> dkey(1);
    DevBuf<DevSplit> dcand(d.E), dsplit(1);
    std::vector<uint32_t> feats(d.E);
    for (int32_t e = 0; e < d.E; e++) feats[e] = e;
    const uint32_t ns[2] = {0u, 0u};
    h2d(s, dhist.p, hh.data(), hh.size());
    h2d(s, dsum.p, sums, 2);
    h2d(s, dq.p, qe, 2);
    h2d(s, dcount.p, &pc, 1);
    h2d(s, dnodes.p, ns, 2);
    h2d(s, dfeats.p, feats.data(), feats.size());
    BST_HIP(hipMemsetAsync(dkey.p, 0, dkey.bytes(), s));
    SplitParams p;
    p.lambda = cfg->lambda;
    p.min_gain = cfg->min_gain;
    p.min_child_weight = cfg->min_child_weight;
    p.min_samples_leaf = cfg->min_samples_leaf;
    p.max_onehot_cats = cfg->max_onehot_cats;
    p.rightmost = (n_threads != 1 && d.E >= 16) ? 1 : 0;
    launch_split(s, d, dnodes.p, dnodes.p + 1, 1, dfeats.p, d.E, dhist.p, slot_elems, dsum.p,
                 dcount.p, dq.p, 1u << 30, p, dkey.p, dcand.p, dsplit.p);
    BST_HIP(hipGetLastError());
    DevSplit r;
    d2h(s, &r, dsplit.p, 1);
    BST_HIP(hipStreamSynchronize(s));
    out->feature = r.feature;
    out->gain = r.gain;
    out->default_left = (r.flags & 1u) ? 1 : 0;
    out->is_categorical = (r.flags & 2u) ? 1 : 0;
    out->bin = r.bin;
    for (int w = 0; w < 8; w++) out->left_cats[w] = r.cats[w];
    (void)nbt;
    return BST_OK;
  });
}

int bst_kernel_partition(bst_ctx* ctx, const bst_dataset* ds, uint32_t* rows, int64_t n,
                         int32_t e, uint32_t bin, int32_t default_left, const uint32_t* cats,
                         int32_t n_words, int64_t* left_count) {
  return guard([&] {
    need(ctx && ds && rows, BST_E_INVALID, "null argument");
    need(e >= 0 && e < ds->d.E, BST_E_INVALID, "bad feature");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    const int64_t m = std::max<int64_t>(n, 1);
    DevBuf<uint32_t> idx(m), idx_tmp(m), tl, toff, il(1), cnt(2);
    DevBuf<float2> ogh(m), ogh_tmp(m);
    DevBuf<uint8_t> flag(m);
    DevBuf<long long> sums(4);
    DevBuf<DevSplit> split(1);
    for (int64_t i = 0; i < n; i++) need(rows[i] < ds->d.n, BST_E_SHAPE, "row id out of range");
    h2d(s, idx.p, rows, n);
    BST_HIP(hipMemsetAsync(ogh.p, 0, ogh.bytes(), s));
    BST_HIP(hipMemsetAsync(sums.p, 0, sums.bytes(), s));
    DevSplit sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.gain = 1.0f;
    sp.feature = (uint32_t)e;
    sp.bin = bin;
    sp.flags = (default_left ? 1u : 0u) | (cats ? 2u : 0u);
    for (int w = 0; cats && w < n_words && w < kMaxCatWords; w++) sp.cats[w] = cats[w];
    h2d(s, split.p, &sp, 1);
    std::vector<PartTile> tiles;
    const uint32_t tile = 4096;
    for (int64_t b = 0; b < n || tiles.empty(); b += tile)
      tiles.push_back({0u, (uint32_t)b, (uint32_t)std::max<int64_t>(0, std::min<int64_t>(tile, n - b)), 0u});
    PartItem it{0u, 0u, (uint32_t)n, 0u, (uint32_t)tiles.size(), 1u, (uint32_t)e, 0u};
    DevBuf<PartItem> ditem(1);
    DevBuf<PartTile> dtiles(tiles.size());
    tl.alloc(tiles.size());
    toff.alloc(tiles.size());
    h2d(s, ditem.p, &it, 1);
    h2d(s, dtiles.p, tiles.data(), tiles.size());
    launch_partition(s, ds->d, ditem.p, 1, dtiles.p, (int)tiles.size(), split.p, idx.p, ogh.p,
                     idx_tmp.p, ogh_tmp.p, flag.p, tl.p, toff.p, il.p, cnt.p, sums.p);
    BST_HIP(hipGetLastError());
    uint32_t L = 0;
    d2h(s, &L, il.p, 1);
    d2h(s, rows, idx.p, n);
    BST_HIP(hipStreamSynchronize(s));
    if (left_count) *left_count = L;
    return BST_OK;
  });
}

int bst_kernel_partition_fused(bst_ctx* ctx, const bst_dataset* ds, uint32_t* rows, int64_t n,
                               int32_t e, uint32_t bin, int32_t default_left,
                               const uint32_t* cats, int32_t n_words, int32_t mode,
                               int64_t* left_count) {
  return guard([&] {
    need(ctx && ds && rows, BST_E_INVALID, "null argument");
    need(e >= 0 && e < ds->d.E, BST_E_INVALID, "bad feature");
    need(n >= 0 && n < (int64_t)0x7FFFFFFF, BST_E_SHAPE, "row count");
    for (int64_t i = 0; i < n; i++) need(rows[i] < ds->d.n, BST_E_SHAPE, "row id out of range");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    const int64_t m = std::max<int64_t>(n, 1);
    DevBuf<uint32_t> idx0(m), idx1(m), cursor(2), dyn(1), tout(1), ticket(2);
    DevBuf<float2> ogh0(m), ogh1(m);
    DevBuf<DevSplit> split(1);
    DevBuf<float> preds(1);
    // The node's rows as the partition stores a right child (last position
    // first) with BST_PART_REVERSED: the same rows in the same order.
    const bool input_reversed = (mode & BST_PART_REVERSED) != 0;
    std::vector<uint32_t> src(rows, rows + n);
    if (input_reversed) std::reverse(src.begin(), src.end());
    BST_HIP(hipMemsetAsync(idx0.p, 0, idx0.bytes(), s));
    h2d(s, idx0.p, src.data(), n);
    BST_HIP(hipMemsetAsync(ogh0.p, 0, ogh0.bytes(), s));
    BST_HIP(hipMemsetAsync(cursor.p, 0, cursor.bytes(), s));
    BST_HIP(hipMemsetAsync(tout.p, 0, tout.bytes(), s));
    BST_HIP(hipMemsetAsync(ticket.p, 0, ticket.bytes(), s));
    DevSplit sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.gain = 1.0f;
    sp.feature = (uint32_t)e;
    sp.bin = bin;
    sp.flags = (default_left ? 1u : 0u) | (cats ? 2u : 0u);
    for (int w = 0; cats && w < n_words && w < kMaxCatWords; w++) sp.cats[w] = cats[w];
    h2d(s, split.p, &sp, 1);
    const uint32_t src_bits = input_reversed ? kTileRev : 0u;
    std::vector<PartTile> tiles;
    for (int64_t b = 0; b < n || tiles.empty(); b += kFusedTileRows) {
      const uint32_t b0 = (uint32_t)b;
      const uint32_t begin = input_reversed ? (uint32_t)(n - 1 - b) : b0;
      tiles.push_back({0u, n > 0 ? begin : 0u,
                       (uint32_t)std::max<int64_t>(0, std::min<int64_t>(kFusedTileRows, n - b)),
                       src_bits});
    }
    DevBuf<unsigned long long> tstat(tiles.size());
    BST_HIP(hipMemsetAsync(tstat.p, 0, tstat.bytes(), s));
    const FeatMeta& fm = ds->fmeta[(size_t)e];
    PartItem it;
    std::memset(&it, 0, sizeof(it));
    it.node = 0;
    it.begin = 0;
    it.count = (uint32_t)n;
    it.tile0 = 0;
    it.n_tiles = (uint32_t)tiles.size();
    it.right_node = 1;
    it.feature = (uint32_t)e;
    it.split_bin = bin;
    it.split_flags = sp.flags;
    it.miss_bin = fm.has_missing ? fm.default_bin : 0xFFFFFFFFu;
    DevBuf<PartItem> ditem(1);
    DevBuf<PartTile> dtiles(tiles.size());
    const uint32_t nt = (uint32_t)tiles.size();
    h2d(s, ditem.p, &it, 1);
    h2d(s, dtiles.p, tiles.data(), tiles.size());
    h2d(s, dyn.p, &nt, 1);
    // BST_PART_TICKETS: half as many blocks as tiles (rounded up).
    const int grid = (mode & BST_PART_TICKETS) ? (int)((nt + 1) / 2) : (int)nt;
    launch_partition_fused(s, ds->d, ditem.p, dtiles.p, grid, split.p, idx0.p, ogh0.p, idx1.p,
                           ogh1.p, cursor.p, tstat.p, ticket.p, ticket.p + 1, dyn.p, tout.p,
                           preds.p, 1, 1u << 30);
    BST_HIP(hipGetLastError());
    uint32_t cur[2] = {0, 0}, tk[2] = {0, 0};
    d2h(s, cur, cursor.p, 2);
    d2h(s, tk, ticket.p, 2);
    d2h(s, rows, idx1.p, n);
    BST_HIP(hipStreamSynchronize(s));
    if (tk[1]) raise(BST_E_DEVICE, "a partition tile's look-back did not finish within its wait bound");
    // The right child's rows are stored from the end of the range down:
    // hand them back in row order.
    std::reverse(rows + cur[0], rows + n);
    if (left_count) *left_count = cur[0];
    return BST_OK;
  });
}

// ---- serving (declared here, implemented in server.cpp) -------------------

}  // extern "C"
