// trainer.cpp — the boosting loop and tree growth, driving the device kernels.
//
// Control flow restates (paths relative to crates/boosters/src/):
//   GBDTTrainer::train          training/gbdt/trainer.rs:162-401
//   TreeGrower::grow            training/gbdt/grower.rs:265-342
//   process_candidate           grower.rs:346-481 (subtraction trick, child sums)
//   finalize_remaining          grower.rs:484-514
//   GrowthState                 training/gbdt/expansion.rs:112-231 (depth-wise level queue,
//                               leaf-wise max-heap by (gain, partition id), n_leaves++ on pop)
//   apply_split_to_builder      grower.rs:576-650 (next_up thresholds, categorical swap)
//   compute_leaf_weight         split/gain.rs:151-168
//   update_predictions          grower.rs:208-219 (partition path) / trainer.rs:331-339
//
// Schedules: PLAIN expands one node per launch sequence (partition -> sync ->
// histogram/subtract/split -> sync); BATCHED expands every chosen node of all
// trees grown together (a depth level, or one pop per class tree of a
// softmax round) with one launch per kernel.  Both produce identical trees:
// node processing order, partition ids and tie rules are the same.
#include "limits.h"
#include "trainer.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "grow_device.h"
#include "sample_keys.h"

namespace bst {

namespace {

constexpr uint32_t kHistTile = 8192;   // rows per histogram block
constexpr uint32_t kPartTile = 4096;   // rows per partition block
constexpr uint32_t kSegTile = 16384;   // rows per sums / leaf-update block
constexpr uint32_t kDirect = 0x80000000u;

// The k candidates 0..n-1 with the smallest (sampling key, index), in index
// order (the draw rule of every sampling decision; sample_keys.h).
std::vector<uint32_t> smallest_keys(uint32_t n, uint64_t base, uint32_t k) {
  std::vector<uint64_t> kv(n);
  for (uint32_t i = 0; i < n; i++) kv[i] = ((uint64_t)sample_key(base, i) << 32) | i;
  k = std::min(k, n);
  if (k < n) std::nth_element(kv.begin(), kv.begin() + k, kv.end());
  std::vector<uint32_t> out(k);
  for (uint32_t j = 0; j < k; j++) out[j] = (uint32_t)(kv[j] & 0xFFFFFFFFu);
  std::sort(out.begin(), out.end());
  return out;
}

DevSplit invalid_split() {
  DevSplit s;
  std::memset(&s, 0, sizeof(s));
  s.gain = -std::numeric_limits<float>::infinity();
  return s;
}

template <class T>
void upload(hipStream_t s, DevBuf<T>& d, const std::vector<T>& h) {
  if (h.empty()) return;
  d.ensure(h.size());
  h2d(s, d.p, h.data(), h.size());
}

// default_metric_for_objective, metrics/mod.rs:213-232.
int default_metric(int obj) {
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

// compute_weighted_quantile (objectives/regression.rs:307-333): pairs sorted
// by value (a stable sort), total and running weights summed in f32 in
// sorted order; the first value whose running weight reaches alpha * total.
float weighted_quantile(const float* v, const float* w, int64_t n, float alpha) {
  if (n == 0) return 0.0f;
  std::vector<std::pair<float, float>> sorted((size_t)n);
  for (int64_t i = 0; i < n; i++) sorted[(size_t)i] = {v[i], w ? w[i] : 1.0f};
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

}  // namespace



struct Trainer::TreeState {
  int t = 0;
  int output = 0;
  TreeBuild tb;
  std::vector<uint32_t> pbegin, pcount, pslot;
  uint32_t n_part = 1;
  std::vector<uint32_t> free_slots;
  std::vector<float> leaf_lr;
  std::vector<std::pair<uint32_t, uint32_t>> leaf_map;  // (tree node, partition id) per leaf
  std::vector<Cand> level, next;
  std::priority_queue<Cand, std::vector<Cand>, CandLess> heap;
  uint32_t n_leaves = 1;
  int64_t rows = 0;
  bool sampled = false;
};

Trainer::Trainer(bst_ctx* ctx, bst_dataset* ds, const TrainOptions& opt)
    : ctx_(ctx), ds_(ds), opt_(opt), s_(ctx->stream), n_(ds->d.n) {
  const auto& c = opt_.cfg;
  sched_ = opt_.ext.schedule;
  lr_ = c.learning_rate;
  const uint32_t max_leaves = c.growth == BST_GROW_DEPTHWISE ? (1u << c.max_depth) : c.max_leaves;
  MN_ = (uint32_t)std::max<uint64_t>(2ull * max_leaves - 1, 63) + 2;
  SPT_ = max_leaves + 2;
  const bool col_sampling = !(std::fabs(c.colsample_bytree - 1.0f) < 1e-6f) || level_sampling();
  // Class trees of a round grow together, except on the host-driven path with
  // column sampling (its split launch takes one feature list).
  // Positions of the trees of a round are 32-bit (t * n + i): a round whose
  // K * n does not fit grows its class trees one at a time.
  T_ = class_batch_trees(opt_.K, n_,
                         sched_ != BST_SCHED_PLAIN && (graph_capable() || !col_sampling));
  tune_ = choose_hist_tuning(ctx, ds->d, c, opt_.ext, T_);
  ctx->last_tuning = tune_;
  dd_ = ds->d;
  dd_.FG = tune_.fg;
  dd_.hist_priv = tune_.priv;
  dd_.NG = (dd_.E + dd_.FG - 1) / dd_.FG;
  slot_elems_ = (int64_t)dd_.NG * 256 * dd_.FG * 2;
  rightmost_ = opt_.ext.n_threads != 1 && ds->d.E >= 16;
  sp_.lambda = (double)c.lambda;
  sp_.min_gain = (double)c.min_gain;
  sp_.min_child_weight = (double)c.min_child_weight;
  sp_.min_samples_leaf = c.min_samples_leaf;
  sp_.max_onehot_cats = c.max_onehot_cats;
  sp_.rightmost = rightmost_ ? 1 : 0;

  const int K = opt_.K;
  preds_.alloc((size_t)K * n_);
  g_.alloc((size_t)K * n_);
  h_.alloc((size_t)K * n_);
  absmax_.alloc(2 * K);
  qexp_.alloc(2 * K);
  idx_.alloc((size_t)T_ * n_);
  idx_tmp_.alloc((size_t)T_ * n_);
  ogh_.alloc((size_t)T_ * n_);
  ogh_tmp_.alloc((size_t)T_ * n_);
  flag_.alloc((size_t)T_ * n_);
  node_sum_.alloc((size_t)2 * T_ * MN_);
  node_count_.alloc((size_t)T_ * MN_);
  node_key_.alloc((size_t)T_ * MN_);
  node_split_.alloc((size_t)T_ * MN_);
  cand_.alloc((size_t)T_ * MN_ * ds->d.E);
  hist_.alloc((size_t)T_ * SPT_ * slot_elems_);
  leaf_val_.alloc((size_t)T_ * MN_);
  tree_output_.alloc(T_);
  feats_.alloc(ds->d.E);
  BST_HIP(hipMemsetAsync(node_key_.p, 0, node_key_.bytes(), s_));
  if (c.objective == BST_OBJ_LAMBDARANK) {
    group_off_.alloc(opt_.group_off.size());
    h2d(s_, group_off_.p, opt_.group_off.data(), opt_.group_off.size());
    rank_disc_.alloc(opt_.rank_disc.size());
    h2d(s_, rank_disc_.p, opt_.rank_disc.data(), opt_.rank_disc.size());
  }
}

void Trainer::launch_objective(const float* y, const float* w) {
  const auto& cfg = opt_.cfg;
  if (cfg.objective == BST_OBJ_LAMBDARANK)
    launch_lambdarank(s_, preds_.p, y, w, group_off_.p, (int64_t)opt_.group_off.size() - 1,
                      rank_disc_.p, g_.p, h_.p, absmax_.p);
  else
    launch_gradients(s_, cfg.objective, opt_.obj, preds_.p, y, w, n_, opt_.K, g_.p, h_.p,
                     absmax_.p);
}

uint32_t Trainer::alloc_slot(TreeState& ts) {
  if (ts.free_slots.empty()) raise(BST_E_OOM, "histogram slots exhausted");
  const uint32_t s = ts.free_slots.back();
  ts.free_slots.pop_back();
  return s;
}

void Trainer::make_leaf(TreeState& ts, const Cand& c) {
  const double G = std::ldexp((double)c.g, -qexp_h_[2 * ts.output]);
  const double H = std::ldexp((double)c.h, -qexp_h_[2 * ts.output + 1]);
  const auto& cfg = opt_.cfg;
  float w;
  const double lam = (double)cfg.lambda, al = (double)cfg.alpha;
  if (al == 0.0) {
    w = (float)(-G / (H + lam));
  } else {
    const double ag = std::fabs(G);
    w = ag <= al ? 0.0f : (float)((G > 0.0 ? -1.0 : 1.0) * (ag - al) / (H + lam));
  }
  ts.tb.is_leaf[c.tree_node] = 1;
  ts.tb.leaf[c.tree_node] = w;
  ts.tb.cover[c.tree_node] = (float)H;
  ts.leaf_lr[c.node] = w * lr_;
  ts.leaf_map.push_back({c.tree_node, c.node});
  if (ts.pslot[c.node] != UINT32_MAX) {
    ts.free_slots.push_back(ts.pslot[c.node]);
    ts.pslot[c.node] = UINT32_MAX;
  }
}

// Root histograms and splits for every tree grown this round.
void Trainer::root(std::vector<TreeState>& trees) {
  std::vector<HistItem> items;
  std::vector<uint32_t> zero, nodes, slots;
  std::vector<uint32_t> counts((size_t)T_ * MN_, 0);
  bool gather = false;
  for (auto& ts : trees) {
    const uint32_t gnode = ts.t * MN_;
    ts.pslot[0] = alloc_slot(ts);
    const uint32_t gslot = ts.t * SPT_ + ts.pslot[0];
    const uint32_t m = (uint32_t)ts.rows;
    const uint32_t nt = std::max<uint32_t>(1, (m + kHistTile - 1) / kHistTile);
    if (nt > 1 || m == 0) zero.push_back(gslot);
    for (uint32_t k = 0; k < nt && m > 0; k++) {
      const uint32_t b = k * kHistTile, cnt = std::min(kHistTile, m - b);
      for (int gr = 0; gr < dd_.NG; gr++)
        items.push_back({gnode, gslot, (uint32_t)(ts.t * n_) + b, cnt, (uint32_t)gr,
                         (uint32_t)ts.t | (nt == 1 ? kDirect : 0u)});
    }
    counts[gnode] = m;
    nodes.push_back(gnode);
    slots.push_back(gslot);
    // Without gather the kernel takes the position as the row id, which only
    // holds for tree slot 0 (slot t's positions start at t * n).
    gather = gather || ts.sampled || ts.t > 0;
  }
  upload(s_, d_hist_items_, items);
  upload(s_, d_u32a_, zero);
  upload(s_, d_u32b_, nodes);
  upload(s_, d_u32c_, slots);
  h2d(s_, node_count_.p, counts.data(), counts.size());
  const int32_t* qexp = qexp_.p + 2 * trees[0].output;
  launch_zero_slots(s_, d_u32a_.p, (int)zero.size(), hist_.p, slot_elems_);
  double rows = 0;
  for (auto& ts : trees) rows += (double)ts.rows;
  int pi = ctx_->prof_begin(0, hist_bytes(rows, (int)trees.size(), gather));
  launch_hist(s_, dd_, d_hist_items_.p, (int)items.size(), idx_.p, ogh_.p, 0, qexp, hist_.p,
              slot_elems_, node_sum_.p, gather, 0);
  ctx_->prof_end(pi);
  pi = ctx_->prof_begin(1, split_bytes((int)nodes.size()));
  launch_split(s_, dd_, d_u32b_.p, d_u32c_.p, (int)nodes.size(), feats_.p, (int)feats_h_.size(),
               hist_.p, slot_elems_, node_sum_.p, node_count_.p, qexp, MN_, sp_, node_key_.p,
               cand_.p, node_split_.p);
  ctx_->prof_end(pi);
  std::vector<long long> sums(2 * nodes.size());
  std::vector<DevSplit> splits(nodes.size());
  for (size_t i = 0; i < nodes.size(); i++) {
    d2h(s_, &sums[2 * i], node_sum_.p + 2 * nodes[i], 2);
    d2h(s_, &splits[i], node_split_.p + nodes[i], 1);
  }
  qexp_h_.resize(2 * opt_.K);
  d2h(s_, qexp_h_.data(), qexp_.p, 2 * opt_.K);
  BST_HIP(hipStreamSynchronize(s_));
  ctx_->prof_drain();
  stats.hist_rows += 0;
  for (size_t i = 0; i < trees.size(); i++) {
    auto& ts = trees[i];
    stats.hist_rows += ts.rows;
    const uint32_t root = ts.tb.alloc();
    Cand c{0, root, 0, splits[i], sums[2 * i], sums[2 * i + 1], (uint32_t)ts.rows};
    if (opt_.cfg.growth == BST_GROW_DEPTHWISE)
      ts.level.push_back(c);
    else
      ts.heap.push(c);
  }
}

namespace {
struct Pending {
  int t;
  Cand c;
  uint32_t lt, rt, right;
};
}  // namespace

void Trainer::expand_batch(std::vector<TreeState>& trees,
                           std::vector<std::pair<int, Cand>>& chosen) {
  const auto& cfg = opt_.cfg;
  std::vector<PartItem> pitems;
  std::vector<PartTile> ptiles;
  std::vector<Pending> pend;
  for (auto& tc : chosen) {
    TreeState& ts = trees[tc.first];
    const Cand& c = tc.second;
    if (!(c.split.gain > -std::numeric_limits<float>::infinity()) || c.split.gain <= cfg.min_gain) {
      make_leaf(ts, c);
      continue;
    }
    const double H = std::ldexp((double)c.h, -qexp_h_[2 * ts.output + 1]);
    ts.tb.gain[c.tree_node] = c.split.gain;
    ts.tb.cover[c.tree_node] = (float)H;
    // apply_split_to_builder
    const uint32_t e = c.split.feature;
    const uint32_t of = ds_->eff_to_orig[e];
    const uint32_t a = ts.tb.alloc(), b = ts.tb.alloc();
    ts.tb.split_index[c.tree_node] = of;
    ts.tb.left[c.tree_node] = a;
    ts.tb.right[c.tree_node] = b;
    uint32_t lt, rt;
    if (!(c.split.flags & 2u)) {
      ts.tb.thr[c.tree_node] = ds_->thr_up[(size_t)e * 256 + c.split.bin];
      ts.tb.dleft[c.tree_node] = (uint8_t)(c.split.flags & 1u);
      ts.tb.stype[c.tree_node] = 0;
      lt = a;
      rt = b;
    } else {
      int maxc = -1;
      for (int k = 0; k < 256; k++)
        if ((c.split.cats[k >> 5] >> (k & 31)) & 1u) maxc = k;
      std::vector<uint32_t> words;
      if (maxc >= 0) {
        words.assign((size_t)(maxc >> 5) + 1, 0u);
        for (int k = 0; k <= maxc; k++)
          if ((c.split.cats[k >> 5] >> (k & 31)) & 1u) words[k >> 5] |= 1u << (k & 31);
      }
      ts.tb.thr[c.tree_node] = 0.0f;
      ts.tb.dleft[c.tree_node] = (c.split.flags & 1u) ? 0 : 1;
      ts.tb.stype[c.tree_node] = 1;
      ts.tb.cat_nodes.push_back({c.tree_node, words});
      lt = b;
      rt = a;
    }
    const uint32_t right = ts.n_part++;
    const uint32_t begin = (uint32_t)(ts.t * n_) + ts.pbegin[c.node];
    const uint32_t nt = std::max<uint32_t>(1, (c.count + kPartTile - 1) / kPartTile);
    pitems.push_back({ts.t * MN_ + c.node, begin, c.count, (uint32_t)ptiles.size(), nt,
                      ts.t * MN_ + right, e, 0});
    for (uint32_t k = 0; k < nt; k++) {
      const uint32_t b0 = k * kPartTile;
      ptiles.push_back({(uint32_t)pitems.size() - 1, begin + b0,
                        c.count > b0 ? std::min(kPartTile, c.count - b0) : 0u, 0});
    }
    pend.push_back({tc.first, c, lt, rt, right});
  }
  if (pend.empty()) return;

  upload(s_, d_part_items_, pitems);
  upload(s_, d_part_tiles_, ptiles);
  tile_left_.ensure(ptiles.size());
  tile_off_.ensure(ptiles.size());
  item_left_.ensure(pitems.size());
  double prow = 0;
  for (auto& it : pitems) prow += it.count;
  int pi = ctx_->prof_begin(2, prow * (4 + 8 + 12 + 1));
  launch_partition(s_, dd_, d_part_items_.p, (int)pitems.size(), d_part_tiles_.p,
                   (int)ptiles.size(), node_split_.p, idx_.p, ogh_.p, idx_tmp_.p, ogh_tmp_.p,
                   flag_.p, tile_left_.p, tile_off_.p, item_left_.p, node_count_.p, node_sum_.p);
  ctx_->prof_end(pi);
  std::vector<uint32_t> left(pitems.size());
  d2h(s_, left.data(), item_left_.p, left.size());
  BST_HIP(hipStreamSynchronize(s_));
  ctx_->prof_drain();

  std::vector<HistItem> hitems;
  std::vector<uint32_t> zero, quads_full, quads_sums, snodes, sslots;
  std::vector<Segment> segs;
  bool leafwise = cfg.growth == BST_GROW_LEAFWISE;
  std::vector<uint8_t> need(pend.size());
  for (size_t i = 0; i < pend.size(); i++) {
    Pending& p = pend[i];
    TreeState& ts = trees[p.t];
    const uint32_t node = p.c.node, right = p.right;
    const uint32_t L = left[i], R = p.c.count - L;
    ts.pbegin[right] = ts.pbegin[node] + L;
    ts.pcount[right] = R;
    ts.pcount[node] = L;
    const bool left_small = L <= R;
    const uint32_t small = left_small ? node : right, large = left_small ? right : node;
    const uint32_t scount = left_small ? L : R;
    ts.pslot[large] = ts.pslot[node];
    if (large != node) ts.pslot[node] = UINT32_MAX;
    const uint32_t child_depth = p.c.depth + 1;
    need[i] = leafwise || child_depth < cfg.max_depth;
    const uint32_t gsmall = ts.t * MN_ + small, glarge = ts.t * MN_ + large;
    const uint32_t sbegin = (uint32_t)(ts.t * n_) + ts.pbegin[small];
    stats.hist_rows += scount;
    if (need[i]) {
      ts.pslot[small] = alloc_slot(ts);
      const uint32_t gss = ts.t * SPT_ + ts.pslot[small], gls = ts.t * SPT_ + ts.pslot[large];
      const uint32_t nt = std::max<uint32_t>(1, (scount + kHistTile - 1) / kHistTile);
      if (nt > 1 || scount == 0) zero.push_back(gss);
      for (uint32_t k = 0; k < nt && scount > 0; k++) {
        const uint32_t b0 = k * kHistTile, cnt = std::min(kHistTile, scount - b0);
        for (int gr = 0; gr < dd_.NG; gr++)
          hitems.push_back({gsmall, gss, sbegin + b0, cnt, (uint32_t)gr,
                            (uint32_t)ts.t | (nt == 1 ? kDirect : 0u)});
      }
      quads_full.insert(quads_full.end(), {gls, gss, glarge, gsmall});
      const uint32_t gl = ts.t * MN_ + node, grt = ts.t * MN_ + right;
      snodes.push_back(gl);
      sslots.push_back(ts.t * SPT_ + ts.pslot[node]);
      snodes.push_back(grt);
      sslots.push_back(ts.t * SPT_ + ts.pslot[right]);
    } else {
      ts.pslot[small] = UINT32_MAX;
      for (uint32_t b0 = 0; b0 < scount; b0 += kSegTile)
        segs.push_back({gsmall, sbegin + b0, std::min(kSegTile, scount - b0), (uint32_t)ts.t});
      quads_sums.insert(quads_sums.end(), {0u, 0u, glarge, gsmall});
    }
  }
  const int32_t* qexp = qexp_.p + 2 * trees[0].output;
  upload(s_, d_u32a_, zero);
  upload(s_, d_hist_items_, hitems);
  upload(s_, d_segs_, segs);
  upload(s_, d_u32b_, quads_full);
  upload(s_, d_u32c_, quads_sums);
  launch_zero_slots(s_, d_u32a_.p, (int)zero.size(), hist_.p, slot_elems_);
  if (!hitems.empty()) {
    double hrows = 0;
    int hnodes = 0;
    for (auto& it : hitems)
      if (it.group == 0) hrows += it.count;
    for (size_t k = 0; k < quads_full.size(); k += 4) hnodes++;
    pi = ctx_->prof_begin(0, hist_bytes(hrows, hnodes, true));
    launch_hist(s_, dd_, d_hist_items_.p, (int)hitems.size(), idx_.p, ogh_.p, 0, qexp, hist_.p,
                slot_elems_, node_sum_.p, true, 0);
    ctx_->prof_end(pi);
  }
  launch_node_sums(s_, d_segs_.p, (int)segs.size(), ogh_.p, 0, qexp, node_sum_.p);
  launch_subtract(s_, d_u32b_.p, (int)quads_full.size() / 4, hist_.p, slot_elems_, node_sum_.p);
  launch_subtract(s_, d_u32c_.p, (int)quads_sums.size() / 4, hist_.p, 0, node_sum_.p);
  if (!snodes.empty() && level_sampling()) {
    // The children's level (sample_level(new_depth), grower.rs:443-444): one
    // tree at a time on this path, and its expanded nodes share a depth.
    uint32_t depth = 0;
    for (auto& tc : chosen) depth = std::max(depth, tc.second.depth + 1);
    set_level_features(depth);
  }
  if (!snodes.empty()) {
    std::vector<uint32_t> both(snodes);
    both.insert(both.end(), sslots.begin(), sslots.end());
    upload(s_, d_u32d_, both);
    pi = ctx_->prof_begin(1, split_bytes((int)snodes.size()));
    launch_split(s_, dd_, d_u32d_.p, d_u32d_.p + snodes.size(), (int)snodes.size(), feats_.p,
                 (int)feats_h_.size(), hist_.p, slot_elems_, node_sum_.p, node_count_.p, qexp, MN_,
                 sp_, node_key_.p, cand_.p, node_split_.p);
    ctx_->prof_end(pi);
  }
  std::vector<long long> sums((size_t)2 * T_ * MN_);
  std::vector<DevSplit> splits((size_t)T_ * MN_);
  d2h(s_, sums.data(), node_sum_.p, sums.size());
  if (!snodes.empty()) d2h(s_, splits.data(), node_split_.p, splits.size());
  BST_HIP(hipStreamSynchronize(s_));
  ctx_->prof_drain();
  for (size_t i = 0; i < pend.size(); i++) {
    Pending& p = pend[i];
    TreeState& ts = trees[p.t];
    const uint32_t node = p.c.node, right = p.right;
    const uint32_t gl = ts.t * MN_ + node, grt = ts.t * MN_ + right;
    Cand cl{node, p.lt, p.c.depth + 1, need[i] ? splits[gl] : invalid_split(), sums[2 * gl],
            sums[2 * gl + 1], ts.pcount[node]};
    Cand cr{right, p.rt, p.c.depth + 1, need[i] ? splits[grt] : invalid_split(), sums[2 * grt],
            sums[2 * grt + 1], ts.pcount[right]};
    if (cfg.growth == BST_GROW_DEPTHWISE) {
      ts.next.push_back(cl);
      ts.next.push_back(cr);
    } else {
      ts.heap.push(cl);
      ts.heap.push(cr);
    }
  }
}

void Trainer::grow_round(std::vector<TreeState>& trees) {
  const auto& cfg = opt_.cfg;
  root(trees);
  const bool plain = sched_ == BST_SCHED_PLAIN;
  if (cfg.growth == BST_GROW_DEPTHWISE) {
    for (uint32_t depth = 0; depth < cfg.max_depth; depth++) {
      std::vector<std::pair<int, Cand>> chosen;
      for (size_t i = 0; i < trees.size(); i++)
        for (auto& c : trees[i].level) chosen.push_back({(int)i, c});
      if (chosen.empty()) break;
      for (auto& ts : trees) ts.next.clear();
      if (plain) {
        for (auto& one : chosen) {
          std::vector<std::pair<int, Cand>> v{one};
          expand_batch(trees, v);
        }
      } else {
        expand_batch(trees, chosen);
      }
      for (auto& ts : trees) ts.level.swap(ts.next);
    }
    for (auto& ts : trees) {
      for (auto& c : ts.level) make_leaf(ts, c);
      ts.level.clear();
    }
  } else {
    for (;;) {
      std::vector<std::pair<int, Cand>> chosen;
      for (size_t i = 0; i < trees.size(); i++) {
        TreeState& ts = trees[i];
        if (ts.n_leaves < cfg.max_leaves && !ts.heap.empty()) {
          chosen.push_back({(int)i, ts.heap.top()});
          ts.heap.pop();
          ts.n_leaves++;
        }
      }
      if (chosen.empty()) break;
      if (plain) {
        for (auto& one : chosen) {
          std::vector<std::pair<int, Cand>> v{one};
          expand_batch(trees, v);
        }
      } else {
        expand_batch(trees, chosen);
      }
    }
    for (auto& ts : trees)
      while (!ts.heap.empty()) {
        make_leaf(ts, ts.heap.top());
        ts.heap.pop();
      }
  }
  for (auto& ts : trees)
    for (size_t i = 0; i < ts.tb.leaf.size(); i++)
      if (ts.tb.is_leaf[i]) ts.tb.leaf[i] = ts.tb.leaf[i] * lr_;
}

// Leaf renewal of one round's trees (trainer.rs:265-291): every leaf with
// rows gets lr x the alpha-quantile of its rows' residuals, in the tree and
// in the value the prediction update adds.
void Trainer::renew_round(std::vector<TreeState>& trees, const float* y, const float* w) {
  std::vector<RenewLeaf> leaves;
  std::vector<std::pair<int, size_t>> where;  // (tree, leaf_map index) per entry
  for (size_t ti = 0; ti < trees.size(); ti++) {
    const TreeState& ts = trees[ti];
    const float alpha =
        opt_.cfg.objective == BST_OBJ_PINBALL ? opt_.alphas[(size_t)ts.output] : 0.5f;
    for (size_t li = 0; li < ts.leaf_map.size(); li++) {
      const uint32_t p = ts.leaf_map[li].second;
      if (ts.pcount[p] == 0) continue;
      leaves.push_back({(uint32_t)(ts.t * n_) + ts.pbegin[p], ts.pcount[p],
                        (uint32_t)ts.output, alpha});
      where.push_back({(int)ti, li});
    }
  }
  if (leaves.empty()) return;
  const int nl = (int)leaves.size();
  const int64_t n_pos = (int64_t)T_ * n_;
  upload(s_, d_renew_leaves_, leaves);
  renew_keys_.ensure((size_t)n_pos);
  renew_sorted_.ensure((size_t)n_pos);
  renew_seg_.ensure((size_t)2 * nl);
  renew_out_.ensure((size_t)nl);
  const size_t tb = renew_temp_bytes(n_pos, nl);
  renew_temp_.ensure(std::max<size_t>(tb, 1));
  renew_leaves(s_, d_renew_leaves_.p, nl, n_pos, idx_.p, y, preds_.p, w, n_, renew_keys_.p,
               renew_sorted_.p, renew_seg_.p, renew_seg_.p + nl, renew_temp_.p, tb,
               renew_out_.p);
  std::vector<float> vals((size_t)nl);
  d2h(s_, vals.data(), renew_out_.p, (size_t)nl);
  BST_HIP(hipStreamSynchronize(s_));
  for (int i = 0; i < nl; i++) {
    if (std::isnan(vals[(size_t)i])) continue;
    TreeState& ts = trees[(size_t)where[(size_t)i].first];
    const auto& lm = ts.leaf_map[where[(size_t)i].second];
    const float sv = vals[(size_t)i] * lr_;
    ts.tb.leaf[lm.first] = sv;
    ts.leaf_lr[lm.second] = sv;
  }
}

bool Trainer::level_sampling() const {
  return !(std::fabs(opt_.cfg.colsample_bylevel - 1.0f) < 1e-6f);
}

// ColSampler::sample_level (column.rs:199-224): the level's features are the
// ceil(|tree features| x colsample_bylevel) tree features with the smallest
// sampling keys of (round, output, depth); without level sampling, the tree's.
void Trainer::set_level_features(uint32_t depth) {
  if (level_sampling()) {
    const uint32_t kk = colsample_count((uint32_t)tree_feats_h_.size(), opt_.cfg.colsample_bylevel);
    const uint64_t base = level_key_base(opt_.cfg.seed, cur_round_, (uint32_t)cur_output_, depth);
    std::vector<uint64_t> kv(tree_feats_h_.size());
    for (size_t j = 0; j < tree_feats_h_.size(); j++)
      kv[j] = ((uint64_t)sample_key(base, tree_feats_h_[j]) << 32) | tree_feats_h_[j];
    std::sort(kv.begin(), kv.end());
    feats_h_.resize(kk);
    for (uint32_t j = 0; j < kk; j++) feats_h_[j] = (uint32_t)(kv[j] & 0xFFFFFFFFu);
    std::sort(feats_h_.begin(), feats_h_.end());
  } else {
    feats_h_ = tree_feats_h_;
  }
  h2d(s_, feats_.p, feats_h_.data(), feats_h_.size());
  // The split kernel's tie rule depends on the number of candidate features.
  sp_.rightmost = (opt_.ext.n_threads != 1 && feats_h_.size() >= 16) ? 1 : 0;
}

int Trainer::metric_kind() const {
  const auto& cfg = opt_.cfg;
  return cfg.metric == BST_METRIC_DEFAULT ? default_metric(cfg.objective) : cfg.metric;
}

void Trainer::metric_launch(const float* preds, const float* y, const float* w, int64_t n,
                            double* host_part) {
  const auto& cfg = opt_.cfg;
  // PredictionKind::Margin metrics read untransformed scores (eval.rs:150-184).
  const int mk = metric_kind();
  const int transform = mk == BST_METRIC_MARGIN_ACCURACY ? BST_TRANSFORM_IDENTITY
                        : cfg.objective == BST_OBJ_LOGISTIC ? BST_TRANSFORM_SIGMOID
                        : cfg.objective == BST_OBJ_SOFTMAX  ? BST_TRANSFORM_SOFTMAX
                                                            : BST_TRANSFORM_IDENTITY;
  const int nb = kMetricBlocks;
  d_metric_.ensure((size_t)2 * nb * opt_.K);
  MetricParams mp;
  mp.accuracy_threshold = cfg.accuracy_threshold;
  mp.huber_delta = cfg.metric_huber_delta;
  for (size_t i = 0; i < opt_.alphas.size() && i < 16; i++) mp.alphas[i] = opt_.alphas[i];
  mp.n_alphas = (int)std::min<size_t>(opt_.alphas.size(), 16);
  launch_metric(s_, mk, transform, mp, preds, opt_.K, y, w, n, d_metric_.p, nb);
  d2h(s_, host_part, d_metric_.p, (size_t)2 * nb * opt_.K);
}

double Trainer::metric_finish(const double* part) const {
  const int nb = kMetricBlocks;
  const int mk = metric_kind();
  if (mk == BST_METRIC_MAE) {  // mean over outputs of each output's weighted mean
    double acc = 0;
    for (int k = 0; k < opt_.K; k++) {
      double sa = 0, sw = 0;
      for (int b = 0; b < nb; b++) {
        sa += part[(size_t)(k * nb + b) * 2];
        sw += part[(size_t)(k * nb + b) * 2 + 1];
      }
      acc += sw > 0.0 ? sa / sw : 0.0;
    }
    return acc / opt_.K;
  }
  if (mk == BST_METRIC_QUANTILE) {  // pooled over the outputs' sums
    double sl = 0, sw = 0;
    for (int k = 0; k < opt_.K; k++)
      for (int b = 0; b < nb; b++) {
        sl += part[(size_t)(k * nb + b) * 2];
        sw += part[(size_t)(k * nb + b) * 2 + 1];
      }
    return sw > 0.0 ? sl / sw : 0.0;
  }
  if (mk == BST_METRIC_RMSE) {
    double acc = 0;
    for (int k = 0; k < opt_.K; k++) {
      double ss = 0, sw = 0;
      for (int b = 0; b < nb; b++) {
        ss += part[(size_t)(k * nb + b) * 2];
        sw += part[(size_t)(k * nb + b) * 2 + 1];
      }
      acc += sw > 0.0 ? std::sqrt(ss / sw) : 0.0;
    }
    return acc / opt_.K;
  }
  double sl = 0, sw = 0;
  for (int b = 0; b < nb; b++) {
    sl += part[(size_t)b * 2];
    sw += part[(size_t)b * 2 + 1];
  }
  if (mk == BST_METRIC_MAPE) return sw > 0 ? (sl / sw) * 100.0 : 0.0;
  return sw > 0 ? sl / sw : 0.0;
}

double Trainer::eval_metric(const float* preds, const float* y, const float* w, int64_t n) {
  if (metric_kind() == BST_METRIC_NONE) return std::numeric_limits<double>::quiet_NaN();
  std::vector<double> part((size_t)2 * kMetricBlocks * opt_.K);
  metric_launch(preds, y, w, n, part.data());
  BST_HIP(hipStreamSynchronize(s_));
  return metric_finish(part.data());
}

bst_model* Trainer::train(const float* y, const float* w, const float* vX, const float* vy,
                          int64_t n_valid) {
  const auto& cfg = opt_.cfg;
  const int K = opt_.K;
  const auto t0 = std::chrono::steady_clock::now();
  auto* m = new bst_model();
  m->info.n_groups = K;
  m->info.n_features = ds_->d.F;
  m->info.objective = cfg.objective;
  m->info.output_transform = cfg.objective == BST_OBJ_LOGISTIC ? BST_TRANSFORM_SIGMOID
                             : cfg.objective == BST_OBJ_SOFTMAX ? BST_TRANSFORM_SOFTMAX
                                                                : BST_TRANSFORM_IDENTITY;
  // Base score (objectives: regression.rs:44-68, classification.rs:54-76,189-222),
  // accumulated in f64 on the host from the device targets.
  {
    // Targets come back through the context's pinned staging buffer (a
    // pageable copy of 1M floats costs milliseconds).
    float* yh = (float*)ctx_->pinned_staging((size_t)(w ? 2 : 1) * n_ * sizeof(float));
    float* wh = w ? yh + n_ : nullptr;
    // A kernel writes the mapped pinned buffer directly: after the GPU has
    // been idle, DMA copies of this size have been seen to take milliseconds.
    void* ydev = nullptr;
    BST_HIP(hipHostGetDevicePointer(&ydev, yh, 0));
    launch_copy_u32(s_, (uint32_t*)ydev, (const uint32_t*)y, n_);
    if (w) launch_copy_u32(s_, (uint32_t*)ydev + n_, (const uint32_t*)w, n_);
    BST_HIP(hipStreamSynchronize(s_));
    if (std::getenv("BST_TRACE_SETUP"))
      std::fprintf(stderr, "[bst] targets to host %.3f ms\n",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                       .count());
    auto wt = [&](int64_t i) { return w ? wh[i] : 1.0f; };
    m->base_score.assign(K, 0.0f);
    if (cfg.objective == BST_OBJ_LAMBDARANK || cfg.objective == BST_OBJ_HINGE) {
      // Scores start at 0 (ranking has no prior; hinge, classification.rs:118-123).
    } else if (cfg.objective == BST_OBJ_ABSOLUTE || cfg.objective == BST_OBJ_PSEUDO_HUBER) {
      // Weighted median, regression.rs:106-117,218-230.
      m->base_score[0] = weighted_quantile(yh, wh, n_, 0.5f);
    } else if (cfg.objective == BST_OBJ_PINBALL) {
      // One weighted quantile per alpha, regression.rs:158-172.
      for (int k = 0; k < K; k++) m->base_score[k] = weighted_quantile(yh, wh, n_, opt_.alphas[k]);
    } else if (cfg.objective == BST_OBJ_POISSON) {
      // log of the weighted mean target, regression.rs:272-296.
      double sw = 0, swy = 0;
      for (int64_t i = 0; i < n_; i++) {
        sw += (double)wt(i);
        swy += (double)wt(i) * (double)yh[i];
      }
      const double mean = sw > 0.0 ? std::max(swy / sw, 1e-7) : 1.0;
      m->base_score[0] = (float)std::log(mean);
    } else if (cfg.objective == BST_OBJ_SQUARED) {
      double sw = 0, swy = 0;
      for (int64_t i = 0; i < n_; i++) {
        sw += (double)wt(i);
        swy += (double)wt(i) * (double)yh[i];
      }
      m->base_score[0] = sw > 0 ? (float)(swy / sw) : 0.0f;
    } else if (cfg.objective == BST_OBJ_LOGISTIC) {
      double pos = 0, tot = 0;
      for (int64_t i = 0; i < n_; i++) {
        pos += (double)wt(i) * (double)yh[i];
        tot += (double)wt(i);
      }
      const double p = std::min(std::max(pos / tot, 1e-7), 1.0 - 1e-7);
      m->base_score[0] = (float)std::log(p / (1.0 - p));
    } else {
      std::vector<double> cw(K, 0.0);
      double tot = 0;
      for (int64_t i = 0; i < n_; i++) {
        const size_t l = (size_t)yh[i];
        if (l < (size_t)K) cw[l] += (double)wt(i);
        tot += (double)wt(i);
      }
      for (int k = 0; k < K; k++) {
        const double p = std::min(std::max(cw[k] / tot, 1e-7), 1.0 - 1e-7);
        m->base_score[k] = (float)std::log(p);
      }
    }
  }
  {
    auto bits = [](float v) {
      uint32_t u;
      std::memcpy(&u, &v, 4);
      return u;
    };
    for (int k = 0; k < K; k++)
      launch_fill_u32(s_, (uint32_t*)preds_.p + (size_t)k * n_, bits(m->base_score[k]), n_);
    if (vX && n_valid > 0) {
      vpreds_.alloc((size_t)K * n_valid);
      for (int k = 0; k < K; k++)
        launch_fill_u32(s_, (uint32_t*)vpreds_.p + (size_t)k * n_valid, bits(m->base_score[k]),
                        n_valid);
    }
  }
  const uint32_t E = (uint32_t)ds_->d.E;
  feats_h_.resize(E);
  for (uint32_t i = 0; i < E; i++) feats_h_[i] = i;
  h2d(s_, feats_.p, feats_h_.data(), E);
  const bool bytree_sampling = !(std::fabs(cfg.colsample_bytree - 1.0f) < 1e-6f);
  const bool col_sampling = bytree_sampling || level_sampling();
  const bool uniform = !(std::fabs(cfg.subsample - 1.0f) < 1e-6f);


  // One grown tree as a single-tree forest (leaf values already carry the learning rate).
  auto single_tree = [&](const TreeBuild& tb, int output) {
    bst_model one;
    one.info.n_trees = 1;
    one.info.n_groups = K;
    one.info.n_features = ds_->d.F;
    one.base_score.assign(K, 0.0f);
    const size_t nn = tb.split_index.size();
    one.tree_offset = {0, (int64_t)nn};
    one.tree_group = {(uint32_t)output};
    one.split_index = tb.split_index;
    one.split_threshold = tb.thr;
    one.left = tb.left;
    one.right = tb.right;
    one.default_left = tb.dleft;
    one.is_leaf = tb.is_leaf;
    one.split_type = tb.stype;
    one.leaf_value.assign(nn, 0.0f);
    for (size_t i = 0; i < nn; i++)
      if (tb.is_leaf[i]) one.leaf_value[i] = tb.leaf[i];
    one.gain = tb.gain;
    one.cover = tb.cover;
    one.cat_start.assign(nn, 0);
    one.cat_size.assign(nn, 0);
    auto cn = tb.cat_nodes;
    std::sort(cn.begin(), cn.end(), [](auto& a, auto& b) { return a.first < b.first; });
    for (auto& c : cn) {
      one.cat_start[c.first] = (uint32_t)one.cat_words.size();
      one.cat_size[c.first] = (uint32_t)c.second.size();
      one.cat_words.insert(one.cat_words.end(), c.second.begin(), c.second.end());
    }
    one.info.total_nodes = (int64_t)nn;
    one.info.total_cat_words = (int64_t)one.cat_words.size();
    return one;
  };
  // Validation predictions += the tree (needs one.upload first).
  auto valid_update = [&](bst_model& one) {
    launch_predict_accum(s_, one.view(), vX, n_valid, n_valid, ds_->d.F, vpreds_.p, n_valid,
                         ctx_->lds_per_block);
  };
  auto append_tree = [&](bst_model* fm, const bst_model& one, int output) {
    const size_t nn = one.split_index.size();
    const int64_t off = fm->tree_offset.empty() ? 0 : fm->tree_offset.back();
    if (fm->tree_offset.empty()) fm->tree_offset.push_back(0);
    fm->tree_offset.push_back(off + (int64_t)nn);
    fm->tree_group.push_back((uint32_t)output);
    const uint32_t woff = (uint32_t)fm->cat_words.size();
    for (size_t i = 0; i < nn; i++) {
      fm->split_index.push_back(one.split_index[i]);
      fm->split_threshold.push_back(one.split_threshold[i]);
      fm->left.push_back(one.left[i]);
      fm->right.push_back(one.right[i]);
      fm->default_left.push_back(one.default_left[i]);
      fm->is_leaf.push_back(one.is_leaf[i]);
      fm->split_type.push_back(one.split_type[i]);
      fm->leaf_value.push_back(one.leaf_value[i]);
      fm->gain.push_back(one.gain[i]);
      fm->cover.push_back(one.cover[i]);
      fm->cat_start.push_back(one.cat_size[i] ? one.cat_start[i] + woff : 0);
      fm->cat_size.push_back(one.cat_size[i]);
    }
    fm->cat_words.insert(fm->cat_words.end(), one.cat_words.begin(), one.cat_words.end());
    stats.n_trees++;
  };
  const bool on_device = device_schedule_ok();
  if (opt_.ext.goss_top_rate > 0.0f && !on_device)
    raise(BST_E_CONFIG, "GOSS row sampling runs on the graph schedule");
  if (on_device) {
    // The split kernel's tie rule depends on the number of candidate features.
    const uint32_t nf = col_sampling
                            ? colsample_count(E, cfg.colsample_bytree)
                            : E;
    sp_.rightmost = (opt_.ext.n_threads != 1 && std::min(nf, E) >= 16) ? 1 : 0;
  }
  if (std::getenv("BST_TRACE_SETUP"))
    std::fprintf(stderr, "[bst] train set-up before the first round %.3f ms\n",
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                     .count());

  std::vector<std::vector<float>> leafv;
  bool have_best = false;
  double best_val = 0;
  size_t best_round = 0, cur_round = 0, best_n_trees = 0;
  std::vector<uint32_t> sampled;
  // Device rounds are pipelined: round r+1 is queued before the host reads
  // round r, so the host's tree assembly overlaps the device.  The round's
  // graph also adds each new tree to the validation predictions (a traversal
  // of the validation rows on the device), and the round's metric (on the
  // validation set when there is one, eval.rs:218-270) runs in the queue
  // right after it: one host wait per round, for the oldest round in flight.
  // Profiled runs read each round's kernel events right after it, unpipelined.
  valid_X_ = vX && vy && n_valid > 0 ? vX : nullptr;
  valid_y_ = valid_X_ ? vy : nullptr;
  valid_n_ = valid_X_ ? n_valid : 0;
  // With early stopping, a round queued past the stop is discarded; its
  // prediction updates would reach the kept training predictions, so a train
  // that keeps them (and may stop early) runs unpipelined.
  const bool pipelined =
      on_device && !ctx_->profile && !(ctx_->keep_preds && cfg.early_stopping_rounds > 0);
  const bool queue_metric = pipelined && metric_kind() != BST_METRIC_NONE;
  if (pipelined && cfg.n_trees > 0) device_launch(y, w, queue_metric);
  for (uint32_t round = 0; round < cfg.n_trees; round++) {
    double queued_value = std::numeric_limits<double>::quiet_NaN();
    if (on_device) {
      std::vector<TreeBuild> tbs;
      if (pipelined) {
        if (round + 1 < cfg.n_trees) device_launch(y, w, queue_metric);
        device_collect(tbs, &queued_value);
      } else {
        device_launch(y, w, false);
        device_collect(tbs, nullptr);
      }
      for (int t = 0; t < T_; t++) {
        bst_model one = single_tree(tbs[t], t);
        append_tree(m, one, t);
      }
    } else {
    BST_HIP(hipMemsetAsync(absmax_.p, 0, absmax_.bytes(), s_));
    launch_objective(y, w);
    launch_qscale(s_, absmax_.p, K, n_, qexp_.p);
    for (int k0 = 0; k0 < K; k0 += T_) {
      const int nb = std::min(T_, K - k0);  // the last batch of a round may be short
      std::vector<TreeState> trees(nb);
      BST_HIP(hipMemsetAsync(node_sum_.p, 0, node_sum_.bytes(), s_));
      std::vector<uint32_t> outs(nb);
      for (int t = 0; t < nb; t++) {
        TreeState& ts = trees[t];
        ts.t = t;
        ts.output = k0 + t;
        outs[t] = ts.output;
        ts.pbegin.assign(MN_, 0);
        ts.pcount.assign(MN_, 0);
        ts.pslot.assign(MN_, UINT32_MAX);
        ts.leaf_lr.assign(MN_, std::numeric_limits<float>::quiet_NaN());
        ts.leaf_map.clear();
        ts.free_slots.clear();
        for (uint32_t s = SPT_; s-- > 0;) ts.free_slots.push_back(s);
        // Column sampling per tree (column.rs:180-186): the k columns with
        // the smallest sampling keys of (round, output); per level
        // (column.rs:199-224) a subset of those for each depth.
        if (col_sampling) {
          if (bytree_sampling) {
            const uint32_t kk = colsample_count(E, cfg.colsample_bytree);
            tree_feats_h_ = smallest_keys(E, sample_key_base(cfg.seed, kStreamCols, round, (uint32_t)ts.output), kk);
          } else {
            tree_feats_h_.resize(E);
            for (uint32_t i = 0; i < E; i++) tree_feats_h_[i] = i;
          }
          cur_round_ = round;
          cur_output_ = ts.output;
          set_level_features(0);
        } else {
          sp_.rightmost = (opt_.ext.n_threads != 1 && E >= 16) ? 1 : 0;
        }
        // Uniform row sampling (row.rs:170-208): target_count rows, the
        // smallest sampling keys of (round, output).
        if (uniform) {
          const size_t tc = std::min<size_t>(std::max<size_t>((size_t)((float)n_ * cfg.subsample), 1), n_);
          sampled = smallest_keys((uint32_t)n_, sample_key_base(cfg.seed, kStreamRows, round, (uint32_t)ts.output), (uint32_t)tc);
          upload(s_, d_sampled_, sampled);
          ts.rows = (int64_t)tc;
          ts.sampled = true;
          launch_init_rows(s_, idx_.p + t * n_, ogh_.p + t * n_, g_.p + (size_t)ts.output * n_,
                           h_.p + (size_t)ts.output * n_, d_sampled_.p, ts.rows);
        } else {
          ts.rows = n_;
          launch_init_rows(s_, idx_.p + t * n_, ogh_.p + t * n_, g_.p + (size_t)ts.output * n_,
                           h_.p + (size_t)ts.output * n_, nullptr, n_);
        }
        ts.pcount[0] = (uint32_t)ts.rows;
      }
      sp_.rightmost = sp_.rightmost;
      grow_round(trees);
      if (needs_leaf_renew(cfg.objective)) renew_round(trees, y, w);
      // Prediction update and model append, trees in output order.
      h2d(s_, tree_output_.p, outs.data(), outs.size());
      for (auto& ts : trees) {
        bst_model one = single_tree(ts.tb, ts.output);
        const bool need_dev = ts.sampled || (vX && n_valid > 0);
        if (need_dev) one.upload(s_);
        if (!ts.sampled) {
          std::vector<Segment> segs;
          std::vector<float> lv((size_t)T_ * MN_, 0.0f);
          for (uint32_t p = 0; p < ts.n_part; p++) {
            const float v = ts.leaf_lr[p];
            if (std::isnan(v)) continue;
            lv[ts.t * MN_ + p] = v;
            for (uint32_t b0 = 0; b0 < ts.pcount[p]; b0 += kSegTile)
              segs.push_back({ts.t * MN_ + p, (uint32_t)(ts.t * n_) + ts.pbegin[p] + b0,
                              std::min(kSegTile, ts.pcount[p] - b0), (uint32_t)ts.t});
          }
          h2d(s_, leaf_val_.p, lv.data(), lv.size());
          upload(s_, d_segs_, segs);
          launch_leaf_update(s_, d_segs_.p, (int)segs.size(), leaf_val_.p, idx_.p, 0, preds_.p,
                             tree_output_.p, n_);
          BST_HIP(hipStreamSynchronize(s_));
        } else {
          launch_predict_accum(s_, one.view(), ds_->d.raw, n_, n_, ds_->d.F, preds_.p, n_,
                               ctx_->lds_per_block);
        }
        if (vX && n_valid > 0) valid_update(one);
        if (need_dev) BST_HIP(hipStreamSynchronize(s_));
        append_tree(m, one, ts.output);
      }
    }
    }
    // evaluate_round (eval.rs:218-260) and early stopping (callback.rs:78-118).
    double value;
    if (pipelined)
      value = queued_value;
    else if (vX && vy && n_valid > 0)
      value = eval_metric(vpreds_.p, vy, nullptr, n_valid);
    else
      value = eval_metric(preds_.p, y, w, n_);
    metric_history.push_back(value);
    if (cfg.early_stopping_rounds > 0) {
      const bool hib = higher_is_better(metric_kind());
      const bool imp = !have_best || (hib ? value > best_val : value < best_val);
      if (imp) {
        best_val = value;
        have_best = true;
        best_round = cur_round;
        best_n_trees = m->tree_group.size();
      }
      cur_round++;
      if (cur_round - best_round > cfg.early_stopping_rounds) break;
    }
  }
  if (pipelined) BST_HIP(hipStreamSynchronize(s_));  // a round queued past an early stop
  if (ctx_->keep_preds) {
    // Training predictions after the last round trained (before any
    // early-stopping truncation): base score plus every tree's leaf, added
    // by the partition's row ranges (or by traversal for sampled rounds).
    ctx_->last_preds.resize((size_t)K * n_);
    d2h(s_, ctx_->last_preds.data(), preds_.p, ctx_->last_preds.size());
    BST_HIP(hipStreamSynchronize(s_));
  }
  if (cfg.early_stopping_rounds > 0 && best_n_trees > 0 && best_n_trees < m->tree_group.size()) {
    const int64_t keep = m->tree_offset[best_n_trees];
    m->tree_offset.resize(best_n_trees + 1);
    m->tree_group.resize(best_n_trees);
    for (auto* v : {&m->split_index, &m->left, &m->right, &m->cat_start, &m->cat_size})
      v->resize(keep);
    for (auto* v : {&m->split_threshold, &m->leaf_value, &m->gain, &m->cover}) v->resize(keep);
    for (auto* v : {&m->default_left, &m->is_leaf, &m->split_type}) v->resize(keep);
    uint32_t words = 0;
    for (int64_t i = 0; i < keep; i++)
      if (m->cat_size[i]) words = std::max(words, m->cat_start[i] + m->cat_size[i]);
    m->cat_words.resize(words);
  }
  if (m->tree_offset.empty()) m->tree_offset.push_back(0);
  m->info.n_trees = (int64_t)m->tree_group.size();
  m->info.total_nodes = m->tree_offset.back();
  m->info.total_cat_words = (int64_t)m->cat_words.size();
  stats.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return m;
}

}  // namespace bst
