// trainer_device.cpp — the graph schedule: one boosting round as a fixed
// launch sequence driven by the device-side grower, captured in a hipGraph.
//
// Round = fill(absmax) -> gradients -> qscale -> init rows (per tree) ->
// controller init -> root histogram + split -> steps x (plan, partition,
// children, zero, histogram, sums, subtract, split[, advance]) -> finalize
// -> partition-range prediction update -> copy of the tree arrays to pinned
// host memory.  No host synchronisation happens inside a round; the only one
// is after the graph launch, to read the grown trees (and the metric).
#include <algorithm>
#include <cmath>
#include <limits>

#include "grow_device.h"
#include "trainer.h"

namespace bst {

namespace {
constexpr uint32_t kHistTileD = 8192;
constexpr uint32_t kPartTileD = 4096;
constexpr uint32_t kSegTileD = 16384;

template <class T>
T* pinned(size_t n) {
  T* p = nullptr;
  BST_HIP(hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
  return p;
}
}  // namespace

struct Trainer::DevState {
  CtlParams P{};
  DeviceBufs B{};
  int steps = 0;
  uint32_t cap_items = 0, cap_tiles = 0, cap_hist = 0, cap_zero = 0, cap_quads = 0,
           cap_segs = 0, cap_split = 0, cap_lsegs = 0;
  DevBuf<TreeCtl> ctl;
  DevBuf<uint32_t> open_node, open_tnode, open_depth, next_node, next_tnode, next_depth, ch_node,
      ch_tnode, ch_depth, pbegin, pslot, free_slot;
  DevBuf<uint32_t> tr_split, tr_left, tr_right, tr_cats;
  DevBuf<float> tr_thr, tr_leaf, tr_gain, tr_cover;
  DevBuf<uint8_t> tr_dleft, tr_isleaf, tr_stype;
  DevBuf<PartItem> items;
  DevBuf<PartTile> tiles;
  DevBuf<uint32_t> pend;
  DevBuf<HistItem> hitems;
  DevBuf<uint32_t> zero, quads, squads, split_nodes, split_slots, counts, rows, eff_to_orig;
  DevBuf<Segment> segs, leaf_segs;
  DevBuf<float> thr_up;
  // Pinned mirrors of the round's results.
  TreeCtl* h_ctl = nullptr;
  uint32_t *h_split = nullptr, *h_left = nullptr, *h_right = nullptr, *h_cats = nullptr,
           *h_counts = nullptr;
  float *h_thr = nullptr, *h_leaf = nullptr, *h_gain = nullptr, *h_cover = nullptr;
  uint8_t *h_dleft = nullptr, *h_isleaf = nullptr, *h_stype = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~DevState() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : {(void*)h_ctl, (void*)h_split, (void*)h_left, (void*)h_right, (void*)h_cats,
                    (void*)h_counts, (void*)h_thr, (void*)h_leaf, (void*)h_gain, (void*)h_cover,
                    (void*)h_dleft, (void*)h_isleaf, (void*)h_stype})
      if (p) (void)hipHostFree(p);
  }
};

void Trainer::DevStateDelete::operator()(DevState* p) const { delete p; }

bool Trainer::device_schedule_ok() const {
  const auto& c = opt_.cfg;
  const bool sampling = !(std::fabs(c.subsample - 1.0f) < 1e-6f) ||
                        !(std::fabs(c.colsample_bytree - 1.0f) < 1e-6f) ||
                        opt_.ext.goss_top_rate > 0.0f;
  const bool depth_ok = c.growth != BST_GROW_DEPTHWISE || c.max_depth <= 12;
  return sched_ == BST_SCHED_GRAPH && !sampling && depth_ok && T_ == opt_.K;
}

void Trainer::device_setup() {
  std::unique_ptr<DevState, DevStateDelete> st(new DevState());
  const auto& c = opt_.cfg;
  CtlParams& P = st->P;
  P.T = (uint32_t)T_;
  P.MN = MN_;
  P.SPT = SPT_;
  const uint32_t max_leaves = c.growth == BST_GROW_DEPTHWISE ? (1u << c.max_depth) : c.max_leaves;
  P.MNT = 2 * max_leaves + 1;
  P.growth = c.growth == BST_GROW_DEPTHWISE ? 0 : 1;
  P.W = P.growth == 0 ? std::max<uint32_t>(1, max_leaves / 2) : 1;
  P.WO = P.growth == 0 ? max_leaves : max_leaves + 1;
  P.max_depth = c.max_depth;
  P.max_leaves = c.max_leaves;
  P.NG = (uint32_t)ds_->d.NG;
  P.E = (uint32_t)ds_->d.E;
  P.FG = (uint32_t)ds_->d.FG;
  P.hist_tile = kHistTileD;
  P.part_tile = kPartTileD;
  P.seg_tile = kSegTileD;
  P.stride = n_;
  P.lr = lr_;
  P.lambda = (double)c.lambda;
  P.alpha = (double)c.alpha;
  P.min_gain = (double)c.min_gain;
  P.out0 = 0;
  st->steps = P.growth == 0 ? (int)c.max_depth : (int)(c.max_leaves > 0 ? c.max_leaves - 1 : 0);
  const uint32_t T = P.T;
  const uint32_t tiles_n = (uint32_t)((n_ + kPartTileD - 1) / kPartTileD);
  const uint32_t hist_n = (uint32_t)((n_ + kHistTileD - 1) / kHistTileD);
  const uint32_t segs_n = (uint32_t)((n_ + kSegTileD - 1) / kSegTileD);
  st->cap_items = T * P.W;
  st->cap_tiles = T * (tiles_n + P.W);
  st->cap_hist = T * (hist_n + P.W) * P.NG;
  st->cap_zero = T * (P.W + 1);
  st->cap_quads = T * P.W;
  st->cap_segs = T * (segs_n + P.W);
  st->cap_split = std::max<uint32_t>(T, 2 * T * P.W);
  st->cap_lsegs = T * (segs_n + MN_);
  auto& s = *st;
  s.ctl.alloc(T);
  for (auto* b : {&s.open_node, &s.open_tnode, &s.open_depth}) b->alloc((size_t)T * P.WO);
  for (auto* b : {&s.next_node, &s.next_tnode, &s.next_depth}) b->alloc((size_t)T * 2 * P.W);
  for (auto* b : {&s.ch_node, &s.ch_tnode, &s.ch_depth}) b->alloc((size_t)T * P.W);
  s.pbegin.alloc((size_t)T * MN_);
  s.pslot.alloc((size_t)T * MN_);
  s.free_slot.alloc((size_t)T * SPT_);
  for (auto* b : {&s.tr_split, &s.tr_left, &s.tr_right}) b->alloc((size_t)T * P.MNT);
  s.tr_cats.alloc((size_t)T * P.MNT * kMaxCatWords);
  for (auto* b : {&s.tr_thr, &s.tr_leaf, &s.tr_gain, &s.tr_cover}) b->alloc((size_t)T * P.MNT);
  for (auto* b : {&s.tr_dleft, &s.tr_isleaf, &s.tr_stype}) b->alloc((size_t)T * P.MNT);
  s.items.alloc(s.cap_items);
  s.tiles.alloc(s.cap_tiles);
  s.pend.alloc((size_t)5 * s.cap_items);
  s.hitems.alloc(s.cap_hist);
  s.zero.alloc(s.cap_zero);
  s.quads.alloc((size_t)4 * s.cap_quads);
  s.squads.alloc((size_t)4 * s.cap_quads);
  s.split_nodes.alloc(s.cap_split);
  s.split_slots.alloc(s.cap_split);
  s.counts.alloc(C_NCOUNTS);
  s.rows.alloc(T);
  s.segs.alloc(s.cap_segs);
  s.leaf_segs.alloc(s.cap_lsegs);
  s.thr_up.alloc(ds_->thr_up.size());
  h2d(s_, s.thr_up.p, ds_->thr_up.data(), ds_->thr_up.size());
  s.eff_to_orig.alloc(ds_->eff_to_orig.size());
  h2d(s_, s.eff_to_orig.p, ds_->eff_to_orig.data(), ds_->eff_to_orig.size());
  tile_left_.ensure(s.cap_tiles);
  tile_off_.ensure(s.cap_tiles);
  item_left_.ensure(s.cap_items);
  std::vector<uint32_t> rows(T, (uint32_t)n_);
  h2d(s_, s.rows.p, rows.data(), rows.size());
  std::vector<uint32_t> outs(T);
  for (uint32_t t = 0; t < T; t++) outs[t] = t;
  h2d(s_, tree_output_.p, outs.data(), outs.size());
  DeviceBufs& B = s.B;
  B.ctl = s.ctl.p;
  B.open_node = s.open_node.p;
  B.open_tnode = s.open_tnode.p;
  B.open_depth = s.open_depth.p;
  B.next_node = s.next_node.p;
  B.next_tnode = s.next_tnode.p;
  B.next_depth = s.next_depth.p;
  B.ch_node = s.ch_node.p;
  B.ch_tnode = s.ch_tnode.p;
  B.ch_depth = s.ch_depth.p;
  B.pbegin = s.pbegin.p;
  B.pslot = s.pslot.p;
  B.free_slot = s.free_slot.p;
  B.leaf_val = leaf_val_.p;
  B.tr_split = s.tr_split.p;
  B.tr_left = s.tr_left.p;
  B.tr_right = s.tr_right.p;
  B.tr_cats = s.tr_cats.p;
  B.tr_thr = s.tr_thr.p;
  B.tr_leaf = s.tr_leaf.p;
  B.tr_gain = s.tr_gain.p;
  B.tr_cover = s.tr_cover.p;
  B.tr_dleft = s.tr_dleft.p;
  B.tr_isleaf = s.tr_isleaf.p;
  B.tr_stype = s.tr_stype.p;
  B.items = s.items.p;
  B.tiles = s.tiles.p;
  B.pend = s.pend.p;
  B.hitems = s.hitems.p;
  B.zero = s.zero.p;
  B.quads = s.quads.p;
  B.squads = s.squads.p;
  B.split_nodes = s.split_nodes.p;
  B.split_slots = s.split_slots.p;
  B.segs = s.segs.p;
  B.leaf_segs = s.leaf_segs.p;
  B.counts = s.counts.p;
  B.node_sum = node_sum_.p;
  B.node_count = node_count_.p;
  B.node_split = node_split_.p;
  B.thr_up = s.thr_up.p;
  B.eff_to_orig = s.eff_to_orig.p;
  B.qexp = qexp_.p;
  B.item_left = item_left_.p;
  const size_t tn = (size_t)T * P.MNT;
  s.h_ctl = pinned<TreeCtl>(T);
  s.h_split = pinned<uint32_t>(tn);
  s.h_left = pinned<uint32_t>(tn);
  s.h_right = pinned<uint32_t>(tn);
  s.h_cats = pinned<uint32_t>(tn * kMaxCatWords);
  s.h_counts = pinned<uint32_t>(C_NCOUNTS);
  s.h_thr = pinned<float>(tn);
  s.h_leaf = pinned<float>(tn);
  s.h_gain = pinned<float>(tn);
  s.h_cover = pinned<float>(tn);
  s.h_dleft = pinned<uint8_t>(tn);
  s.h_isleaf = pinned<uint8_t>(tn);
  s.h_stype = pinned<uint8_t>(tn);
  BST_HIP(hipStreamSynchronize(s_));
  dev_ = std::move(st);
}

// Every launch of one round (T trees = all K outputs), no host sync.
void Trainer::device_enqueue_round(const float* y, const float* w) {
  DevState& S = *dev_;
  const CtlParams& P = S.P;
  const DeviceBufs& B = S.B;
  const auto& cfg = opt_.cfg;
  const int K = opt_.K;
  const auto& d = ds_->d;
  launch_fill_u32(s_, absmax_.p, 0u, 2 * K);
  launch_gradients(s_, cfg.objective, preds_.p, y, w, n_, K, g_.p, h_.p, absmax_.p);
  launch_qscale(s_, absmax_.p, K, n_, qexp_.p);
  for (uint32_t t = 0; t < P.T; t++)
    launch_init_rows(s_, idx_.p + t * n_, ogh_.p + t * n_, g_.p + (size_t)t * n_,
                     h_.p + (size_t)t * n_, nullptr, n_);
  const int g_hist = (int)std::min<uint32_t>(S.cap_hist, 2048);
  const int g_tiles = (int)std::min<uint32_t>(S.cap_tiles, 2048);
  const int g_items = (int)std::min<uint32_t>(S.cap_items, 1024);
  const int g_quads = (int)std::min<uint32_t>(std::max<uint32_t>(S.cap_quads, 1), 256);
  const int g_zero = (int)std::min<uint32_t>(S.cap_zero, 256);
  const int g_segs = (int)std::min<uint32_t>(std::max<uint32_t>(S.cap_segs, 1), 1024);
  const int g_split = (int)std::min<uint32_t>(S.cap_split, 512);
  const int g_lsegs = (int)std::min<uint32_t>(S.cap_lsegs, 2048);
  uint32_t* cnt = S.counts.p;
  launch_ctl_init(s_, P, B, S.rows.p);
  launch_zero_slots(s_, B.zero, g_zero, hist_.p, slot_elems_, cnt + C_ZERO);
  int pi = ctx_->prof_begin(0, 0.0);
  launch_hist(s_, d, B.hitems, g_hist, idx_.p, ogh_.p, 0, qexp_.p, hist_.p, slot_elems_,
              node_sum_.p, P.T > 1, 0, cnt + C_HIST);
  ctx_->prof_end(pi);
  pi = ctx_->prof_begin(1, 0.0);
  launch_split(s_, d, B.split_nodes, B.split_slots, g_split, feats_.p, (int)feats_h_.size(),
               hist_.p, slot_elems_, node_sum_.p, node_count_.p, qexp_.p, MN_, sp_, node_key_.p,
               cand_.p, node_split_.p, cnt + C_SPLIT);
  ctx_->prof_end(pi);
  for (int step = 0; step < S.steps; step++) {
    launch_ctl_plan(s_, P, B);
    pi = ctx_->prof_begin(2, 0.0);
    launch_partition(s_, d, B.items, g_items, B.tiles, g_tiles, node_split_.p, idx_.p, ogh_.p,
                     idx_tmp_.p, ogh_tmp_.p, flag_.p, tile_left_.p, tile_off_.p, item_left_.p,
                     node_count_.p, node_sum_.p, cnt + C_ITEMS, cnt + C_TILES);
    ctx_->prof_end(pi);
    launch_ctl_children(s_, P, B);
    launch_zero_slots(s_, B.zero, g_zero, hist_.p, slot_elems_, cnt + C_ZERO);
    pi = ctx_->prof_begin(0, 0.0);
    launch_hist(s_, d, B.hitems, g_hist, idx_.p, ogh_.p, 0, qexp_.p, hist_.p, slot_elems_,
                node_sum_.p, true, 0, cnt + C_HIST);
    ctx_->prof_end(pi);
    launch_node_sums(s_, B.segs, g_segs, ogh_.p, 0, qexp_.p, node_sum_.p, cnt + C_SEGS);
    launch_subtract(s_, B.quads, g_quads, hist_.p, slot_elems_, node_sum_.p, cnt + C_QUADS);
    launch_subtract(s_, B.squads, g_quads, hist_.p, 0, node_sum_.p, cnt + C_SQUADS);
    pi = ctx_->prof_begin(1, 0.0);
    launch_split(s_, d, B.split_nodes, B.split_slots, g_split, feats_.p, (int)feats_h_.size(),
                 hist_.p, slot_elems_, node_sum_.p, node_count_.p, qexp_.p, MN_, sp_,
                 node_key_.p, cand_.p, node_split_.p, cnt + C_SPLIT);
    ctx_->prof_end(pi);
    if (P.growth == 0) launch_ctl_advance(s_, P, B);
  }
  launch_ctl_finalize(s_, P, B);
  launch_leaf_update(s_, B.leaf_segs, g_lsegs, leaf_val_.p, idx_.p, 0, preds_.p,
                     tree_output_.p, n_, cnt + C_LSEGS);
  const size_t tn = (size_t)P.T * P.MNT;
  d2h(s_, S.h_ctl, S.ctl.p, P.T);
  d2h(s_, S.h_counts, S.counts.p, C_NCOUNTS);
  d2h(s_, S.h_split, S.tr_split.p, tn);
  d2h(s_, S.h_left, S.tr_left.p, tn);
  d2h(s_, S.h_right, S.tr_right.p, tn);
  d2h(s_, S.h_cats, S.tr_cats.p, tn * kMaxCatWords);
  d2h(s_, S.h_thr, S.tr_thr.p, tn);
  d2h(s_, S.h_leaf, S.tr_leaf.p, tn);
  d2h(s_, S.h_gain, S.tr_gain.p, tn);
  d2h(s_, S.h_cover, S.tr_cover.p, tn);
  d2h(s_, S.h_dleft, S.tr_dleft.p, tn);
  d2h(s_, S.h_isleaf, S.tr_isleaf.p, tn);
  d2h(s_, S.h_stype, S.tr_stype.p, tn);
}

// Launch (or capture and launch) one round, wait, and return the trees.
void Trainer::device_round(const float* y, const float* w, std::vector<TreeBuild>& out) {
  if (!dev_) device_setup();
  DevState& S = *dev_;
  const bool use_graph = !ctx_->profile;  // profiled runs time each launch with events
  if (use_graph) {
    if (!S.exec) {
      BST_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
      device_enqueue_round(y, w);
      BST_HIP(hipStreamEndCapture(s_, &S.graph));
      BST_HIP(hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0));
    }
    BST_HIP(hipGraphLaunch(S.exec, s_));
  } else {
    device_enqueue_round(y, w);
  }
  BST_HIP(hipStreamSynchronize(s_));
  BST_HIP(hipGetLastError());
  if (ctx_->profile) {
    // Algorithmic bytes of the round's histogram launches, spread evenly.
    const double rows = (double)S.h_counts[C_HROWS];
    const int nodes = (int)S.h_counts[C_HNODES];
    const double bytes = hist_bytes(rows, nodes, true);
    int nh = 0;
    for (auto& p : ctx_->pending)
      if (p.kind == 0) nh++;
    for (auto& p : ctx_->pending)
      if (p.kind == 0) p.bytes = bytes / std::max(nh, 1);
    ctx_->prof_drain();
  }
  stats.hist_rows += S.h_counts[C_HROWS];
  const CtlParams& P = S.P;
  out.assign(P.T, TreeBuild());
  for (uint32_t t = 0; t < P.T; t++) {
    TreeBuild& tb = out[t];
    const uint32_t nn = S.h_ctl[t].n_tnodes;
    const size_t o = (size_t)t * P.MNT;
    for (uint32_t i = 0; i < nn; i++) {
      tb.alloc();
      tb.split_index[i] = S.h_split[o + i];
      tb.left[i] = S.h_left[o + i];
      tb.right[i] = S.h_right[o + i];
      tb.thr[i] = S.h_thr[o + i];
      tb.leaf[i] = S.h_isleaf[o + i] ? S.h_leaf[o + i] : 0.0f;
      tb.gain[i] = S.h_gain[o + i];
      tb.cover[i] = S.h_cover[o + i];
      tb.dleft[i] = S.h_dleft[o + i];
      tb.is_leaf[i] = S.h_isleaf[o + i];
      tb.stype[i] = S.h_stype[o + i];
      if (!tb.is_leaf[i] && tb.stype[i] == 1) {
        const uint32_t* c = S.h_cats + (o + i) * kMaxCatWords;
        int maxc = -1;
        for (int k = 0; k < 256; k++)
          if ((c[k >> 5] >> (k & 31)) & 1u) maxc = k;
        std::vector<uint32_t> words;
        if (maxc >= 0) words.assign(c, c + (maxc >> 5) + 1);
        tb.cat_nodes.push_back({i, words});
      }
    }
  }
}

}  // namespace bst
