// trainer_device.cpp — the graph schedule: one boosting round as a fixed
// launch sequence driven by the device-side grower, captured in a hipGraph.
//
// Round = gradients -> qscale -> controller init -> rows + root totals ->
// root histogram -> finish -> split -> steps x (plan, one-pass partition into
// the other buffer of the idx/ogh pair, children, histogram, finish, split)
// -> finalize -> partition-range prediction update -> one copy of the
// round's trees to pinned host memory.  No host synchronisation happens
// inside a round; the only one is after the graph launch, to read the trees.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "grow_device.h"
#include "limits.h"
#include "sample_device.h"
#include "trainer.h"

namespace bst {

namespace {
constexpr uint32_t kHistTileD = 65536;  // largest histogram tile (rows per item)
constexpr uint32_t kPartTileD = kFusedTileRows;
constexpr uint32_t kSegTileD = 4096;
constexpr int kMaxGrid = 2048;

int grid_for(uint32_t cap, int max_grid) {
  return (int)std::max<uint32_t>(1, std::min<uint32_t>(cap, (uint32_t)max_grid));
}
}  // namespace

struct Trainer::DevState {
  CtlParams P{};
  DeviceBufs B{};
  TreeOut O{};
  int steps = 0;
  uint32_t cap_items = 0, cap_tiles = 0, cap_hist = 0, cap_fin = 0, cap_split = 0,
           cap_lsegs = 0;
  size_t out_bytes = 0;
  DevBuf<uint32_t> open_node, open_tnode, open_depth, next_node, next_tnode, next_depth, ch_node,
      ch_tnode, ch_depth, pbegin, pslot, free_slot, cursor;
  DevBuf<uint8_t> pbuf;
  DevBuf<char> out;  // TreeOut arrays, back to back
  DevBuf<PartItem> items;
  DevBuf<PartTile> tiles;
  DevBuf<unsigned long long> tstat;  // look-back status per partition tile
  DevBuf<uint32_t> pend;
  DevBuf<HistItem> hitems;
  DevBuf<uint32_t> fin, split_nodes, split_slots, split_minus, rows, eff_to_orig;
  DevBuf<Segment> leaf_segs;
  DevBuf<float> thr_up;
  DevBuf<long long> partial;    // one histogram group per multi-tile item
  // Sampling: row draws (uniform / GOSS) and per-tree column lists.
  SampleParams SP{};
  int sample_mode = kSampleNone;
  uint32_t goss_warmup = 0;
  int n_feats = 0;     // candidate columns per tree
  bool col_sample = false;
  uint32_t host_round = 0;
  DevBuf<uint32_t> round;  // boosting round, read by the sampling kernels
  DevBuf<SampleState> samp;
  DevBuf<uint32_t> bcount, feats_dev;
  // Two pinned mirrors of `out` (and of the metric partials): round r reads
  // slot r % 2 while round r + 1 may be in flight.
  char* h_out[2] = {nullptr, nullptr};
  TreeOut H[2]{};
  double* h_metric[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  uint64_t launched = 0, collected = 0;  // rounds queued / read
  // [0]: rounds without row sampling, [1]: rounds with it.
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  BufPool* pool = nullptr;  // where the pinned mirrors came from (null: hipHostMalloc)
  size_t metric_bytes = 0;
  void* pinned(size_t bytes) {
    if (pool)
      if (void* q = pool->take_host(bytes)) return q;
    void* q = nullptr;
    BST_HIP(hipHostMalloc(&q, bytes, hipHostMallocDefault));
    return q;
  }
  void unpin(void* q, size_t bytes) {
    if (!q) return;
    if (pool)
      pool->put_host(q, bytes);
    else
      (void)hipHostFree(q);
  }
  ~DevState() {
    for (int i = 0; i < 2; i++) {
      if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    for (int i = 0; i < 2; i++) {
      unpin(h_out[i], out_bytes);
      unpin(h_metric[i], metric_bytes);
      if (done[i]) (void)hipEventDestroy(done[i]);
    }
  }
};

void Trainer::DevStateDelete::operator()(DevState* p) const { delete p; }

bool Trainer::graph_capable() const {
  const auto& c = opt_.cfg;
  const bool depth_ok = c.growth != BST_GROW_DEPTHWISE || c.max_depth <= 12;
  // Leaf renewal (quantile objectives) reads each leaf's rows in the stable
  // partition order of the host-driven schedule, which the graph's fused
  // partition does not keep.
  // Column sampling by level draws a feature subset per depth on the host.
  const bool bylevel = !(std::fabs(c.colsample_bylevel - 1.0f) < 1e-6f);
  return sched_ == BST_SCHED_GRAPH && depth_ok && !needs_leaf_renew(c.objective) && !bylevel;
}

bool Trainer::device_schedule_ok() const { return graph_capable() && T_ == opt_.K; }

void Trainer::device_setup() {
  std::unique_ptr<DevState, DevStateDelete> st(new DevState());
  const auto& c = opt_.cfg;
  CtlParams& P = st->P;
  P.T = (uint32_t)T_;
  P.MN = MN_;
  P.SPT = SPT_;
  if (P.T > kCtlMaxTrees) raise(BST_E_CONFIG, "more trees per round than the device grower holds");
  const uint32_t max_leaves = c.growth == BST_GROW_DEPTHWISE ? (1u << c.max_depth) : c.max_leaves;
  P.MNT = 2 * max_leaves + 1;
  P.growth = c.growth == BST_GROW_DEPTHWISE ? 0 : 1;
  P.W = P.growth == 0 ? std::max<uint32_t>(1, max_leaves / 2) : 1;
  P.WO = P.growth == 0 ? max_leaves : max_leaves + 1;
  P.max_depth = c.max_depth;
  P.max_leaves = c.max_leaves;
  P.NG = (uint32_t)dd_.NG;
  P.E = (uint32_t)dd_.E;
  P.FG = (uint32_t)dd_.FG;
  P.hist_tile = kHistTileD;
  P.hist_tile_min = tune_.tile_min;
  P.hist_target = (uint32_t)std::max(64, (int)tune_.blocks_per_cu * ctx_->n_cu);
  if (const char* e = std::getenv("BST_HIST_TILE")) {
    P.hist_tile = (uint32_t)std::max(512, std::atoi(e));
    P.hist_tile_min = P.hist_tile;  // fixed tile
  }
  P.part_tile = kPartTileD;
  P.part_tile_small = kFusedTileRowsSmall;
  P.seg_tile = kSegTileD;
  P.rightmost = (uint32_t)sp_.rightmost;
  P.stride = n_;
  P.lr = lr_;
  P.lambda = (double)c.lambda;
  P.alpha = (double)c.alpha;
  P.min_gain = (double)c.min_gain;
  P.out0 = 0;
  st->steps = P.growth == 0 ? (int)c.max_depth : (int)(c.max_leaves > 0 ? c.max_leaves - 1 : 0);
  const uint32_t T = P.T;
  // Tiles of a step: at most one per kFusedTileRowsSmall rows plus one per node.
  const uint32_t tiles_n = (uint32_t)((n_ + kFusedTileRowsSmall - 1) / kFusedTileRowsSmall);
  const uint32_t hist_n = (uint32_t)((n_ + P.hist_tile - 1) / P.hist_tile);
  const uint32_t segs_n = (uint32_t)((n_ + kSegTileD - 1) / kSegTileD);
  // Worst cases per step: every tree expands W nodes whose rows sum to n.
  st->cap_items = T * P.W;
  st->cap_tiles = T * (tiles_n + P.W);
  // Adaptive tiles give at most ~hist_target items for the rows, or the
  // largest-tile count, plus one partial tile per node.
  st->cap_hist = std::max<uint32_t>(P.hist_target + P.NG, T * hist_n * P.NG) + T * P.W * P.NG + P.NG;
  st->cap_fin = std::max<uint32_t>(T, T * P.W);
  st->cap_split = std::max<uint32_t>(T, 2 * T * P.W);
  st->cap_lsegs = T * (segs_n + MN_);
  auto& s = *st;
  for (auto* b : {&s.open_node, &s.open_tnode, &s.open_depth}) b->alloc((size_t)T * P.WO);
  for (auto* b : {&s.next_node, &s.next_tnode, &s.next_depth}) b->alloc((size_t)T * 2 * P.W);
  for (auto* b : {&s.ch_node, &s.ch_tnode, &s.ch_depth}) b->alloc((size_t)T * P.W);
  s.pbegin.alloc((size_t)T * MN_);
  s.pslot.alloc((size_t)T * MN_);
  s.free_slot.alloc((size_t)T * SPT_);
  s.pbuf.alloc((size_t)T * MN_);
  s.cursor.alloc((size_t)2 * s.cap_items);
  // Row sampling (row.rs:141-270) and column sampling (column.rs:266-283).
  {
    SampleParams& SP = s.SP;
    SP.T = T;
    SP.n = n_;
    SP.stride = n_;
    SP.seed = c.seed;
    const bool uniform = !(std::fabs(c.subsample - 1.0f) < 1e-6f);
    if (opt_.ext.goss_top_rate > 0.0f) {
      const size_t n = (size_t)n_;
      const size_t top_k =
          std::min<size_t>(std::max<size_t>((size_t)((float)n * opt_.ext.goss_top_rate), 1), n - 1);
      const size_t other_k = std::max<size_t>((size_t)((float)n * opt_.ext.goss_other_rate), 1);
      s.sample_mode = kSampleGoss;
      SP.k_mag = (uint32_t)(n - top_k);
      SP.k_key = (uint32_t)(other_k - 1);
      SP.mult = (float)(n - top_k) / (float)other_k;
      s.goss_warmup = (uint32_t)(1.0f / c.learning_rate);
    } else if (uniform) {
      const size_t tc =
          std::min<size_t>(std::max<size_t>((size_t)((float)n_ * c.subsample), 1), (size_t)n_);
      s.sample_mode = kSampleUniform;
      SP.k_key = (uint32_t)(tc - 1);
    }
    SP.mode = s.sample_mode;
    s.col_sample = !(std::fabs(c.colsample_bytree - 1.0f) < 1e-6f);
    s.n_feats = (int)P.E;
    if (s.col_sample) {
      s.n_feats = (int)std::min<uint32_t>(
          P.E, colsample_count(P.E, c.colsample_bytree));
      s.feats_dev.alloc((size_t)T * P.E);
    }
    s.round.alloc(1);
    launch_fill_u32(s_, s.round.p, 0u, 1);
    if (s.sample_mode != kSampleNone) {
      s.samp.alloc(T);
      const int64_t nb = (n_ + kSampleBlockRows - 1) / kSampleBlockRows;
      s.bcount.alloc((size_t)T * 4 * nb);
    }
  }
  s.out_bytes = tree_out_layout(T, P.MNT, nullptr, nullptr);
  s.out.alloc(s.out_bytes);
  tree_out_layout(T, P.MNT, s.out.p, &s.O);
  s.pool = tls_pool;
  s.metric_bytes = sizeof(double) * 2 * kMetricBlocks * opt_.K;
  for (int i = 0; i < 2; i++) {
    s.h_out[i] = (char*)s.pinned(s.out_bytes);
    tree_out_layout(T, P.MNT, s.h_out[i], &s.H[i]);
    s.h_metric[i] = (double*)s.pinned(s.metric_bytes);
    BST_HIP(hipEventCreateWithFlags(&s.done[i], hipEventDisableTiming));
  }
  s.items.alloc(s.cap_items);
  s.tiles.alloc(s.cap_tiles);
  s.tstat.alloc(s.cap_tiles);
  s.pend.alloc((size_t)5 * s.cap_items);
  s.hitems.alloc(s.cap_hist);
  s.fin.alloc((size_t)4 * s.cap_fin);
  s.split_nodes.alloc(s.cap_split);
  s.split_slots.alloc(s.cap_split);
  s.split_minus.alloc(s.cap_split);
  s.rows.alloc(T);
  s.leaf_segs.alloc(s.cap_lsegs);
  s.partial.alloc((size_t)s.cap_hist * 256 * P.FG * 2);
  s.thr_up.alloc(ds_->thr_up.size());
  h2d(s_, s.thr_up.p, ds_->thr_up.data(), ds_->thr_up.size());
  s.eff_to_orig.alloc(ds_->eff_to_orig.size());
  h2d(s_, s.eff_to_orig.p, ds_->eff_to_orig.data(), ds_->eff_to_orig.size());
  std::vector<uint32_t> rows(T, (uint32_t)n_);
  h2d(s_, s.rows.p, rows.data(), rows.size());
  std::vector<uint32_t> outs(T);
  for (uint32_t t = 0; t < T; t++) outs[t] = t;
  h2d(s_, tree_output_.p, outs.data(), outs.size());
  // The captured round expects absmax zeroed (each round's qscale re-zeroes it).
  launch_fill_u32(s_, absmax_.p, 0u, (int64_t)absmax_.n);
  DeviceBufs& B = s.B;
  B.ctl = s.O.ctl;
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
  B.pbuf = s.pbuf.p;
  B.free_slot = s.free_slot.p;
  B.leaf_val = leaf_val_.p;
  B.tr_split = s.O.tr_split;
  B.tr_left = s.O.tr_left;
  B.tr_right = s.O.tr_right;
  B.tr_cats = s.O.tr_cats;
  B.tr_thr = s.O.tr_thr;
  B.tr_leaf = s.O.tr_leaf;
  B.tr_gain = s.O.tr_gain;
  B.tr_cover = s.O.tr_cover;
  B.tr_dleft = s.O.tr_dleft;
  B.tr_isleaf = s.O.tr_isleaf;
  B.tr_stype = s.O.tr_stype;
  B.items = s.items.p;
  B.tiles = s.tiles.p;
  B.tstat = s.tstat.p;
  B.pend = s.pend.p;
  B.hitems = s.hitems.p;
  B.fin = s.fin.p;
  B.split_nodes = s.split_nodes.p;
  B.split_slots = s.split_slots.p;
  B.split_minus = s.split_minus.p;
  B.leaf_segs = s.leaf_segs.p;
  B.counts = s.O.counts;
  B.node_sum = node_sum_.p;
  B.node_count = node_count_.p;
  B.node_split = node_split_.p;
  B.node_key = node_key_.p;
  B.cand = cand_.p;
  B.thr_up = s.thr_up.p;
  B.eff_to_orig = s.eff_to_orig.p;
  B.fmeta = dd_.meta;
  B.qexp = qexp_.p;
  B.cursor = s.cursor.p;
  BST_HIP(hipStreamSynchronize(s_));
  dev_ = std::move(st);
}

// Every launch of one round (T trees = all K outputs), no host sync.
void Trainer::device_enqueue_round(const float* y, const float* w, bool sampled) {
  DevState& S = *dev_;
  // Depth-wise rounds over every row apply the last level's leaf values in
  // its partition; sampled rounds update every row by traversal instead.
  CtlParams P = S.P;
  P.leaf_apply = !sampled && P.growth == 0 ? 1u : 0u;
  const DeviceBufs& B = S.B;
  const auto& cfg = opt_.cfg;
  const int K = opt_.K;
  const auto& d = dd_;
  // The histogram kernel keeps one block per CU resident (two with 16-feature
  // groups, LDS-limited), so a grid of that many blocks walking the items is
  // as fast as a larger one and costs nothing on steps with few items.  A
  // multiple of 8 keeps the XCD mapping.
  const int g_hist = grid_for(S.cap_hist, (hist_blocks_resident(d.FG) * ctx_->n_cu + 7) / 8 * 8);
  const int g_tiles = grid_for(S.cap_tiles, kMaxGrid);
  const int g_split = grid_for(S.cap_split, 512);
  const int g_lsegs = grid_for(S.cap_lsegs, kMaxGrid);
  uint32_t* cnt = B.counts;
  const int n_feats = S.n_feats;
  // Without column sampling every tree scans columns 0 .. E-1 (no list).
  const uint32_t* feats = S.col_sample ? S.feats_dev.p : nullptr;
  const int feat_stride = S.col_sample ? (int)P.E : 0;
  launch_objective(y, w);
  if (sampled) {
    // Sampled rows (index order) and their (g, h) go straight to slot
    // positions; rows[t] and the quantisation maxima cover only them.
    launch_sample_rows(s_, S.SP, S.round.p, g_.p, h_.p, S.samp.p, S.bcount.p, idx_.p, ogh_.p,
                       S.rows.p, absmax_.p);
  }
  // Unsampled trees take every row (no row counts to read).
  launch_ctl_init(s_, P, B, sampled ? S.rows.p : nullptr, absmax_.p, 2 * K, n_, qexp_.p);
  if (!sampled)
    launch_init_rows_sums(s_, idx_.p, ogh_.p, g_.p, h_.p, n_, (int)P.T, qexp_.p, MN_,
                          node_sum_.p);
  if (S.col_sample)
    launch_sample_columns(s_, opt_.cfg.seed, S.round.p, (int)P.T, (int)P.E, n_feats,
                          S.feats_dev.p);
  auto hist_step = [&](bool gather, bool root_sums) {
    int pi = ctx_->prof_begin(0, 0.0);
    launch_hist(s_, d, B.hitems, g_hist, idx_.p, ogh_.p, 0, qexp_.p, hist_.p, slot_elems_,
                root_sums ? node_sum_.p : nullptr, gather, 0, cnt + C_HIST, S.partial.p,
                idx_tmp_.p, ogh_tmp_.p);
    ctx_->prof_end(pi);
    launch_hist_finish(s_, B.fin, (int)S.cap_fin, (int)P.NG, (int)P.FG, S.partial.p, hist_.p,
                       slot_elems_, cnt + C_FIN);
    pi = ctx_->prof_begin(1, 0.0);
    launch_split(s_, d, B.split_nodes, B.split_slots, g_split, feats, n_feats, hist_.p,
                 slot_elems_, node_sum_.p, node_count_.p, qexp_.p, MN_, sp_, node_key_.p,
                 cand_.p, node_split_.p, cnt + C_SPLIT, /*finalize=*/false, feat_stride,
                 B.split_minus);
    ctx_->prof_end(pi);
  };
  // Positions of tree slot t start at t * n, so only slot 0 of an unsampled
  // round may skip the gather.  Sampled roots take their totals from the
  // histogram (the rows were chosen after the quantisation scale).
  hist_step(sampled || P.T > 1, sampled);
  for (int step = 0; step < S.steps; step++) {
    launch_ctl_plan(s_, P, B, /*advance=*/step > 0);
    int pi = ctx_->prof_begin(2, 0.0);
    launch_partition_fused(s_, d, B.items, B.tiles, g_tiles, node_split_.p, idx_.p, ogh_.p,
                           idx_tmp_.p, ogh_tmp_.p, B.cursor, B.tstat, cnt + C_TICKET,
                           cnt + C_ERR, cnt + C_TILES, tree_output_.p, preds_.p, n_, P.MN);
    ctx_->prof_end(pi);
    // The last step's children are never split (depth limit, or the leaf
    // budget is spent), so their histograms are not built; when its
    // partition applied their values, the finalize makes them leaves and
    // this step's children controller has nothing left to do.
    const bool last = step + 1 == S.steps;
    if (!(last && P.leaf_apply)) launch_ctl_children(s_, P, B);
    if (!last) hist_step(true, false);
  }
  launch_ctl_finalize(s_, P, B, S.steps == 0 ? 0 : P.leaf_apply ? 2 : 1, S.round.p);
  if (sampled) {
    // Rows outside the sample get the tree too: traverse the raw features
    // (trainer.rs:331-339).
    launch_tree_update_raw(s_, (int)P.T, P.MNT, B.tr_split, B.tr_thr, B.tr_left, B.tr_right,
                           B.tr_dleft, B.tr_isleaf, B.tr_stype, B.tr_cats, B.tr_leaf, d.raw, n_,
                           tree_output_.p, preds_.p);
  } else {
    launch_leaf_update(s_, B.leaf_segs, g_lsegs, leaf_val_.p, idx_.p, 0, preds_.p,
                       tree_output_.p, n_, cnt + C_LSEGS, idx_tmp_.p);
  }
  if (valid_n_ > 0) {
    // Validation predictions += the round's trees (trainer.rs:341-352):
    // the same traversal of raw features, over the validation rows.
    launch_tree_update_raw(s_, (int)P.T, P.MNT, B.tr_split, B.tr_thr, B.tr_left, B.tr_right,
                           B.tr_dleft, B.tr_isleaf, B.tr_stype, B.tr_cats, B.tr_leaf, valid_X_,
                           valid_n_, tree_output_.p, vpreds_.p);
  }
}

// Launch (or capture and launch) one round, wait, and return the trees.
void Trainer::device_launch(const float* y, const float* w, bool metric) {
  using clk = std::chrono::steady_clock;
  static const bool trace = std::getenv("BST_TRACE_SETUP") != nullptr;
  const auto t0 = clk::now();
  if (!dev_) device_setup();
  DevState& S = *dev_;
  const auto t1 = clk::now();
  const bool sampled = S.sample_mode != kSampleNone &&
                       !(S.sample_mode == kSampleGoss && S.host_round < S.goss_warmup);
  S.host_round++;
  const int gi = sampled ? 1 : 0;
  const bool use_graph = !ctx_->profile;  // profiled runs time each launch with events
  if (use_graph) {
    if (!S.exec[gi]) {
      BST_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
      device_enqueue_round(y, w, sampled);
      BST_HIP(hipStreamEndCapture(s_, &S.graph[gi]));
      const auto t2 = clk::now();
      BST_HIP(hipGraphInstantiate(&S.exec[gi], S.graph[gi], nullptr, nullptr, 0));
      if (trace) {
        const auto t3 = clk::now();
        std::fprintf(stderr, "[bst] device setup %.3f ms, capture %.3f ms, instantiate %.3f ms\n",
                     std::chrono::duration<double, std::milli>(t1 - t0).count(),
                     std::chrono::duration<double, std::milli>(t2 - t1).count(),
                     std::chrono::duration<double, std::milli>(t3 - t2).count());
      }
    }
    BST_HIP(hipGraphLaunch(S.exec[gi], s_));
  } else {
    device_enqueue_round(y, w, sampled);
  }
  const int slot = (int)(S.launched & 1u);
  if (metric) {
    if (valid_n_ > 0)  // the validation metric is unweighted
      metric_launch(vpreds_.p, valid_y_, nullptr, valid_n_, S.h_metric[slot]);
    else
      metric_launch(preds_.p, y, w, n_, S.h_metric[slot]);
  }
  d2h(s_, S.h_out[slot], S.out.p, S.out_bytes);
  BST_HIP(hipEventRecord(S.done[slot], s_));
  S.launched++;
}

void Trainer::device_collect(std::vector<TreeBuild>& out, double* metric_value) {
  DevState& S = *dev_;
  if (S.collected >= S.launched) raise(BST_E_INVALID, "no device round in flight");
  const int slot = (int)(S.collected & 1u);
  BST_HIP(hipEventSynchronize(S.done[slot]));
  BST_HIP(hipGetLastError());
  S.collected++;
  const TreeOut& H = S.H[slot];
  if (H.counts[C_ERR])
    raise(BST_E_DEVICE, "a partition tile's look-back did not finish within its wait bound");
  if (metric_value) *metric_value = metric_finish(S.h_metric[slot]);
  if (ctx_->profile) {
    // Algorithmic bytes of the round's histogram launches, spread evenly.
    const double rows = (double)H.counts[C_HROWS];
    const int nodes = (int)H.counts[C_HNODES];
    const double bytes = hist_bytes(rows, nodes, true);
    int nh = 0;
    for (auto& p : ctx_->pending)
      if (p.kind == 0) nh++;
    for (auto& p : ctx_->pending)
      if (p.kind == 0) p.bytes = bytes / std::max(nh, 1);
    ctx_->prof_drain();
  }
  stats.hist_rows += H.counts[C_HROWS];
  const CtlParams& P = S.P;
  out.assign(P.T, TreeBuild());
  for (uint32_t t = 0; t < P.T; t++) {
    TreeBuild& tb = out[t];
    const uint32_t nn = H.ctl[t].n_tnodes;
    const size_t o = (size_t)t * P.MNT;
    for (uint32_t i = 0; i < nn; i++) {
      tb.alloc();
      tb.split_index[i] = H.tr_split[o + i];
      tb.left[i] = H.tr_left[o + i];
      tb.right[i] = H.tr_right[o + i];
      tb.thr[i] = H.tr_thr[o + i];
      tb.leaf[i] = H.tr_isleaf[o + i] ? H.tr_leaf[o + i] : 0.0f;
      tb.gain[i] = H.tr_gain[o + i];
      tb.cover[i] = H.tr_cover[o + i];
      tb.dleft[i] = H.tr_dleft[o + i];
      tb.is_leaf[i] = H.tr_isleaf[o + i];
      tb.stype[i] = H.tr_stype[o + i];
      if (!tb.is_leaf[i] && tb.stype[i] == 1) {
        const uint32_t* c = H.tr_cats + (o + i) * kMaxCatWords;
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
