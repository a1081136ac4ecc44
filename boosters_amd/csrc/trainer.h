// trainer.h — host orchestration of histogram GBDT training on the device.
#pragma once
#include <memory>
#include <queue>

#include "host_common.h"

namespace bst {

// MutableTree (repr/gbdt/mutable_tree.rs:16-34) on the host.
struct TreeBuild {
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

// NodeCandidate (training/gbdt/expansion.rs:242-320); sums are exact fixed point.
struct Cand {
  uint32_t node;       // partition id (tree-local)
  uint32_t tree_node;  // tree builder id
  uint32_t depth;
  DevSplit split;
  long long g, h;
  uint32_t count;
};
struct CandLess {  // BinaryHeap order: gain, then partition id
  bool operator()(const Cand& a, const Cand& b) const {
    if (a.split.gain < b.split.gain) return true;
    if (a.split.gain > b.split.gain) return false;
    return a.node < b.node;
  }
};

// Objective::needs_leaf_renew, objectives/mod.rs:341-355.
inline bool needs_leaf_renew(int objective) {
  return objective == BST_OBJ_ABSOLUTE || objective == BST_OBJ_PINBALL;
}

struct TrainOptions {
  bst_gbdt_config cfg;
  bst_ext_config ext;
  int K = 1;  // outputs
  ObjParams obj;              // objective payloads (pseudo-Huber delta, pinball alphas)
  std::vector<float> alphas;  // the config's alphas, owned (pinball outputs, quantile metric)
  // LambdaRank: query group row offsets [n_groups + 1] and rank discounts
  // 1 / log2(2 + r) for r below the largest group.
  std::vector<int64_t> group_off;
  std::vector<double> rank_disc;
};

struct TrainStats {
  double ms_total = 0;
  int64_t hist_rows = 0;
  int64_t n_trees = 0;
};

class Trainer {
 public:
  Trainer(bst_ctx* ctx, bst_dataset* ds, const TrainOptions& opt);
  // targets/weights/valid are device pointers; valid may be null.
  bst_model* train(const float* y_dev, const float* w_dev, const float* vX_dev,
                   const float* vy_dev, int64_t n_valid);
  TrainStats stats;
  std::vector<double> metric_history;

 private:
  struct TreeState;
  void grow_round(std::vector<TreeState>& trees);
  void expand_batch(std::vector<TreeState>& trees, std::vector<std::pair<int, Cand>>& chosen);
  void root(std::vector<TreeState>& trees);
  void make_leaf(TreeState& ts, const Cand& c);
  void renew_round(std::vector<TreeState>& trees, const float* y, const float* w);
  // Column sampling by level (host-driven schedules): the tree's features and
  // the (round, output) whose level subsets set_level_features draws.
  bool level_sampling() const;
  void set_level_features(uint32_t depth);
  std::vector<uint32_t> tree_feats_h_;
  // Validation set of the current train (device pointers; null without one).
  const float* valid_X_ = nullptr;
  const float* valid_y_ = nullptr;
  int64_t valid_n_ = 0;
  uint32_t cur_round_ = 0;
  int cur_output_ = 0;
  void finish_trees(std::vector<TreeState>& trees, bst_model* m);
  // Training metric: weights w (nullable); validation metric: unweighted.
  double eval_metric(const float* preds, const float* y, const float* w, int64_t n);
  uint32_t alloc_slot(TreeState& ts);
  // Graph schedule (trainer_device.cpp): the whole round on the device.
  struct DevState;
  struct DevStateDelete {
    void operator()(DevState* p) const;
  };
  std::unique_ptr<DevState, DevStateDelete> dev_;
  bool graph_capable() const;  // the configuration can run on the graph schedule
  bool device_schedule_ok() const;
  void device_setup();
  void device_enqueue_round(const float* y, const float* w, bool sampled);
  // Queue one boosting round (graph replay, then the training metric's
  // partial sums if `metric`, then the trees' copy to a host buffer of its
  // own); rounds queued back to back run while the host reads earlier ones.
  void device_launch(const float* y, const float* w, bool metric);
  // Wait for the oldest queued round and read its trees (and metric partials).
  void device_collect(std::vector<TreeBuild>& out, double* metric_value);
  // Training metric (eval.rs:218-260): partial sums on the device into
  // `host_part` (2 doubles per block and output, nb blocks), and the value
  // from them.
  int metric_kind() const;
  void metric_launch(const float* preds, const float* y, const float* w, int64_t n,
                     double* host_part);
  double metric_finish(const double* part) const;
  static constexpr int kMetricBlocks = 1024;
  // Algorithmic bytes: each histogrammed row reads its E bins (1 B, or
  // 1.125 B with the NaN hi plane) plus its gradient pair (8 B) and row id
  // (4 B when gathered); each built node writes E x 256 (g, h) int64 pairs.
  double hist_bytes(double rows, int nodes, bool gather) const {
    const double b = ds_->d.hi_off >= 0 ? 1.125 : 1.0;
    return rows * (ds_->d.E * b + (gather ? 12.0 : 8.0)) + (double)nodes * ds_->d.E * 256.0 * 16.0;
  }
  double split_bytes(int nodes) const { return (double)nodes * ds_->d.E * 256.0 * 16.0; }

  bst_ctx* ctx_;
  bst_dataset* ds_;
  HistTuning tune_;  // histogram launch shape (autotuned or fixed)
  DatasetDev dd_;    // the dataset's device view with this run's feature grouping
  TrainOptions opt_;
  hipStream_t s_;
  int64_t n_;
  int T_;              // trees grown together (K for class batching, else 1)
  uint32_t MN_;        // node ids per tree
  uint32_t SPT_;       // histogram slots per tree
  int64_t slot_elems_;
  SplitParams sp_;
  bool rightmost_;
  int sched_;
  float lr_;
  std::vector<int32_t> qexp_h_;

  DevBuf<float> preds_, g_, h_, vpreds_;
  DevBuf<int64_t> group_off_;  // LambdaRank
  DevBuf<double> rank_disc_;
  // The round's gradient launch (objective dispatch).
  void launch_objective(const float* y, const float* w);
  DevBuf<uint32_t> absmax_;
  DevBuf<int32_t> qexp_;
  DevBuf<uint32_t> idx_, idx_tmp_;
  DevBuf<float2> ogh_, ogh_tmp_;
  DevBuf<uint8_t> flag_;
  DevBuf<long long> node_sum_, hist_;
  DevBuf<uint32_t> node_count_;
  DevBuf<unsigned long long> node_key_;
  DevBuf<DevSplit> node_split_, cand_;
  DevBuf<uint32_t> tile_left_, tile_off_, item_left_;
  DevBuf<float> leaf_val_;
  DevBuf<uint32_t> tree_output_;
  DevBuf<uint32_t> feats_;
  // staging for per-step work lists
  DevBuf<HistItem> d_hist_items_;
  DevBuf<PartItem> d_part_items_;
  DevBuf<PartTile> d_part_tiles_;
  DevBuf<uint32_t> d_u32a_, d_u32b_, d_u32c_, d_u32d_;
  DevBuf<Segment> d_segs_;
  DevBuf<uint32_t> d_sampled_;
  DevBuf<double> d_metric_;
  // Leaf renewal (quantile objectives): leaves, keys, sorted keys, segment
  // offsets, sort scratch, renewed values.
  DevBuf<RenewLeaf> d_renew_leaves_;
  DevBuf<unsigned long long> renew_keys_, renew_sorted_;
  DevBuf<uint32_t> renew_seg_;
  DevBuf<uint8_t> renew_temp_;
  DevBuf<float> renew_out_;
  std::vector<uint32_t> feats_h_;
};

}  // namespace bst
