/*
 * bst_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of egordm/boosters' histogram-GBDT path, used as the parity
 * checker for the HIP engine and as the CPU baseline timed beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * engine (boosters_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * crates/boosters/src/).  Shared POD structs come from include/boosters_amd.h.
 */
#ifndef BST_ORACLE_H
#define BST_ORACLE_H

#include "../include/boosters_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_binned ora_binned;
typedef struct ora_model ora_model;

/* A split decision, SplitInfo (training/gbdt/split/types.rs:34-55). */
typedef struct ora_split {
  uint32_t feature; /* effective column */
  float gain;       /* -inf = invalid   */
  int32_t default_left;
  int32_t is_categorical;
  uint32_t bin;           /* numeric: rows with bin <= bin go left */
  uint32_t left_cats[8];  /* categorical: bin ids going left (<= 256) */
} ora_split;

const char* ora_last_error(void);
int ora_num_threads(void);
void ora_set_num_threads(int n);
/* 0 = glibc expf (reference-faithful), 1 = f64 exp rounded to f32 (engine). */
void ora_set_exp_mode(int mode);

/* Binning: BinnedDataset::from_dataset, data/binned/dataset.rs:207-309. */
int ora_bin(const float* feat_major, int64_t n_rows, int32_t n_feat, const uint8_t* is_categorical,
            const bst_binning_config* cfg, ora_binned** out);
void ora_binned_free(ora_binned* b);
int ora_feature_info(const ora_binned* b, int32_t feature, uint32_t* n_bins, int32_t* missing_type,
                     uint32_t* default_bin, int32_t* is_categorical, int32_t* is_trivial);
int ora_bin_bounds(const ora_binned* b, int32_t feature, double* out, int64_t cap);
int ora_get_bins(const ora_binned* b, uint16_t* out);
int ora_hist_layout(const ora_binned* b, int32_t* n_effective, uint32_t* eff_to_orig,
                    uint32_t* eff_n_bins, uint32_t* eff_offset);

/* BinMapper::value_to_bin on one feature's mapper, bin_mapper.rs:210-258. */
uint32_t ora_value_to_bin(const ora_binned* b, int32_t feature, double value);
/* GainParams::compute_leaf_weight, split/gain.rs:151-168. */
float ora_leaf_weight(const bst_gbdt_config* cfg, double grad_sum, double hess_sum);

/* Objectives: objectives/mod.rs:236-298. */
int ora_gradients(int32_t objective, int32_t n_outputs, const float* preds, const float* targets,
                  const float* weights, int64_t n, float* grad, float* hess);
/* LambdaRank (no reference counterpart; spec in bst_oracle.cpp). */
int ora_lambdarank_gradients(const float* preds, const float* targets, const float* weights,
                             int64_t n, const int32_t* group_sizes, int64_t n_groups,
                             float* grad, float* hess);
int ora_base_score(int32_t objective, int32_t n_outputs, const float* targets,
                   const float* weights, int64_t n, float* out);
/* The same with the objective's payload from a config (pinball alphas,
 * pseudo-Huber delta); n_outputs follows Objective::n_outputs. */
int ora_gradients_cfg(const bst_gbdt_config* cfg, const float* preds, const float* targets,
                      const float* weights, int64_t n, float* grad, float* hess);
int ora_base_score_cfg(const bst_gbdt_config* cfg, const float* targets, const float* weights,
                       int64_t n, float* out);
/* compute_weighted_quantile (objectives/regression.rs:307-333, base scores)
 * and the leaf-renewal quantiles over a leaf's rows (utils.rs:17-126;
 * weights nullable = unweighted). */
float ora_weighted_quantile(const float* values, const float* weights, int64_t n, float alpha);
float ora_quantile_indexed(const uint32_t* indices, uint32_t n, const float* values,
                           const float* weights, float alpha);

/* Histogram of one node in partition order, (f64, f64) bins: ops.rs:79-283. */
int ora_histogram(const ora_binned* b, const float* grad, const float* hess, const uint32_t* rows,
                  int64_t n_rows, double* out_gh);

/* GreedySplitter::find_split, split/find.rs:113-270. hist_gh in the layout of
 * ora_hist_layout; features nullable = all effective columns. */
int ora_find_split(const ora_binned* b, const double* hist_gh, double parent_grad,
                   double parent_hess, uint32_t parent_count, const bst_gbdt_config* cfg,
                   int32_t n_threads, const uint32_t* features, int32_t n_features,
                   ora_split* out);

/* RowPartitioner::split, partition.rs:296-382 (stable; returns left count). */
int ora_partition(const ora_binned* b, uint32_t* rows, int64_t n_rows, const ora_split* split,
                  int64_t* left_count);

/* GBDTTrainer::train, training/gbdt/trainer.rs:162-401. */
int ora_train(const ora_binned* b, const float* feat_major, const float* targets,
              int64_t n_targets, const float* weights, const float* valid_feat_major,
              const float* valid_targets, int64_t n_valid, const bst_gbdt_config* cfg,
              const bst_ext_config* ext, ora_model** out);
int ora_model_info(const ora_model* m, bst_model_info* info);
int ora_model_export(const ora_model* m, bst_forest_arrays* out);
int ora_model_metric(const ora_model* m, double* out, int64_t cap, int64_t* n_rounds);
int ora_model_stats(const ora_model* m, int64_t* hist_rows_built, double* ms_total);
void ora_model_free(ora_model* m);

/* Prediction over a forest in exchange format: predictor.rs:197-302,
 * tree_view.rs:119-160, transform.rs:23-115. */
int ora_predict(const bst_model_info* info, const bst_forest_arrays* f, const float* feat_major,
                int64_t n_rows, int32_t n_feat, int32_t raw, float* out, uint32_t* leaf_idx);

/* Training predictions as the reference's trainer forms them (trainer.rs:
 * 323-339): base score plus, tree by tree in forest order, the leaf each
 * training row of `b` reaches by the partitioner's rule on its bins
 * (partition.rs:321-337; grower.rs:208-219), or, for trees with
 * by_threshold[t] != 0 (rounds trained on a row sample), by traversing the
 * raw features X with the thresholds.  out: [n_groups][n_rows]. */
int ora_partition_predict(const ora_binned* b, const bst_model_info* info,
                          const bst_forest_arrays* f, const uint8_t* by_threshold,
                          const float* X, float* out);

/* Sampling rules, exposed for their pins.  SplitMix64 steps from a state
 * (the mixer of dataset.rs:598-629's edge sample); the edge sample itself. */
void ora_splitmix64_stream(uint64_t state, int64_t count, uint64_t* out);
int64_t ora_binning_sample(int64_t n, uint64_t cnt, uint32_t* out);
/* The EFB row sample's generator (bundling.rs:203-251): xoshiro256++ outputs
 * from an explicit state [4] (nullable: seed_from_u64(seed) instead), rand's
 * inclusive range sampling over scripted draws (value, draws consumed), the
 * row sample, and the bundles a binned dataset made (sizes [n_bundles],
 * features concatenated, original indices). */
void ora_xoshiro256pp(const uint64_t* state, uint64_t seed, int64_t count, uint64_t* out);
int ora_sample_inclusive(const uint64_t* draws, int64_t n_draws, uint64_t low, uint64_t high,
                         uint64_t* value, int64_t* used);
int64_t ora_efb_sample_rows(int64_t n, int64_t max_samples, uint64_t seed, int64_t* out,
                            int64_t cap);
int32_t ora_bundles(const ora_binned* b, int32_t* sizes, uint32_t* features, int64_t cap);
/* The row sample of (round, output) as ora_train draws it: uniform
 * (row.rs:170-208) or GOSS (row.rs:211-270, after the warm-up of :109-112).
 * g_out/h_out: the edited pairs; rows_out [n] the kept rows in order;
 * *n_rows_out = -1 when the round uses every row. */
int ora_sample_rows(const float* g, const float* h, int64_t n, const bst_gbdt_config* cfg,
                    const bst_ext_config* ext, uint32_t round, uint32_t output, float* g_out,
                    float* h_out, uint32_t* rows_out, int64_t* n_rows_out);
/* The columns of tree (round, output), column.rs:180-186,266-283. */
int ora_sample_cols(uint32_t n_feat, const bst_gbdt_config* cfg, uint32_t round, uint32_t output,
                    uint32_t* out, int32_t* k);

/* The level subset of a tree's features at `depth` (column.rs:199-224). */
int ora_sample_level(const uint32_t* tree_feats, uint32_t n_tree_feats, const bst_gbdt_config* cfg,
                     uint32_t round, uint32_t output, uint32_t depth, uint32_t* out, int32_t* k);

/* Metrics over transformed predictions [K][n]: metrics/{classification,regression}.rs. */
double ora_metric(int32_t metric, const float* preds, int32_t n_outputs, const float* targets,
                  const float* weights, int64_t n);
/* With the metric's payload (quantile alphas, accuracy threshold, Huber
 * delta) from a config. */
double ora_metric_cfg(const bst_gbdt_config* cfg, int32_t metric, const float* preds,
                      int32_t n_outputs, const float* targets, const float* weights, int64_t n);
/* default_metric_for_objective, metrics/mod.rs:213-232. */
int ora_default_metric(int32_t objective);

#ifdef __cplusplus
}
#endif
#endif
