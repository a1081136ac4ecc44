/*
 * boosters_amd.h — C ABI of the MI355X (gfx950) histogram-GBDT engine.
 *
 * This is the drop-in boundary for egordm/boosters' histogram training and
 * tree-ensemble prediction path.  Every entry point takes plain pointers and
 * sizes, returns an int status (0 = ok, negative = error) and never lets a C++
 * exception or abort cross the boundary; bst_last_error() returns a
 * thread-local message for the last failure on the calling thread.
 *
 * Reference seams replaced (paths relative to crates/boosters/src/):
 *   training   GBDTModel::train            model/gbdt/model.rs:115-143
 *              GBDTTrainer::train           training/gbdt/trainer.rs:162-170
 *   predict    GBDTModel::predict/_raw      model/gbdt/model.rs:217-258
 *              UnrolledPredictor6::predict_into  inference/gbdt/predictor.rs:197-247
 *   config     GBDTConfig                    model/gbdt/config.rs:111-226
 *   binning    BinnedDataset::from_dataset   data/binned/dataset.rs:207-309
 *   result     Forest<ScalarLeaf> / Tree::new repr/gbdt/forest.rs:31-36, tree.rs:65-100
 *
 * Host buffers passed in are borrowed and copied; device memory is owned by
 * the handles.  A bst_ctx owns one HIP stream and is not re-entrant; distinct
 * contexts are independent.  bst_server_submit is thread-safe.
 */
#ifndef BOOSTERS_AMD_H
#define BOOSTERS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BST_OK 0
#define BST_E_CONFIG (-1)  /* invalid config   (reference: ConfigError, config.rs:39-48) */
#define BST_E_TARGETS (-2) /* too few targets  (reference: trainer.rs:175-177 -> None)   */
#define BST_E_SHAPE (-3)   /* shape mismatch   (reference: predictor.rs:209-213 panic)  */
#define BST_E_OOM (-4)     /* device memory / serving admission budget exceeded          */
#define BST_E_DEVICE (-5)  /* HIP runtime error or no gfx950 device                      */
#define BST_E_FORMAT (-6)  /* corrupt / unsupported model bytes (persist/error.rs)       */
#define BST_E_INVALID (-7) /* null handle, bad argument                                  */

/* ---- enums mirroring the reference ------------------------------------- */
/* Objective enum, training/objectives/mod.rs:196-219.  Payloads live in
 * bst_gbdt_config: softmax n_classes, pinball alphas (one output per alpha),
 * pseudo-Huber delta. */
#define BST_OBJ_SQUARED 0
#define BST_OBJ_LOGISTIC 1
#define BST_OBJ_SOFTMAX 2
#define BST_OBJ_LAMBDARANK 3 /* extension: no counterpart in the reference */
#define BST_OBJ_ABSOLUTE 4   /* AbsoluteLoss: sign gradients, leaves renewed to medians */
#define BST_OBJ_HINGE 5      /* HingeLoss: labels {0, 1} as {-1, +1} */
#define BST_OBJ_PINBALL 6    /* PinballLoss { alphas }: leaves renewed to alpha-quantiles */
#define BST_OBJ_PSEUDO_HUBER 7 /* PseudoHuberLoss { delta } */
#define BST_OBJ_POISSON 8    /* PoissonLoss: log-space scores */

/* GrowthStrategy, training/gbdt/expansion.rs:46-57 */
#define BST_GROW_DEPTHWISE 0
#define BST_GROW_LEAFWISE 1

/* Metric selection; BST_METRIC_DEFAULT follows default_metric_for_objective,
 * training/metrics/mod.rs:213-232. */
#define BST_METRIC_DEFAULT (-1)
#define BST_METRIC_NONE 0
#define BST_METRIC_RMSE 1
#define BST_METRIC_LOGLOSS 2
#define BST_METRIC_MLOGLOSS 3
#define BST_METRIC_MAE 4                 /* regression.rs:55-87 */
#define BST_METRIC_MAPE 5                /* regression.rs:89-124 */
#define BST_METRIC_QUANTILE 6            /* regression.rs:126-174, alphas from the config */
#define BST_METRIC_HUBER 7               /* regression.rs:216-250, metric_huber_delta */
#define BST_METRIC_POISSON_DEVIANCE 8    /* regression.rs:176-214 */
#define BST_METRIC_ACCURACY 9            /* classification.rs:46-83, accuracy_threshold */
#define BST_METRIC_MARGIN_ACCURACY 10    /* classification.rs:85-122 */
#define BST_METRIC_MULTICLASS_ACCURACY 11 /* classification.rs:124-192 */

/* Output transform, model/transform.rs:23-31 */
#define BST_TRANSFORM_IDENTITY 0
#define BST_TRANSFORM_SIGMOID 1
#define BST_TRANSFORM_SOFTMAX 2

/* Engine schedule (extension; see bst_ext_config). */
#define BST_SCHED_PLAIN 0   /* one launch per node: hist, split, partition   */
#define BST_SCHED_BATCHED 1 /* level/class-batched fused launches            */
#define BST_SCHED_GRAPH 2   /* batched, boosting iteration replayed as a hipGraph */

/* ---- configuration ----------------------------------------------------- */
/* BinningConfig, data/binned/feature_analysis.rs:29-63 (field-for-field). */
typedef struct bst_binning_config {
  uint32_t max_bins;                    /* default 256     */
  float sparsity_threshold;             /* default 0.9     */
  int32_t enable_bundling;              /* default 1       */
  uint32_t max_categorical_cardinality; /* default 0       */
  uint64_t sample_cnt;                  /* default 200000  */
} bst_binning_config;

/* GBDTConfig, model/gbdt/config.rs:111-226 (field-for-field; enums as ints). */
typedef struct bst_gbdt_config {
  int32_t objective;          /* BST_OBJ_*, default squared              */
  uint32_t n_classes;         /* softmax only                            */
  int32_t metric;             /* BST_METRIC_*, default BST_METRIC_DEFAULT */
  uint32_t n_trees;           /* default 100 */
  float learning_rate;        /* default 0.3 */
  int32_t growth;             /* BST_GROW_*, default depth-wise */
  uint32_t max_depth;         /* default 6   */
  uint32_t max_leaves;        /* default 31  */
  uint32_t max_onehot_cats;   /* default 4   */
  float lambda;               /* default 1.0 */
  float alpha;                /* default 0.0 */
  float min_child_weight;     /* default 1.0 */
  float min_gain;             /* default 0.0 */
  uint32_t min_samples_leaf;  /* default 1   */
  float subsample;            /* default 1.0 */
  float colsample_bytree;     /* default 1.0 */
  float colsample_bylevel;    /* default 1.0 */
  bst_binning_config binning; /* default BinningConfig::default() */
  uint32_t early_stopping_rounds; /* 0 = disabled (reference: Option<u32>) */
  uint64_t cache_size;        /* default 8 (accepted; device keeps every node histogram) */
  uint64_t seed;              /* default 42 */
  int32_t verbosity;          /* default 0  */
  /* Objective and metric payloads (enum data of Objective / Metric). */
  float huber_delta;          /* PseudoHuberLoss { delta }, default 1.0 */
  uint32_t n_alphas;          /* PinballLoss { alphas } / Metric::Quantile: count */
  const float* alphas;        /* [n_alphas], each in (0, 1); borrowed for the call */
  float accuracy_threshold;   /* Metric::Accuracy { threshold }, default 0.5 */
  double metric_huber_delta;  /* Metric::Huber { delta }, default 1.0 */
} bst_gbdt_config;

/* Extension knobs that have no counterpart in GBDTConfig. */
typedef struct bst_ext_config {
  int32_t schedule;        /* BST_SCHED_*, default BST_SCHED_GRAPH            */
  int32_t n_threads;       /* reference n_threads: 1 = sequential tie rule,
                              anything else = rayon (parallel) tie rule       */
  float goss_top_rate;     /* > 0 enables GOSS (row.rs:211-270)               */
  float goss_other_rate;
  const int32_t* query_group_sizes; /* LambdaRank: ragged groups, sum == n_rows */
  int64_t n_query_groups;
  int32_t autotune;        /* 1 = measure histogram launch shapes on this data
                              and keep the fastest (cached per ctx and data
                              shape); 0 = fixed shape                         */
  int32_t hist_fg;            /* 0 = tuned / default 32; else 16 or 32
                                 features per histogram block                 */
  int32_t hist_rows_min;      /* 0 = tuned / default 1024; else fewest rows per
                                 histogram block, a power of two 512..8192    */
  int32_t hist_blocks_per_cu; /* 0 = tuned / default 2; else 1..8 histogram
                                 blocks per CU that a step's rows spread over */
  int32_t hist_priv;          /* 0 = tuned / default 1; 1 = per-block
                                 sub-histograms in LDS, 2 = in global memory
                                 (device atomics at L2)                       */
} bst_ext_config;

/* ---- forest exchange format -------------------------------------------- */
/* SoA of every tree concatenated, exactly what Tree::new takes
 * (repr/gbdt/tree.rs:65-100) plus gains/covers.  Node ids inside a tree are
 * tree-local; cat_start is an index into cat_words. */
typedef struct bst_model_info {
  int64_t n_trees;
  int64_t n_groups;
  int64_t n_features;
  int64_t total_nodes;
  int64_t total_cat_words;
  int32_t output_transform; /* BST_TRANSFORM_* */
  int32_t objective;
} bst_model_info;

typedef struct bst_forest_arrays {
  float* base_score;      /* [n_groups]        */
  int64_t* tree_offset;   /* [n_trees + 1]     */
  uint32_t* tree_group;   /* [n_trees]         */
  uint32_t* split_index;  /* [total_nodes]     */
  float* split_threshold; /* [total_nodes]     */
  uint32_t* left;         /* [total_nodes]     */
  uint32_t* right;        /* [total_nodes]     */
  uint8_t* default_left;  /* [total_nodes]     */
  uint8_t* is_leaf;       /* [total_nodes]     */
  uint8_t* split_type;    /* [total_nodes] 0 numeric, 1 categorical */
  float* leaf_value;      /* [total_nodes]     */
  float* gain;            /* [total_nodes]     */
  float* cover;           /* [total_nodes]     */
  uint32_t* cat_start;    /* [total_nodes]     */
  uint32_t* cat_size;     /* [total_nodes]     */
  uint32_t* cat_words;    /* [total_cat_words] */
} bst_forest_arrays;

/* ---- opaque handles ---------------------------------------------------- */
typedef struct bst_ctx bst_ctx;
typedef struct bst_dataset bst_dataset;
typedef struct bst_model bst_model;
typedef struct bst_server bst_server;

/* ---- library ----------------------------------------------------------- */
const char* bst_last_error(void);
const char* bst_version(void);
void bst_gbdt_config_default(bst_gbdt_config* cfg);   /* config.rs:111-226 defaults */
void bst_ext_config_default(bst_ext_config* ext);
int bst_gbdt_config_validate(const bst_gbdt_config* cfg); /* GBDTConfig::validate, config.rs:252-310 */

/* ---- context ----------------------------------------------------------- */
int bst_ctx_create(int device, bst_ctx** out);
int bst_ctx_destroy(bst_ctx* ctx);
int bst_ctx_sync(bst_ctx* ctx);
/* The ctx stream as a hipStream_t cast to void*, for callers that enqueue
 * around it (benchmarks time kernels on this stream). */
void* bst_ctx_stream(bst_ctx* ctx);
/* Kernel profiling (no counterpart in the reference): when enabled, every
 * histogram launch of bst_train is bracketed by HIP events on the ctx
 * stream; bst_ctx_kernel_stats returns the launch count, summed device
 * milliseconds and summed algorithmic bytes (rows x (features x bytes per
 * stored bin + 12) + node histograms written) since the last reset.
 * kind 0 = histogram build, 1 = split scan, 2 = partition. */
int bst_ctx_set_profiling(bst_ctx* ctx, int32_t enable);
int bst_ctx_kernel_stats(bst_ctx* ctx, int32_t kind, int64_t* launches, double* ms,
                         double* alg_bytes);
/* A plain streaming copy kernel (dst = src, bytes multiple of 16) on the ctx
 * stream: the reference point that scales HBM byte counters (no counterpart
 * in the reference). */
int bst_kernel_copy(bst_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* Device buffers for callers without their own allocator (plumbing). */
int bst_device_alloc(bst_ctx* ctx, int64_t bytes, void** out);
int bst_device_free(bst_ctx* ctx, void* ptr);
int bst_memcpy_h2d(bst_ctx* ctx, void* dst, const void* src, int64_t bytes);
int bst_memcpy_d2h(bst_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* ---- dataset (quantisation) -------------------------------------------- */
/* BinnedDataset::from_dataset (data/binned/dataset.rs:207-309).
 * feat_major: [n_feat][n_rows] f32, NaN = missing.  is_categorical: [n_feat]
 * (nullable = all numeric).  binning nullable = defaults.  _device variant:
 * feat_major is a device pointer that must stay valid for the call only. */
int bst_dataset_create(bst_ctx* ctx, const float* feat_major, int64_t n_rows, int32_t n_feat,
                       const uint8_t* is_categorical, const bst_binning_config* binning,
                       bst_dataset** out);
int bst_dataset_create_device(bst_ctx* ctx, const float* feat_major_dev, int64_t n_rows,
                              int32_t n_feat, const uint8_t* is_categorical,
                              const bst_binning_config* binning, bst_dataset** out);
int bst_dataset_destroy(bst_dataset* ds);
/* Per original feature: n_bins, missing type (0 none, 2 NaN), default bin,
 * categorical flag, trivial flag; BinMapper accessors, bin_mapper.rs:147-200. */
int bst_dataset_feature_info(const bst_dataset* ds, int32_t feature, uint32_t* n_bins,
                             int32_t* missing_type, uint32_t* default_bin, int32_t* is_categorical,
                             int32_t* is_trivial);
/* Upper bounds (numeric, f64, n_bins values) or sorted categories (as f64). */
int bst_dataset_bin_bounds(const bst_dataset* ds, int32_t feature, double* out, int64_t cap);
/* Bin of every (feature, row) as u16, [n_feat][n_rows]; trivial features
 * report 0.  BinMapper::value_to_bin, bin_mapper.rs:210-258 (bit-exact). */
int bst_dataset_get_bins(const bst_dataset* ds, uint16_t* out);

/* ---- training ---------------------------------------------------------- */
/* GBDTModel::train (model.rs:115-143) over an already-binned dataset.
 * targets: [n_outputs][n_rows] host f32 (softmax: class ids in row 0).
 * weights nullable.  valid/valid_targets nullable: validation set for the
 * metric and early stopping (trainer.rs:341-390).  ext nullable = defaults. */
int bst_train(bst_ctx* ctx, bst_dataset* ds, const float* targets, int64_t n_targets,
              const float* weights, const float* valid_feat_major, const float* valid_targets,
              int64_t n_valid, const bst_gbdt_config* cfg, const bst_ext_config* ext,
              bst_model** out);
/* Same, with targets/weights as device pointers (inputs resident in HBM). */
int bst_train_device(bst_ctx* ctx, bst_dataset* ds, const float* targets_dev, int64_t n_targets,
                     const float* weights_dev, const bst_gbdt_config* cfg,
                     const bst_ext_config* ext, bst_model** out);
/* Timing of the last bst_train on this ctx: binning excluded. */
int bst_last_train_stats(bst_ctx* ctx, double* ms_total, double* ms_per_tree,
                         int64_t* hist_rows_built);
/* Histogram launch shape of the last train on this ctx (fixed or chosen by
 * the autotuner); tuned = 1 if it was measured on this data shape, tune_ms =
 * time the measurement took in that call (0 when cached).  No counterpart in
 * the reference. */
int bst_last_train_tuning(bst_ctx* ctx, int32_t* hist_fg, int32_t* hist_rows_min,
                          int32_t* hist_blocks_per_cu, int32_t* hist_priv, int32_t* tuned,
                          double* tune_ms);
/* Per-round metric values recorded by the last train ([n_rounds]). */
int bst_last_train_metric(bst_ctx* ctx, double* out, int64_t cap, int64_t* n_rounds);
/* Training predictions (raw margins, [n_outputs][n_rows]) after the last
 * round of the last bst_train on this ctx, kept only while enabled: the
 * reference keeps them inside GBDTTrainer::train (trainer.rs:232-339,
 * updated from the partitioner's leaf ranges, grower.rs:208-219) and never
 * returns them.  *count = number of floats; out may be NULL to query it. */
int bst_ctx_set_keep_predictions(bst_ctx* ctx, int32_t enable);
int bst_last_train_predictions(bst_ctx* ctx, float* out, int64_t cap, int64_t* count);

/* ---- model ------------------------------------------------------------- */
int bst_model_info_get(const bst_model* model, bst_model_info* info);
int bst_model_export(const bst_model* model, bst_forest_arrays* out);
int bst_model_import(bst_ctx* ctx, const bst_model_info* info, const bst_forest_arrays* in,
                     bst_model** out);
int bst_model_destroy(bst_model* model);

/* ---- persistence: .bstr files ------------------------------------------ */
/* GBDTModel::write_into (persist/mod.rs:105-111; binary.rs:52-99): the .bstr
 * bytes of the model -- 32-byte header ("BSTR", version 2, format, model type
 * 1), MessagePack of the model schema (persist/schema.rs:27-149), 16-byte
 * trailer (payload length, CRC32C).  compression_level 0 = uncompressed
 * (format byte 1); 1..22 = zstd level (format byte 2, needs libzstd.so.1);
 * the reference's default is 3 (binary.rs:16).  *out is allocated by the
 * library: release it with bst_buffer_free. */
int bst_model_save_bstr(const bst_model* model, int32_t compression_level, uint8_t** out,
                        int64_t* size);
/* GBDTModel::read_from (persist/mod.rs:114-117; binary.rs:101-146;
 * convert.rs:207-470): header, trailer length and CRC32C, optional zstd, the
 * schema, then the forest checks of Tree::validate / Forest::validate.  Any
 * defect returns BST_E_FORMAT.  ctx may be NULL (the model is host-side until
 * its first prediction). */
int bst_model_load_bstr(bst_ctx* ctx, const uint8_t* buf, int64_t size, bst_model** out);
/* The same through a file path. */
int bst_model_save_bstr_file(const bst_model* model, const char* path, int32_t compression_level);
int bst_model_load_bstr_file(bst_ctx* ctx, const char* path, bst_model** out);
/* write_json_into / read_json_from (persist/mod.rs:120-143; json.rs:50-75):
 * {"bstr_version":2,"model_type":"gbdt","model":{schema}}.  pretty != 0
 * indents by two spaces.  *out is NUL-terminated, *size excludes the NUL;
 * release it with bst_buffer_free. */
int bst_model_save_json(const bst_model* model, int32_t pretty, char** out, int64_t* size);
int bst_model_load_json(bst_ctx* ctx, const char* buf, int64_t size, bst_model** out);
/* Frees a buffer returned by bst_model_save_bstr / bst_model_save_json. */
void bst_buffer_free(void* p);

/* ---- prediction -------------------------------------------------------- */
/* GBDTModel::predict / predict_raw (model.rs:217-258).  feat_major host
 * [n_feat][n_rows]; out [n_groups][n_rows] caller-allocated; leaf_idx
 * nullable [n_trees][n_rows] (tree-local node id of the reached leaf). */
int bst_predict(bst_ctx* ctx, const bst_model* model, const float* feat_major, int64_t n_rows,
                int32_t n_feat, int32_t raw, float* out, uint32_t* leaf_idx);
/* Device-pointer variant: feat_major_dev / out_dev / leaf_idx_dev on device. */
int bst_predict_device(bst_ctx* ctx, const bst_model* model, const float* feat_major_dev,
                       int64_t n_rows, int32_t n_feat, int32_t raw, float* out_dev,
                       uint32_t* leaf_idx_dev);

/* Allocation-free variant of bst_predict_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, same results bit for bit (leaf indices too).
 * It calls no hipMalloc / hipFree: the model must already be resident on the
 * ctx's device (any earlier bst_predict / bst_predict_device /
 * bst_server_create with it uploads it; otherwise BST_E_INVALID), and the
 * two-phase launch that small batches take keeps its per-row leaf values in
 * a static 32 MiB device array (8,388,608 floats, one copy per GPU for the
 * life of the process), processing rows in batches of 8388608 / n_trees.
 * Forests with more trees than that are rejected (BST_E_OOM).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (two ctxs of
 * the same device on different threads or streams) race. */
int bst_predict_device_noalloc(bst_ctx* ctx, const bst_model* model, const float* feat_major_dev,
                               int64_t n_rows, int32_t n_feat, int32_t raw, float* out_dev,
                               uint32_t* leaf_idx_dev);

/* ---- serving loop (extension: continuous batching) ---------------------
 * One persistent traversal launch serves every request.  Requests are
 * coalesced into up to max_inflight (at most 16) batches held in pinned host
 * arenas; a batch too small to occupy the device is spread over it through a
 * per-slot leaf-value buffer in HBM.  The arenas and leaf buffers are sized
 * from hbm_budget_bytes (capped at 90% of the device's free HBM) at creation;
 * nothing is allocated afterwards, and a request larger than a batch is
 * refused (BST_E_OOM).  Outputs equal bst_predict's raw margins bit for bit,
 * whether a request ran alone or batched with others. */
int bst_server_create(bst_ctx* ctx, const bst_model* model, int64_t hbm_budget_bytes,
                      int32_t max_inflight, bst_server** out);
/* rows: host [n_feat][n_rows]; out: host [n_groups][n_rows] raw margins. */
int bst_server_submit(bst_server* srv, const float* feat_major, int64_t n_rows, int32_t n_feat,
                      float* out, uint64_t* ticket);
int bst_server_wait(bst_server* srv, uint64_t ticket, int64_t* latency_ns);
int bst_server_destroy(bst_server* srv);

/* ---- component entry points (one hot-path kernel each, for parity) ------ */
/* Objective::compute_gradients_into (objectives/mod.rs:236-276).
 * preds [K][n], targets [n] (class ids for softmax), weights nullable,
 * grad/hess out [K][n].  Host buffers. */
int bst_kernel_gradients(bst_ctx* ctx, int32_t objective, int32_t n_outputs, const float* preds,
                         const float* targets, const float* weights, int64_t n, float* grad,
                         float* hess);
/* The same with the objective's payload from a config (pinball alphas, one
 * output each; pseudo-Huber delta); outputs follow Objective::n_outputs.
 * bst_kernel_gradients runs the payload-free objectives (pseudo-Huber with
 * delta 1.0) and rejects pinball. */
int bst_kernel_gradients_cfg(bst_ctx* ctx, const bst_gbdt_config* cfg, const float* preds,
                             const float* targets, const float* weights, int64_t n, float* grad,
                             float* hess);
/* LambdaRank gradients over ragged query groups (group_sizes [n_groups],
 * each 1..4096 rows, summing to n; targets are integer relevance labels
 * 0..31).  No counterpart in the reference (its objectives stop at
 * objectives/mod.rs:196-219): the published LambdaMART gradients with
 * sigma = 1 and pairs truncated to the top 30 positions.  Host buffers. */
int bst_kernel_lambdarank(bst_ctx* ctx, const float* preds, const float* targets,
                          const float* weights, int64_t n, const int32_t* group_sizes,
                          int64_t n_groups, float* grad, float* hess);
/* TreeGrower::build_histogram (grower.rs:662-730) for one node: rows (nullable
 * = all rows in order) with per-row grad/hess indexed by row id.  Output is the
 * node histogram as f64 (g, h) pairs, [n_effective_features][max_bins_eff]
 * with offsets given by bst_dataset_hist_layout. */
int bst_kernel_histogram(bst_ctx* ctx, const bst_dataset* ds, const float* grad, const float* hess,
                         const uint32_t* rows, int64_t n_rows, double* out_gh, int64_t out_len);
int bst_dataset_hist_layout(const bst_dataset* ds, int32_t* n_effective, uint32_t* eff_to_orig,
                            uint32_t* eff_n_bins, uint32_t* eff_offset);
/* RowPartitioner::split (partition.rs:296-382): stable partition of rows on
 * one effective feature.  split_bin for numeric; left_cats bitset (u32 words,
 * may be null) for categorical.  rows is rewritten in place; returns left count. */
int bst_kernel_partition(bst_ctx* ctx, const bst_dataset* ds, uint32_t* rows, int64_t n_rows,
                         int32_t eff_feature, uint32_t split_bin, int32_t default_left,
                         const uint32_t* left_cats, int32_t n_cat_words, int64_t* left_count);

/* The graph schedule's one-pass partition of one node (a test hook of the
 * stable partition of partition.rs:73-145; no counterpart in the reference's
 * API): the same split rule and the same result as bst_kernel_partition
 * (left rows first, both sides in input order), computed by the training
 * kernel in 4096-row tiles.  mode bits: BST_PART_REVERSED stores the node's
 * rows the way the kernel stores a right child (last position first);
 * BST_PART_TICKETS launches fewer blocks than tiles, so that tiles are
 * claimed through the ticket counter as on steps with more tiles than the
 * grid.  Neither may change the result. */
#define BST_PART_REVERSED 1
#define BST_PART_TICKETS 2
int bst_kernel_partition_fused(bst_ctx* ctx, const bst_dataset* ds, uint32_t* rows,
                               int64_t n_rows, int32_t eff_feature, uint32_t split_bin,
                               int32_t default_left, const uint32_t* left_cats,
                               int32_t n_cat_words, int32_t mode, int64_t* left_count);

/* SplitInfo (training/gbdt/split/types.rs:34-55): feature is an effective
 * column; gain = -inf marks "no valid split". */
typedef struct bst_split_info {
  uint32_t feature;
  float gain;
  int32_t default_left;
  int32_t is_categorical;
  uint32_t bin;          /* numeric: bin <= bin goes left      */
  uint32_t left_cats[8]; /* categorical: bin ids that go left  */
} bst_split_info;

/* GreedySplitter::find_split (split/find.rs:113-270) on one node histogram
 * given as f64 (g, h) pairs in the bst_dataset_hist_layout layout.  The
 * histogram is converted to the engine's 64-bit fixed point (power-of-two
 * scale) before the device scan. */
int bst_kernel_find_split(bst_ctx* ctx, const bst_dataset* ds, const double* hist_gh,
                          double parent_grad, double parent_hess, uint32_t parent_count,
                          const bst_gbdt_config* cfg, int32_t n_threads, bst_split_info* out);

#ifdef __cplusplus
}
#endif
#endif /* BOOSTERS_AMD_H */
