//! Raw bindings to `libboosters_amd`, the MI355X engine for the histogram
//! GBDT hot path.  Every item mirrors `include/boosters_amd.h` one to one
//! (same names, field order and widths); the header is the source of truth.
//!
//! The engine replaces two seams of the `boosters` crate:
//! - `GBDTModel::train_inner` (`model/gbdt/model.rs:149-187`), which calls
//!   `GBDTTrainer::train` (`training/gbdt/trainer.rs:162-401`);
//! - `GBDTModel::predict_raw` (`model/gbdt/model.rs:238-258`), which calls
//!   `UnrolledPredictor6::predict_into` (`inference/gbdt/predictor.rs:197-247`).
//!
//! `integration/mi355x.rs` in this crate is the module that wires them.
//!
//! All functions return `BST_OK` (0) or a negative `BST_E_*` code and leave a
//! message for `bst_last_error` (thread-local in the engine).
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

// ---- status codes --------------------------------------------------------
pub const BST_OK: c_int = 0;
/// Invalid config (reference: `ConfigError`, `config.rs:39-48`).
pub const BST_E_CONFIG: c_int = -1;
/// Too few targets (reference: `trainer.rs:175-177` returns `None`).
pub const BST_E_TARGETS: c_int = -2;
/// Shape mismatch (reference: the `predictor.rs:209-213` assertion).
pub const BST_E_SHAPE: c_int = -3;
/// Device memory or serving admission budget exceeded.
pub const BST_E_OOM: c_int = -4;
/// HIP runtime error or no gfx950 device.
pub const BST_E_DEVICE: c_int = -5;
/// Corrupt or unsupported model bytes (`persist/error.rs`).
pub const BST_E_FORMAT: c_int = -6;
/// Null handle or bad argument.
pub const BST_E_INVALID: c_int = -7;

// ---- enums as ints --------------------------------------------------------
pub const BST_OBJ_SQUARED: i32 = 0;
pub const BST_OBJ_LOGISTIC: i32 = 1;
pub const BST_OBJ_SOFTMAX: i32 = 2;
/// Engine extension with no counterpart in the reference.
pub const BST_OBJ_LAMBDARANK: i32 = 3;
pub const BST_OBJ_ABSOLUTE: i32 = 4;
pub const BST_OBJ_HINGE: i32 = 5;
pub const BST_OBJ_PINBALL: i32 = 6;
pub const BST_OBJ_PSEUDO_HUBER: i32 = 7;
pub const BST_OBJ_POISSON: i32 = 8;
pub const BST_GROW_DEPTHWISE: i32 = 0;
pub const BST_GROW_LEAFWISE: i32 = 1;
pub const BST_METRIC_DEFAULT: i32 = -1;
pub const BST_METRIC_NONE: i32 = 0;
pub const BST_METRIC_RMSE: i32 = 1;
pub const BST_METRIC_LOGLOSS: i32 = 2;
pub const BST_METRIC_MLOGLOSS: i32 = 3;
pub const BST_METRIC_MAE: i32 = 4;
pub const BST_METRIC_MAPE: i32 = 5;
pub const BST_METRIC_QUANTILE: i32 = 6;
pub const BST_METRIC_HUBER: i32 = 7;
pub const BST_METRIC_POISSON_DEVIANCE: i32 = 8;
pub const BST_METRIC_ACCURACY: i32 = 9;
pub const BST_METRIC_MARGIN_ACCURACY: i32 = 10;
pub const BST_METRIC_MULTICLASS_ACCURACY: i32 = 11;
pub const BST_TRANSFORM_IDENTITY: i32 = 0;
pub const BST_TRANSFORM_SIGMOID: i32 = 1;
pub const BST_TRANSFORM_SOFTMAX: i32 = 2;
pub const BST_SCHED_PLAIN: i32 = 0;
pub const BST_SCHED_BATCHED: i32 = 1;
pub const BST_SCHED_GRAPH: i32 = 2;
pub const BST_PART_REVERSED: i32 = 1;
pub const BST_PART_TICKETS: i32 = 2;

// ---- configuration --------------------------------------------------------
/// `BinningConfig` (`data/binned/feature_analysis.rs:29-63`), field for field.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bst_binning_config {
    pub max_bins: u32,
    pub sparsity_threshold: f32,
    pub enable_bundling: i32,
    pub max_categorical_cardinality: u32,
    pub sample_cnt: u64,
}

/// `GBDTConfig` (`model/gbdt/config.rs:111-226`), field for field; enums as
/// ints, `early_stopping_rounds` 0 for `None`; the objective's and metric's
/// enum payloads (Huber delta, quantile alphas, accuracy threshold) at the end.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct bst_gbdt_config {
    pub objective: i32,
    pub n_classes: u32,
    pub metric: i32,
    pub n_trees: u32,
    pub learning_rate: f32,
    pub growth: i32,
    pub max_depth: u32,
    pub max_leaves: u32,
    pub max_onehot_cats: u32,
    pub lambda: f32,
    pub alpha: f32,
    pub min_child_weight: f32,
    pub min_gain: f32,
    pub min_samples_leaf: u32,
    pub subsample: f32,
    pub colsample_bytree: f32,
    pub colsample_bylevel: f32,
    pub binning: bst_binning_config,
    pub early_stopping_rounds: u32,
    pub cache_size: u64,
    pub seed: u64,
    pub verbosity: i32,
    pub huber_delta: f32,
    pub n_alphas: u32,
    pub alphas: *const f32,
    pub accuracy_threshold: f32,
    pub metric_huber_delta: f64,
}

impl Default for bst_gbdt_config {
    fn default() -> Self {
        bst_gbdt_config {
            objective: 0,
            n_classes: 0,
            metric: 0,
            n_trees: 0,
            learning_rate: 0.0,
            growth: 0,
            max_depth: 0,
            max_leaves: 0,
            max_onehot_cats: 0,
            lambda: 0.0,
            alpha: 0.0,
            min_child_weight: 0.0,
            min_gain: 0.0,
            min_samples_leaf: 0,
            subsample: 0.0,
            colsample_bytree: 0.0,
            colsample_bylevel: 0.0,
            binning: bst_binning_config::default(),
            early_stopping_rounds: 0,
            cache_size: 0,
            seed: 0,
            verbosity: 0,
            huber_delta: 0.0,
            n_alphas: 0,
            alphas: std::ptr::null(),
            accuracy_threshold: 0.0,
            metric_huber_delta: 0.0,
        }
    }
}

/// Engine knobs with no counterpart in `GBDTConfig`.  `n_threads` keeps the
/// one meaning the reference's thread count has for results: 1 selects the
/// sequential tie rule, anything else the parallel one
/// (`training/gbdt/split/find.rs:125-129,235-269`).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct bst_ext_config {
    pub schedule: i32,
    pub n_threads: i32,
    pub goss_top_rate: f32,
    pub goss_other_rate: f32,
    pub query_group_sizes: *const i32,
    pub n_query_groups: i64,
    pub autotune: i32,
    pub hist_fg: i32,
    pub hist_rows_min: i32,
    pub hist_blocks_per_cu: i32,
    pub hist_priv: i32,
}

impl Default for bst_ext_config {
    fn default() -> Self {
        let mut e = bst_ext_config {
            schedule: 0,
            n_threads: 0,
            goss_top_rate: 0.0,
            goss_other_rate: 0.0,
            query_group_sizes: std::ptr::null(),
            n_query_groups: 0,
            autotune: 0,
            hist_fg: 0,
            hist_rows_min: 0,
            hist_blocks_per_cu: 0,
            hist_priv: 0,
        };
        unsafe { bst_ext_config_default(&mut e) };
        e
    }
}

impl bst_gbdt_config {
    /// The engine's defaults, which are `GBDTConfig::builder().build()`'s.
    pub fn engine_default() -> Self {
        let mut c = Self::default();
        unsafe { bst_gbdt_config_default(&mut c) };
        c
    }
}

// ---- forest exchange format -------------------------------------------------
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bst_model_info {
    pub n_trees: i64,
    pub n_groups: i64,
    pub n_features: i64,
    pub total_nodes: i64,
    pub total_cat_words: i64,
    pub output_transform: i32,
    pub objective: i32,
}

/// The SoA of every tree concatenated: what `Tree::new` takes
/// (`repr/gbdt/tree.rs:65-100`) plus gains and covers.  Node ids inside a
/// tree are tree-local; `cat_start` indexes `cat_words`.  The caller owns
/// every buffer (sizes from `bst_model_info`).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct bst_forest_arrays {
    pub base_score: *mut f32,
    pub tree_offset: *mut i64,
    pub tree_group: *mut u32,
    pub split_index: *mut u32,
    pub split_threshold: *mut f32,
    pub left: *mut u32,
    pub right: *mut u32,
    pub default_left: *mut u8,
    pub is_leaf: *mut u8,
    pub split_type: *mut u8,
    pub leaf_value: *mut f32,
    pub gain: *mut f32,
    pub cover: *mut f32,
    pub cat_start: *mut u32,
    pub cat_size: *mut u32,
    pub cat_words: *mut u32,
}

/// `SplitInfo` (`training/gbdt/split/types.rs:34-55`) for the per-kernel
/// entry point; `gain = -inf` marks "no valid split".
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bst_split_info {
    pub feature: u32,
    pub gain: f32,
    pub default_left: i32,
    pub is_categorical: i32,
    pub bin: u32,
    pub left_cats: [u32; 8],
}

// ---- opaque handles ---------------------------------------------------------
#[repr(C)]
pub struct bst_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bst_dataset {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bst_model {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bst_server {
    _p: [u8; 0],
}

extern "C" {
    // ---- library ----
    pub fn bst_last_error() -> *const c_char;
    pub fn bst_version() -> *const c_char;
    pub fn bst_gbdt_config_default(cfg: *mut bst_gbdt_config);
    pub fn bst_ext_config_default(ext: *mut bst_ext_config);
    pub fn bst_gbdt_config_validate(cfg: *const bst_gbdt_config) -> c_int;

    // ---- context ----
    pub fn bst_ctx_create(device: c_int, out: *mut *mut bst_ctx) -> c_int;
    pub fn bst_ctx_destroy(ctx: *mut bst_ctx) -> c_int;
    pub fn bst_ctx_sync(ctx: *mut bst_ctx) -> c_int;
    pub fn bst_ctx_stream(ctx: *mut bst_ctx) -> *mut c_void;
    pub fn bst_ctx_set_profiling(ctx: *mut bst_ctx, enable: i32) -> c_int;
    pub fn bst_ctx_kernel_stats(
        ctx: *mut bst_ctx,
        kind: i32,
        launches: *mut i64,
        ms: *mut f64,
        alg_bytes: *mut f64,
    ) -> c_int;
    pub fn bst_kernel_copy(ctx: *mut bst_ctx, dst: *mut c_void, src: *const c_void, bytes: i64)
        -> c_int;
    pub fn bst_device_alloc(ctx: *mut bst_ctx, bytes: i64, out: *mut *mut c_void) -> c_int;
    pub fn bst_device_free(ctx: *mut bst_ctx, ptr: *mut c_void) -> c_int;
    pub fn bst_memcpy_h2d(ctx: *mut bst_ctx, dst: *mut c_void, src: *const c_void, bytes: i64)
        -> c_int;
    pub fn bst_memcpy_d2h(ctx: *mut bst_ctx, dst: *mut c_void, src: *const c_void, bytes: i64)
        -> c_int;

    // ---- dataset (quantisation, BinnedDataset::from_dataset) ----
    pub fn bst_dataset_create(
        ctx: *mut bst_ctx,
        feat_major: *const f32,
        n_rows: i64,
        n_feat: i32,
        is_categorical: *const u8,
        binning: *const bst_binning_config,
        out: *mut *mut bst_dataset,
    ) -> c_int;
    pub fn bst_dataset_create_device(
        ctx: *mut bst_ctx,
        feat_major_dev: *const f32,
        n_rows: i64,
        n_feat: i32,
        is_categorical: *const u8,
        binning: *const bst_binning_config,
        out: *mut *mut bst_dataset,
    ) -> c_int;
    pub fn bst_dataset_destroy(ds: *mut bst_dataset) -> c_int;
    pub fn bst_dataset_feature_info(
        ds: *const bst_dataset,
        feature: i32,
        n_bins: *mut u32,
        missing_type: *mut i32,
        default_bin: *mut u32,
        is_categorical: *mut i32,
        is_trivial: *mut i32,
    ) -> c_int;
    pub fn bst_dataset_bin_bounds(ds: *const bst_dataset, feature: i32, out: *mut f64, cap: i64)
        -> c_int;
    pub fn bst_dataset_get_bins(ds: *const bst_dataset, out: *mut u16) -> c_int;

    // ---- training ----
    pub fn bst_train(
        ctx: *mut bst_ctx,
        ds: *mut bst_dataset,
        targets: *const f32,
        n_targets: i64,
        weights: *const f32,
        valid_feat_major: *const f32,
        valid_targets: *const f32,
        n_valid: i64,
        cfg: *const bst_gbdt_config,
        ext: *const bst_ext_config,
        out: *mut *mut bst_model,
    ) -> c_int;
    pub fn bst_train_device(
        ctx: *mut bst_ctx,
        ds: *mut bst_dataset,
        targets_dev: *const f32,
        n_targets: i64,
        weights_dev: *const f32,
        cfg: *const bst_gbdt_config,
        ext: *const bst_ext_config,
        out: *mut *mut bst_model,
    ) -> c_int;
    pub fn bst_last_train_stats(
        ctx: *mut bst_ctx,
        ms_total: *mut f64,
        ms_per_tree: *mut f64,
        hist_rows_built: *mut i64,
    ) -> c_int;
    pub fn bst_last_train_tuning(
        ctx: *mut bst_ctx,
        hist_fg: *mut i32,
        hist_rows_min: *mut i32,
        hist_blocks_per_cu: *mut i32,
        hist_priv: *mut i32,
        tuned: *mut i32,
        tune_ms: *mut f64,
    ) -> c_int;
    pub fn bst_last_train_metric(ctx: *mut bst_ctx, out: *mut f64, cap: i64, n_rounds: *mut i64)
        -> c_int;
    pub fn bst_ctx_set_keep_predictions(ctx: *mut bst_ctx, enable: i32) -> c_int;
    pub fn bst_last_train_predictions(
        ctx: *mut bst_ctx,
        out: *mut f32,
        cap: i64,
        count: *mut i64,
    ) -> c_int;

    // ---- model ----
    pub fn bst_model_info_get(model: *const bst_model, info: *mut bst_model_info) -> c_int;
    pub fn bst_model_export(model: *const bst_model, out: *mut bst_forest_arrays) -> c_int;
    pub fn bst_model_import(
        ctx: *mut bst_ctx,
        info: *const bst_model_info,
        arrays: *const bst_forest_arrays,
        out: *mut *mut bst_model,
    ) -> c_int;
    pub fn bst_model_destroy(model: *mut bst_model) -> c_int;

    // ---- persistence (.bstr / JSON, persist/mod.rs:105-143) ----
    pub fn bst_model_save_bstr(
        model: *const bst_model,
        compression_level: i32,
        out: *mut *mut u8,
        size: *mut i64,
    ) -> c_int;
    pub fn bst_model_load_bstr(
        ctx: *mut bst_ctx,
        buf: *const u8,
        size: i64,
        out: *mut *mut bst_model,
    ) -> c_int;
    pub fn bst_model_save_bstr_file(
        model: *const bst_model,
        path: *const c_char,
        compression_level: i32,
    ) -> c_int;
    pub fn bst_model_load_bstr_file(
        ctx: *mut bst_ctx,
        path: *const c_char,
        out: *mut *mut bst_model,
    ) -> c_int;
    pub fn bst_model_save_json(
        model: *const bst_model,
        pretty: i32,
        out: *mut *mut c_char,
        size: *mut i64,
    ) -> c_int;
    pub fn bst_model_load_json(
        ctx: *mut bst_ctx,
        buf: *const c_char,
        size: i64,
        out: *mut *mut bst_model,
    ) -> c_int;
    pub fn bst_buffer_free(p: *mut c_void);

    // ---- prediction (predict / predict_raw, model.rs:217-258) ----
    pub fn bst_predict(
        ctx: *mut bst_ctx,
        model: *const bst_model,
        feat_major: *const f32,
        n_rows: i64,
        n_feat: i32,
        raw: i32,
        out: *mut f32,
        leaf_idx: *mut u32,
    ) -> c_int;
    pub fn bst_predict_device(
        ctx: *mut bst_ctx,
        model: *const bst_model,
        feat_major_dev: *const f32,
        n_rows: i64,
        n_feat: i32,
        raw: i32,
        out_dev: *mut f32,
        leaf_idx_dev: *mut u32,
    ) -> c_int;
    /// No allocation of any kind: for callers that cannot allocate device
    /// memory.  Shares one static 32 MiB device array per GPU, so calls that
    /// overlap on one GPU race.
    pub fn bst_predict_device_noalloc(
        ctx: *mut bst_ctx,
        model: *const bst_model,
        feat_major_dev: *const f32,
        n_rows: i64,
        n_feat: i32,
        raw: i32,
        out_dev: *mut f32,
        leaf_idx_dev: *mut u32,
    ) -> c_int;

    // ---- serving loop (continuous batching, one persistent launch) ----
    pub fn bst_server_create(
        ctx: *mut bst_ctx,
        model: *const bst_model,
        hbm_budget_bytes: i64,
        max_inflight: i32,
        out: *mut *mut bst_server,
    ) -> c_int;
    pub fn bst_server_submit(
        srv: *mut bst_server,
        feat_major: *const f32,
        n_rows: i64,
        n_feat: i32,
        out: *mut f32,
        ticket: *mut u64,
    ) -> c_int;
    pub fn bst_server_wait(srv: *mut bst_server, ticket: u64, latency_ns: *mut i64) -> c_int;
    pub fn bst_server_destroy(srv: *mut bst_server) -> c_int;

    // ---- one hot-path kernel each (parity entry points) ----
    pub fn bst_kernel_gradients(
        ctx: *mut bst_ctx,
        objective: i32,
        n_outputs: i32,
        preds: *const f32,
        targets: *const f32,
        weights: *const f32,
        n: i64,
        grad: *mut f32,
        hess: *mut f32,
    ) -> c_int;
    pub fn bst_kernel_gradients_cfg(
        ctx: *mut bst_ctx,
        cfg: *const bst_gbdt_config,
        preds: *const f32,
        targets: *const f32,
        weights: *const f32,
        n: i64,
        grad: *mut f32,
        hess: *mut f32,
    ) -> c_int;
    pub fn bst_kernel_lambdarank(
        ctx: *mut bst_ctx,
        preds: *const f32,
        targets: *const f32,
        weights: *const f32,
        n: i64,
        group_sizes: *const i32,
        n_groups: i64,
        grad: *mut f32,
        hess: *mut f32,
    ) -> c_int;
    pub fn bst_kernel_histogram(
        ctx: *mut bst_ctx,
        ds: *const bst_dataset,
        grad: *const f32,
        hess: *const f32,
        rows: *const u32,
        n_rows: i64,
        out_gh: *mut f64,
        out_len: i64,
    ) -> c_int;
    pub fn bst_dataset_hist_layout(
        ds: *const bst_dataset,
        n_effective: *mut i32,
        eff_to_orig: *mut u32,
        eff_n_bins: *mut u32,
        eff_offset: *mut u32,
    ) -> c_int;
    pub fn bst_kernel_partition(
        ctx: *mut bst_ctx,
        ds: *const bst_dataset,
        rows: *mut u32,
        n_rows: i64,
        eff_feature: i32,
        split_bin: u32,
        default_left: i32,
        left_cats: *const u32,
        n_cat_words: i32,
        left_count: *mut i64,
    ) -> c_int;
    pub fn bst_kernel_partition_fused(
        ctx: *mut bst_ctx,
        ds: *const bst_dataset,
        rows: *mut u32,
        n_rows: i64,
        eff_feature: i32,
        split_bin: u32,
        default_left: i32,
        left_cats: *const u32,
        n_cat_words: i32,
        mode: i32,
        left_count: *mut i64,
    ) -> c_int;
    pub fn bst_kernel_find_split(
        ctx: *mut bst_ctx,
        ds: *const bst_dataset,
        hist_gh: *const f64,
        parent_grad: f64,
        parent_hess: f64,
        parent_count: u32,
        cfg: *const bst_gbdt_config,
        n_threads: i32,
        out: *mut bst_split_info,
    ) -> c_int;
}

/// An engine error: the status code and `bst_last_error`'s message.
#[derive(Debug, Clone)]
pub struct Error {
    pub code: c_int,
    pub message: String,
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "boosters_amd error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

/// Turns a status code into `Result`, reading the engine's message on error.
pub fn check(code: c_int) -> Result<(), Error> {
    if code == BST_OK {
        return Ok(());
    }
    let msg = unsafe {
        let p = bst_last_error();
        if p.is_null() {
            String::new()
        } else {
            std::ffi::CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    };
    Err(Error { code, message: msg })
}
