//! MI355X engine seam for the `boosters` crate.
//!
//! Drop this file in as `crates/boosters/src/model/gbdt/mi355x.rs`, declare it
//! in `model/gbdt/mod.rs` as `#[cfg(feature = "mi355x")] pub(crate) mod mi355x;`,
//! add the feature to `crates/boosters/Cargo.toml`:
//!
//! ```toml
//! [features]
//! mi355x = ["dep:boosters-mi355x-sys"]
//! [dependencies]
//! boosters-mi355x-sys = { path = "<engine checkout>/crates/boosters-mi355x-sys", optional = true }
//! ```
//!
//! and change the two call sites in `model/gbdt/model.rs`:
//!
//! ```ignore
//! // train_inner (model.rs:149-187), instead of `trainer.train(...)`:
//! #[cfg(feature = "mi355x")]
//! let forest = super::mi355x::train(dataset, targets, weights, val_set, &config, parallelism)
//!     .unwrap_or_else(|e| panic!("{e}"))?;
//!
//! // predict_raw (model.rs:238-258), instead of `UnrolledPredictor6::predict_into(...)`:
//! #[cfg(feature = "mi355x")]
//! {
//!     let engine = self.engine.get_or_init(|| {
//!         super::mi355x::EngineModel::from_forest(&self.forest, self.meta.n_features)
//!             .unwrap_or_else(|e| panic!("{e}"))
//!     });
//!     super::mi355x::predict_into(engine, data, output_array.view_mut())
//!         .unwrap_or_else(|e| panic!("{e}"));
//! }
//! ```
//!
//! with one field added to `GBDTModel` (`model.rs:40-55`), so that a model
//! is uploaded once however it was made (trained, read by `persist/`):
//!
//! ```ignore
//! #[cfg(feature = "mi355x")]
//! engine: std::sync::OnceLock<super::mi355x::EngineModel>,
//! ```
//!
//! `train_inner` keeps building `ModelMeta` and the output transform from the
//! returned forest exactly as it does for `GBDTTrainer::train`'s.  The engine
//! bins the raw feature columns itself (`bst_dataset_create` is
//! `BinnedDataset::from_dataset`, bit-exact), so the `BinnedDataset` the CPU
//! path builds beforehand is not needed on this path.
//!
//! Unsupported configurations (custom objectives and metrics, AUC, linear
//! leaves) return `Error::Unsupported` rather than silently running on the
//! CPU: the feature flag chose the engine.

use std::cell::RefCell;

use boosters_mi355x_sys as sys;
use ndarray::ArrayViewMut2;

use crate::data::{Dataset, Feature, TargetsView, WeightsView};
use crate::repr::gbdt::{
    CategoriesStorage, Forest, LeafCoefficients, ScalarLeaf, SplitType, Tree, TreeView,
};
use crate::training::gbdt::GrowthStrategy;
use crate::training::{Metric, Objective, Verbosity};
use crate::utils::Parallelism;

use super::GBDTConfig;

/// Why the engine path could not run.
#[derive(Debug)]
pub enum Error {
    /// A configuration the engine does not implement.
    Unsupported(String),
    /// The engine returned an error code.
    Engine(sys::Error),
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        match self {
            Error::Unsupported(s) => write!(f, "mi355x engine: unsupported configuration: {s}"),
            Error::Engine(e) => write!(f, "mi355x engine: {e}"),
        }
    }
}

impl From<sys::Error> for Error {
    fn from(e: sys::Error) -> Self {
        Error::Engine(e)
    }
}

/// One engine context per thread: a `bst_ctx` owns one HIP stream and is not
/// re-entrant, and a thread-local context gives each caller its own.
struct Ctx(*mut sys::bst_ctx);

impl Drop for Ctx {
    fn drop(&mut self) {
        unsafe { sys::bst_ctx_destroy(self.0) };
    }
}

thread_local! {
    static CTX: RefCell<Option<Ctx>> = const { RefCell::new(None) };
}

fn with_ctx<T>(f: impl FnOnce(*mut sys::bst_ctx) -> Result<T, Error>) -> Result<T, Error> {
    CTX.with(|c| {
        let mut c = c.borrow_mut();
        if c.is_none() {
            let mut h = std::ptr::null_mut();
            sys::check(unsafe { sys::bst_ctx_create(0, &mut h) })?;
            *c = Some(Ctx(h));
        }
        f(c.as_ref().unwrap().0)
    })
}

struct Model(*mut sys::bst_model);

impl Drop for Model {
    fn drop(&mut self) {
        unsafe { sys::bst_model_destroy(self.0) };
    }
}

/// A forest held by the engine.  It is immutable once made, and the engine
/// keeps one device copy of it per device, made on the first prediction
/// there; predictions on distinct contexts may run concurrently.
pub struct EngineModel(Model);

unsafe impl Send for EngineModel {}
unsafe impl Sync for EngineModel {}

impl EngineModel {
    /// Uploads a `Forest<ScalarLeaf>` into the engine (`bst_model_import`).
    pub fn from_forest(forest: &Forest<ScalarLeaf>, n_features: usize) -> Result<Self, Error> {
        with_ctx(|ctx| import_forest(ctx, forest, n_features)).map(EngineModel)
    }
}

struct DatasetHandle(*mut sys::bst_dataset);

impl Drop for DatasetHandle {
    fn drop(&mut self) {
        unsafe { sys::bst_dataset_destroy(self.0) };
    }
}

/// The dataset's features as one feature-major `[n_features][n_samples]`
/// buffer (`data/raw/dataset.rs:43-52`); sparse columns are expanded with
/// their default value.
fn feature_major(ds: &Dataset) -> Vec<f32> {
    let n = ds.n_samples();
    let mut out = Vec::with_capacity(ds.n_features() * n);
    for col in ds.feature_columns() {
        match col {
            Feature::Dense(values) => out.extend(values.iter().copied()),
            Feature::Sparse { indices, values, n_samples, default } => {
                let base = out.len();
                out.resize(base + *n_samples, *default);
                for (&i, &v) in indices.iter().zip(values.iter()) {
                    out[base + i as usize] = v;
                }
            }
        }
    }
    out
}

fn categorical_mask(ds: &Dataset) -> Vec<u8> {
    (0..ds.n_features())
        .map(|f| ds.schema().feature_type(f).is_categorical() as u8)
        .collect()
}

/// `GBDTConfig` -> `bst_gbdt_config`, field for field (`config.rs:111-226`).
/// The quantile alphas are borrowed from `cfg` for the call (`alphas` must
/// outlive the returned struct).
fn engine_config<'a>(cfg: &'a GBDTConfig) -> Result<sys::bst_gbdt_config, Error> {
    let mut c = sys::bst_gbdt_config::engine_default();
    match &cfg.objective {
        Objective::SquaredLoss => c.objective = sys::BST_OBJ_SQUARED,
        Objective::AbsoluteLoss => c.objective = sys::BST_OBJ_ABSOLUTE,
        Objective::LogisticLoss => c.objective = sys::BST_OBJ_LOGISTIC,
        Objective::HingeLoss => c.objective = sys::BST_OBJ_HINGE,
        Objective::SoftmaxLoss { n_classes } => {
            c.objective = sys::BST_OBJ_SOFTMAX;
            c.n_classes = *n_classes as u32;
        }
        Objective::PinballLoss { alphas } => {
            c.objective = sys::BST_OBJ_PINBALL;
            c.n_alphas = alphas.len() as u32;
            c.alphas = alphas.as_ptr();
        }
        Objective::PseudoHuberLoss { delta } => {
            c.objective = sys::BST_OBJ_PSEUDO_HUBER;
            c.huber_delta = *delta;
        }
        Objective::PoissonLoss => c.objective = sys::BST_OBJ_POISSON,
        other => return Err(Error::Unsupported(format!("objective {other:?}"))),
    }
    c.metric = match &cfg.metric {
        Metric::None => sys::BST_METRIC_NONE,
        Metric::Rmse => sys::BST_METRIC_RMSE,
        Metric::Mae => sys::BST_METRIC_MAE,
        Metric::Mape => sys::BST_METRIC_MAPE,
        Metric::LogLoss => sys::BST_METRIC_LOGLOSS,
        Metric::Accuracy { threshold } => {
            c.accuracy_threshold = *threshold;
            sys::BST_METRIC_ACCURACY
        }
        Metric::MarginAccuracy => sys::BST_METRIC_MARGIN_ACCURACY,
        Metric::MulticlassLogLoss => sys::BST_METRIC_MLOGLOSS,
        Metric::MulticlassAccuracy => sys::BST_METRIC_MULTICLASS_ACCURACY,
        Metric::Quantile { alphas } => {
            if c.n_alphas == 0 {
                c.n_alphas = alphas.len() as u32;
                c.alphas = alphas.as_ptr();
            }
            sys::BST_METRIC_QUANTILE
        }
        Metric::Huber { delta } => {
            c.metric_huber_delta = *delta;
            sys::BST_METRIC_HUBER
        }
        Metric::PoissonDeviance => sys::BST_METRIC_POISSON_DEVIANCE,
        other => return Err(Error::Unsupported(format!("metric {other:?}"))),
    };
    if cfg.linear_leaves.is_some() {
        return Err(Error::Unsupported("linear leaves".into()));
    }
    c.n_trees = cfg.n_trees;
    c.learning_rate = cfg.learning_rate;
    match cfg.growth_strategy {
        GrowthStrategy::DepthWise { max_depth } => {
            c.growth = sys::BST_GROW_DEPTHWISE;
            c.max_depth = max_depth;
        }
        GrowthStrategy::LeafWise { max_leaves } => {
            c.growth = sys::BST_GROW_LEAFWISE;
            c.max_leaves = max_leaves;
        }
    }
    c.max_onehot_cats = cfg.max_onehot_cats;
    c.lambda = cfg.lambda;
    c.alpha = cfg.alpha;
    c.min_child_weight = cfg.min_child_weight;
    c.min_gain = cfg.min_gain;
    c.min_samples_leaf = cfg.min_samples_leaf;
    c.subsample = cfg.subsample;
    c.colsample_bytree = cfg.colsample_bytree;
    c.colsample_bylevel = cfg.colsample_bylevel;
    c.binning = sys::bst_binning_config {
        max_bins: cfg.binning.max_bins,
        sparsity_threshold: cfg.binning.sparsity_threshold,
        enable_bundling: cfg.binning.enable_bundling as i32,
        max_categorical_cardinality: cfg.binning.max_categorical_cardinality,
        sample_cnt: cfg.binning.sample_cnt as u64,
    };
    c.early_stopping_rounds = cfg.early_stopping_rounds.unwrap_or(0);
    c.cache_size = cfg.cache_size as u64;
    c.seed = cfg.seed;
    c.verbosity = match cfg.verbosity {
        Verbosity::Silent => 0,
        Verbosity::Warning => 1,
        Verbosity::Info => 2,
        Verbosity::Debug => 3,
    };
    Ok(c)
}

/// `GBDTTrainer::train` on the engine (`training/gbdt/trainer.rs:162-401`).
/// Returns `Ok(None)` where the reference returns `None` (too few targets,
/// `trainer.rs:175-177`).
pub fn train(
    dataset: &Dataset,
    targets: TargetsView<'_>,
    weights: WeightsView<'_>,
    val_set: Option<&Dataset>,
    config: &GBDTConfig,
    parallelism: Parallelism,
) -> Result<Option<Forest<ScalarLeaf>>, Error> {
    let cfg = engine_config(config)?;
    let mut ext = sys::bst_ext_config::default();
    // The thread count only changes results through the split tie rule.
    ext.n_threads = match parallelism {
        Parallelism::Sequential => 1,
        Parallelism::Parallel => 0,
    };

    let n = dataset.n_samples();
    let x = feature_major(dataset);
    let is_cat = categorical_mask(dataset);
    // Targets as [n_outputs][n_samples], contiguous.
    let y: Vec<f32> = targets.view().iter().copied().collect();
    let w: Option<Vec<f32>> = weights.as_array().map(|a| a.iter().copied().collect());
    let (vx, vy) = match val_set {
        Some(v) => {
            let vt = v
                .targets()
                .ok_or_else(|| Error::Unsupported("validation set without targets".into()))?;
            (Some(feature_major(v)), Some(vt.view().iter().copied().collect::<Vec<f32>>()))
        }
        None => (None, None),
    };
    let n_valid = val_set.map_or(0, |v| v.n_samples()) as i64;

    with_ctx(|ctx| {
        let mut ds = std::ptr::null_mut();
        sys::check(unsafe {
            sys::bst_dataset_create(
                ctx,
                x.as_ptr(),
                n as i64,
                dataset.n_features() as i32,
                is_cat.as_ptr(),
                &cfg.binning,
                &mut ds,
            )
        })?;
        let ds = DatasetHandle(ds);
        let mut model = std::ptr::null_mut();
        let rc = unsafe {
            sys::bst_train(
                ctx,
                ds.0,
                y.as_ptr(),
                y.len() as i64,
                w.as_ref().map_or(std::ptr::null(), |w| w.as_ptr()),
                vx.as_ref().map_or(std::ptr::null(), |v| v.as_ptr()),
                vy.as_ref().map_or(std::ptr::null(), |v| v.as_ptr()),
                n_valid,
                &cfg,
                &ext,
                &mut model,
            )
        };
        if rc == sys::BST_E_TARGETS {
            return Ok(None);
        }
        sys::check(rc)?;
        let model = Model(model);
        Ok(Some(export_forest(&model)?))
    })
}

/// The engine's forest as the `Forest<ScalarLeaf>` `GBDTTrainer::train`
/// returns: one `Tree::new` per tree (`repr/gbdt/tree.rs:65-100`) with gains
/// and covers, pushed with its output group (`repr/gbdt/forest.rs:62`).
fn export_forest(model: &Model) -> Result<Forest<ScalarLeaf>, Error> {
    let mut info = sys::bst_model_info::default();
    sys::check(unsafe { sys::bst_model_info_get(model.0, &mut info) })?;
    let (t, g, nn, w) = (
        info.n_trees as usize,
        info.n_groups as usize,
        info.total_nodes as usize,
        info.total_cat_words as usize,
    );
    let mut base_score = vec![0f32; g];
    let mut tree_offset = vec![0i64; t + 1];
    let mut tree_group = vec![0u32; t];
    let mut split_index = vec![0u32; nn];
    let mut split_threshold = vec![0f32; nn];
    let mut left = vec![0u32; nn];
    let mut right = vec![0u32; nn];
    let mut default_left = vec![0u8; nn];
    let mut is_leaf = vec![0u8; nn];
    let mut split_type = vec![0u8; nn];
    let mut leaf_value = vec![0f32; nn];
    let mut gain = vec![0f32; nn];
    let mut cover = vec![0f32; nn];
    let mut cat_start = vec![0u32; nn];
    let mut cat_size = vec![0u32; nn];
    let mut cat_words = vec![0u32; w.max(1)];
    let mut arrays = sys::bst_forest_arrays {
        base_score: base_score.as_mut_ptr(),
        tree_offset: tree_offset.as_mut_ptr(),
        tree_group: tree_group.as_mut_ptr(),
        split_index: split_index.as_mut_ptr(),
        split_threshold: split_threshold.as_mut_ptr(),
        left: left.as_mut_ptr(),
        right: right.as_mut_ptr(),
        default_left: default_left.as_mut_ptr(),
        is_leaf: is_leaf.as_mut_ptr(),
        split_type: split_type.as_mut_ptr(),
        leaf_value: leaf_value.as_mut_ptr(),
        gain: gain.as_mut_ptr(),
        cover: cover.as_mut_ptr(),
        cat_start: cat_start.as_mut_ptr(),
        cat_size: cat_size.as_mut_ptr(),
        cat_words: cat_words.as_mut_ptr(),
    };
    sys::check(unsafe { sys::bst_model_export(model.0, &mut arrays) })?;

    let mut forest = Forest::new(g as u32).with_base_score(base_score);
    for k in 0..t {
        let (a, b) = (tree_offset[k] as usize, tree_offset[k + 1] as usize);
        // cat_start indexes the forest-wide word array: rebase per tree.
        let mut words = Vec::new();
        let mut segments = Vec::with_capacity(b - a);
        for i in a..b {
            let (s, z) = (cat_start[i] as usize, cat_size[i] as usize);
            segments.push((words.len() as u32, z as u32));
            words.extend_from_slice(&cat_words[s..s + z]);
        }
        let any_cat = split_type[a..b].iter().any(|&s| s != 0);
        let categories = if any_cat {
            CategoriesStorage::new(words, segments)
        } else {
            CategoriesStorage::empty()
        };
        let tree = Tree::new(
            split_index[a..b].to_vec(),
            split_threshold[a..b].to_vec(),
            left[a..b].to_vec(),
            right[a..b].to_vec(),
            default_left[a..b].iter().map(|&v| v != 0).collect(),
            is_leaf[a..b].iter().map(|&v| v != 0).collect(),
            leaf_value[a..b].iter().map(|&v| ScalarLeaf(v)).collect(),
            split_type[a..b]
                .iter()
                .map(|&v| if v != 0 { SplitType::Categorical } else { SplitType::Numeric })
                .collect(),
            categories,
            LeafCoefficients::empty(),
        )
        .with_stats(gain[a..b].to_vec(), cover[a..b].to_vec());
        forest.push_tree(tree, tree_group[k]);
    }
    Ok(forest)
}

/// Uploads a `Forest<ScalarLeaf>` (trained anywhere, or read by `persist/`)
/// into the engine through `bst_model_import`.
fn import_forest(ctx: *mut sys::bst_ctx, forest: &Forest<ScalarLeaf>, n_features: usize)
    -> Result<Model, Error> {
    let t = forest.n_trees();
    let g = forest.n_groups() as usize;
    let mut base_score = forest.base_score().to_vec();
    let mut tree_offset = Vec::with_capacity(t + 1);
    let mut tree_group = forest.tree_groups().to_vec();
    let (mut split_index, mut split_threshold, mut left, mut right) =
        (Vec::new(), Vec::new(), Vec::new(), Vec::new());
    let (mut default_left, mut is_leaf, mut split_type) = (Vec::new(), Vec::new(), Vec::new());
    let (mut leaf_value, mut gain, mut cover) = (Vec::new(), Vec::new(), Vec::new());
    let (mut cat_start, mut cat_size, mut cat_words) = (Vec::new(), Vec::new(), Vec::new());
    tree_offset.push(0i64);
    for tree in forest.trees() {
        if tree.has_linear_leaves() {
            return Err(Error::Unsupported("linear leaves".into()));
        }
        let n = tree.n_nodes();
        let word_base = cat_words.len() as u32;
        cat_words.extend_from_slice(tree.categories().bitsets());
        let segs = tree.categories().segments();
        for node in 0..n as u32 {
            split_index.push(tree.split_index(node));
            split_threshold.push(tree.split_threshold(node));
            left.push(tree.left_child(node));
            right.push(tree.right_child(node));
            default_left.push(tree.default_left(node) as u8);
            is_leaf.push(tree.is_leaf(node) as u8);
            split_type.push(matches!(tree.split_type(node), SplitType::Categorical) as u8);
            leaf_value.push(tree.leaf_value(node).0);
            gain.push(tree.gains().map_or(0.0, |v| v[node as usize]));
            cover.push(tree.covers().map_or(0.0, |v| v[node as usize]));
            let (s, z) = segs.get(node as usize).copied().unwrap_or((0, 0));
            cat_start.push(word_base + s);
            cat_size.push(z);
        }
        tree_offset.push(*tree_offset.last().unwrap() + n as i64);
    }
    let info = sys::bst_model_info {
        n_trees: t as i64,
        n_groups: g as i64,
        n_features: n_features as i64,
        total_nodes: split_index.len() as i64,
        total_cat_words: cat_words.len() as i64,
        output_transform: sys::BST_TRANSFORM_IDENTITY,
        objective: sys::BST_OBJ_SQUARED,
    };
    if cat_words.is_empty() {
        cat_words.push(0);
    }
    let arrays = sys::bst_forest_arrays {
        base_score: base_score.as_mut_ptr(),
        tree_offset: tree_offset.as_mut_ptr(),
        tree_group: tree_group.as_mut_ptr(),
        split_index: split_index.as_mut_ptr(),
        split_threshold: split_threshold.as_mut_ptr(),
        left: left.as_mut_ptr(),
        right: right.as_mut_ptr(),
        default_left: default_left.as_mut_ptr(),
        is_leaf: is_leaf.as_mut_ptr(),
        split_type: split_type.as_mut_ptr(),
        leaf_value: leaf_value.as_mut_ptr(),
        gain: gain.as_mut_ptr(),
        cover: cover.as_mut_ptr(),
        cat_start: cat_start.as_mut_ptr(),
        cat_size: cat_size.as_mut_ptr(),
        cat_words: cat_words.as_mut_ptr(),
    };
    let mut h = std::ptr::null_mut();
    sys::check(unsafe { sys::bst_model_import(ctx, &info, &arrays, &mut h) })?;
    Ok(Model(h))
}

/// `UnrolledPredictor6::predict_into` on the engine
/// (`inference/gbdt/predictor.rs:197-247`): raw margins into `output`
/// `[n_groups][n_samples]`, base scores included.  The model's output
/// transform stays with the caller (`GBDTModel::predict`, `model.rs:217-224`).
pub fn predict_into(
    model: &EngineModel,
    data: &Dataset,
    mut output: ArrayViewMut2<'_, f32>,
) -> Result<(), Error> {
    let n = data.n_samples();
    let mut info = sys::bst_model_info::default();
    sys::check(unsafe { sys::bst_model_info_get(model.0 .0, &mut info) })?;
    assert_eq!(
        output.shape(),
        &[info.n_groups as usize, n],
        "output shape must match (n_groups, n_samples)"
    );
    if n == 0 {
        return Ok(());
    }
    let x = feature_major(data);
    let out = output
        .as_slice_mut()
        .ok_or_else(|| Error::Unsupported("non-contiguous output".into()))?;
    with_ctx(|ctx| {
        sys::check(unsafe {
            sys::bst_predict(
                ctx,
                model.0 .0,
                x.as_ptr(),
                n as i64,
                data.n_features() as i32,
                1,
                out.as_mut_ptr(),
                std::ptr::null_mut(),
            )
        })?;
        Ok(())
    })
}
