"""Python host API over the C ABI (include/boosters_amd.h).

Names, argument meanings and defaults follow the reference Python bindings:
  GBDTConfig      packages/boosters-python/python/boosters/_boosters_rs.pyi:234-459
  GBDTModel       _boosters_rs.pyi:462-606 (train :497, predict :543, predict_raw :557)
  Dataset         _boosters_rs.pyi:10-80 (features are [n_samples, n_features])
and the Rust core they wrap (crates/boosters/src/model/gbdt/{config.rs,model.rs}).
Predictions are returned as (n_samples, n_outputs), the bindings' convention.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _abi as A


# ---------------------------------------------------------------------------
# Enums mirroring the reference
# ---------------------------------------------------------------------------
class GrowthStrategy:
    Depthwise = "depthwise"
    Leafwise = "leafwise"


class Schedule:
    """Engine schedule (extension): plain per-node launches, batched launches, or
    batched launches with device-side growth replayed as a hipGraph (default)."""

    Plain = A.SCHED_PLAIN
    Batched = A.SCHED_BATCHED
    Graph = A.SCHED_GRAPH


@dataclass(frozen=True)
class Objective:
    """Objective enum of the bindings (_boosters_rs.pyi Objective) with its
    payload: n_classes (softmax), delta (pseudo-Huber), alpha (pinball
    quantiles, one output each)."""

    kind: int
    n_classes: int = 0
    delta: float = 1.0
    alpha: Tuple[float, ...] = ()

    @staticmethod
    def squared() -> "Objective":
        return Objective(A.OBJ_SQUARED)

    @staticmethod
    def absolute() -> "Objective":
        return Objective(A.OBJ_ABSOLUTE)

    @staticmethod
    def poisson() -> "Objective":
        return Objective(A.OBJ_POISSON)

    @staticmethod
    def logistic() -> "Objective":
        return Objective(A.OBJ_LOGISTIC)

    @staticmethod
    def hinge() -> "Objective":
        return Objective(A.OBJ_HINGE)

    @staticmethod
    def huber(delta: float = 1.0) -> "Objective":
        if not delta > 0.0:
            raise ValueError("huber delta must be positive")
        return Objective(A.OBJ_PSEUDO_HUBER, delta=float(delta))

    @staticmethod
    def pinball(alpha: Sequence[float]) -> "Objective":
        a = tuple(float(v) for v in alpha)
        if not a or not all(0.0 < v < 1.0 for v in a):
            raise ValueError("pinball needs at least one alpha, each in (0, 1)")
        return Objective(A.OBJ_PINBALL, alpha=a)

    @staticmethod
    def softmax(n_classes: int) -> "Objective":
        if n_classes < 2:
            raise ValueError("softmax needs n_classes >= 2")
        return Objective(A.OBJ_SOFTMAX, int(n_classes))

    @property
    def n_outputs(self) -> int:
        if self.kind == A.OBJ_SOFTMAX:
            return self.n_classes
        if self.kind == A.OBJ_PINBALL:
            return len(self.alpha)
        return 1


Objective.Squared = Objective(A.OBJ_SQUARED)  # type: ignore[attr-defined]
Objective.Logistic = Objective(A.OBJ_LOGISTIC)  # type: ignore[attr-defined]
Objective.Absolute = Objective(A.OBJ_ABSOLUTE)  # type: ignore[attr-defined]
Objective.Poisson = Objective(A.OBJ_POISSON)  # type: ignore[attr-defined]
Objective.Hinge = Objective(A.OBJ_HINGE)  # type: ignore[attr-defined]


@dataclass(frozen=True)
class Metric:
    """Metric enum of the bindings (_boosters_rs.pyi Metric) with its payload:
    threshold (accuracy), alpha (quantile)."""

    kind: int
    threshold: float = 0.5
    alpha: Tuple[float, ...] = ()

    @staticmethod
    def none() -> "Metric":
        return Metric(A.METRIC_NONE)

    @staticmethod
    def rmse() -> "Metric":
        return Metric(A.METRIC_RMSE)

    @staticmethod
    def mae() -> "Metric":
        return Metric(A.METRIC_MAE)

    @staticmethod
    def mape() -> "Metric":
        return Metric(A.METRIC_MAPE)

    @staticmethod
    def logloss() -> "Metric":
        return Metric(A.METRIC_LOGLOSS)

    @staticmethod
    def accuracy_at(threshold: float = 0.5) -> "Metric":
        if not 0.0 < threshold <= 1.0:
            raise ValueError("accuracy threshold must be in (0, 1]")
        return Metric(A.METRIC_ACCURACY, threshold=float(threshold))

    @staticmethod
    def margin_accuracy() -> "Metric":
        return Metric(A.METRIC_MARGIN_ACCURACY)

    @staticmethod
    def multi_logloss() -> "Metric":
        return Metric(A.METRIC_MLOGLOSS)

    multiclass_logloss = multi_logloss

    @staticmethod
    def multiclass_accuracy() -> "Metric":
        return Metric(A.METRIC_MULTICLASS_ACCURACY)

    @staticmethod
    def quantile(alpha: Sequence[float]) -> "Metric":
        a = tuple(float(v) for v in alpha)
        if not a or not all(0.0 < v < 1.0 for v in a):
            raise ValueError("quantile metric needs at least one alpha, each in (0, 1)")
        return Metric(A.METRIC_QUANTILE, alpha=a)

    @staticmethod
    def poisson_deviance() -> "Metric":
        return Metric(A.METRIC_POISSON_DEVIANCE)


# ---------------------------------------------------------------------------
# Config
# ---------------------------------------------------------------------------
@dataclass
class GBDTConfig:
    """GBDTConfig with the reference bindings' keyword names and defaults."""

    n_estimators: int = 100
    learning_rate: float = 0.3
    objective: Optional[Objective] = None
    metric: Optional[Metric] = None
    growth_strategy: str = GrowthStrategy.Depthwise
    max_depth: int = 6
    n_leaves: int = 31
    max_onehot_cats: int = 4
    l1: float = 0.0
    l2: float = 1.0
    min_gain_to_split: float = 0.0
    min_child_weight: float = 1.0
    min_samples_leaf: int = 1
    subsample: float = 1.0
    colsample_bytree: float = 1.0
    colsample_bylevel: float = 1.0
    max_bins: int = 256
    enable_bundling: bool = True
    sparsity_threshold: float = 0.9
    max_categorical_cardinality: int = 0
    binning_sample_cnt: int = 200_000
    cache_size: int = 8
    early_stopping_rounds: Optional[int] = None
    seed: int = 42
    verbosity: int = 0
    # Extension knobs (no counterpart in the reference).
    schedule: int = Schedule.Graph
    goss_top_rate: float = 0.0
    goss_other_rate: float = 0.0

    def binning_c(self) -> A.BinningConfig:
        return A.BinningConfig(
            int(self.max_bins), float(self.sparsity_threshold), int(bool(self.enable_bundling)),
            int(self.max_categorical_cardinality), int(self.binning_sample_cnt),
        )

    def to_c(self) -> A.GBDTConfigC:
        c = A.GBDTConfigC()
        obj = self.objective or Objective.squared()
        c.objective = obj.kind
        c.n_classes = obj.n_classes
        c.metric = A.METRIC_DEFAULT if self.metric is None else self.metric.kind
        c.n_trees = int(self.n_estimators)
        c.learning_rate = float(self.learning_rate)
        c.growth = A.GROW_DEPTHWISE if self.growth_strategy == GrowthStrategy.Depthwise else A.GROW_LEAFWISE
        c.max_depth = int(self.max_depth)
        c.max_leaves = int(self.n_leaves)
        c.max_onehot_cats = int(self.max_onehot_cats)
        c.lambda_ = float(self.l2)
        c.alpha = float(self.l1)
        c.min_child_weight = float(self.min_child_weight)
        c.min_gain = float(self.min_gain_to_split)
        c.min_samples_leaf = int(self.min_samples_leaf)
        c.subsample = float(self.subsample)
        c.colsample_bytree = float(self.colsample_bytree)
        c.colsample_bylevel = float(self.colsample_bylevel)
        c.binning = self.binning_c()
        c.early_stopping_rounds = int(self.early_stopping_rounds or 0)
        c.cache_size = int(self.cache_size)
        c.seed = int(self.seed)
        c.verbosity = int(self.verbosity)
        c.huber_delta = float(obj.delta)
        c.accuracy_threshold = float(self.metric.threshold) if self.metric is not None else 0.5
        c.metric_huber_delta = 1.0
        alphas = obj.alpha or (self.metric.alpha if self.metric is not None else ())
        if alphas:
            arr = np.ascontiguousarray(alphas, np.float32)
            c._alphas_keep = arr  # the struct borrows the buffer
            c.alphas = A.ptr(arr)
            c.n_alphas = len(arr)
        else:
            c.n_alphas = 0
        return c

    def ext_c(self, n_threads: int = 0) -> A.ExtConfig:
        e = A.ExtConfig()
        e.schedule = int(self.schedule)
        e.n_threads = int(n_threads)
        e.goss_top_rate = float(self.goss_top_rate)
        e.goss_other_rate = float(self.goss_other_rate)
        e.query_group_sizes = None
        e.n_query_groups = 0
        e.autotune = 1
        return e


# ---------------------------------------------------------------------------
# Runtime objects
# ---------------------------------------------------------------------------
class Context:
    """One HIP device + stream (bst_ctx).  Not re-entrant."""

    _default = None
    _lock = threading.Lock()

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        A.check(A.lib().bst_ctx_create(int(device), C.byref(self.h)))
        self.device = device

    @classmethod
    def default(cls) -> "Context":
        with cls._lock:
            if cls._default is None:
                cls._default = Context(0)
            return cls._default

    def sync(self):
        A.check(A.lib().bst_ctx_sync(self.h))

    @property
    def stream(self) -> int:
        return A.lib().bst_ctx_stream(self.h) or 0

    def __del__(self):
        if getattr(self, "h", None) and A._LIB is not None:
            A._LIB.bst_ctx_destroy(self.h)
            self.h = None


class Dataset:
    """Features [n_samples, n_features] (float32, NaN = missing), labels, weights."""

    def __init__(self, features, labels=None, weights=None, feature_names: Optional[Sequence[str]] = None,
                 categorical_features: Optional[Sequence[int]] = None):
        X = np.asarray(features, dtype=np.float32)
        if X.ndim != 2:
            raise ValueError("features must be 2-D [n_samples, n_features]")
        self.n_samples, self.n_features = X.shape
        # Feature-major copy, the layout of the reference Dataset (raw/dataset.rs:43-52).
        self.feature_major = np.ascontiguousarray(X.T)
        self.labels = None if labels is None else np.ascontiguousarray(labels, dtype=np.float32)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        self.feature_names = list(feature_names) if feature_names is not None else None
        self.categorical_features = list(categorical_features or [])

    @classmethod
    def from_feature_major(cls, feature_major, labels=None, weights=None, categorical_features=None):
        obj = cls.__new__(cls)
        fm = np.ascontiguousarray(feature_major, dtype=np.float32)
        obj.n_features, obj.n_samples = fm.shape
        obj.feature_major = fm
        obj.labels = None if labels is None else np.ascontiguousarray(labels, dtype=np.float32)
        obj.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        obj.feature_names = None
        obj.categorical_features = list(categorical_features or [])
        return obj

    @property
    def shape(self):
        return (self.n_samples, self.n_features)

    def is_categorical_mask(self) -> Optional[np.ndarray]:
        if not self.categorical_features:
            return None
        m = np.zeros(self.n_features, dtype=np.uint8)
        m[list(self.categorical_features)] = 1
        return m


class BinnedDataset:
    """Device-resident quantised dataset (bst_dataset)."""

    def __init__(self, ctx: Context, data: Dataset, binning: A.BinningConfig):
        self.ctx = ctx
        self.h = C.c_void_p()
        cat = data.is_categorical_mask()
        A.check(A.lib().bst_dataset_create(ctx.h, A.ptr(data.feature_major), data.n_samples,
                                           data.n_features, A.ptr(cat), C.byref(binning),
                                           C.byref(self.h)))
        self.n_rows, self.n_feat = data.n_samples, data.n_features

    def bins(self) -> np.ndarray:
        out = np.zeros((self.n_feat, self.n_rows), dtype=np.uint16)
        A.check(A.lib().bst_dataset_get_bins(self.h, A.ptr(out)))
        return out

    def feature_info(self, f: int) -> dict:
        nb, mt, db, cat, tr = C.c_uint32(), C.c_int32(), C.c_uint32(), C.c_int32(), C.c_int32()
        A.check(A.lib().bst_dataset_feature_info(self.h, f, C.byref(nb), C.byref(mt), C.byref(db),
                                                 C.byref(cat), C.byref(tr)))
        return dict(n_bins=nb.value, missing=mt.value, default_bin=db.value,
                    categorical=bool(cat.value), trivial=bool(tr.value))

    def bounds(self, f: int) -> np.ndarray:
        out = np.zeros(512, dtype=np.float64)
        k = A.lib().bst_dataset_bin_bounds(self.h, f, A.ptr(out), 512)
        if k < 0:
            A.check(k)
        return out[:k].copy()

    def layout(self) -> dict:
        ne = C.c_int32()
        e2o = np.zeros(self.n_feat, np.uint32)
        nb = np.zeros(self.n_feat, np.uint32)
        off = np.zeros(self.n_feat, np.uint32)
        total = A.lib().bst_dataset_hist_layout(self.h, C.byref(ne), A.ptr(e2o), A.ptr(nb), A.ptr(off))
        if total < 0:
            A.check(total)
        e = ne.value
        return dict(n_eff=e, eff_to_orig=e2o[:e].copy(), n_bins=nb[:e].copy(), offset=off[:e].copy(),
                    total_bins=total)

    def __del__(self):
        if getattr(self, "h", None) and A._LIB is not None:
            A._LIB.bst_dataset_destroy(self.h)
            self.h = None


class GBDTModel:
    """A trained forest (bst_model)."""

    def __init__(self, handle: C.c_void_p, ctx: Context):
        self.h = handle
        self.ctx = ctx
        info = A.ModelInfo()
        A.check(A.lib().bst_model_info_get(self.h, C.byref(info)))
        self.info = info
        self.metric_history = None
        self.train_stats = None

    # -- GBDTModel.train (model.rs:115-143; bindings _boosters_rs.pyi:497) --
    @staticmethod
    def train(train: Dataset, config: Optional[GBDTConfig] = None, val_set: Optional[Dataset] = None,
              n_threads: int = 0, ctx: Optional[Context] = None) -> "GBDTModel":
        cfg = config or GBDTConfig()
        if train.labels is None:
            raise ValueError("training dataset must have labels")
        ctx = ctx or Context.default()
        binned = BinnedDataset(ctx, train, cfg.binning_c())
        return GBDTModel.train_binned(binned, train.labels, cfg, val_set=val_set,
                                      weights=train.weights, n_threads=n_threads)

    @staticmethod
    def train_binned(binned: BinnedDataset, labels, cfg: GBDTConfig, val_set: Optional[Dataset] = None,
                     weights=None, n_threads: int = 0) -> "GBDTModel":
        y = np.ascontiguousarray(labels, dtype=np.float32).reshape(-1)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        vX = vy = None
        nv = 0
        if val_set is not None:
            vX, vy, nv = val_set.feature_major, val_set.labels, val_set.n_samples
        h = C.c_void_p()
        cc, ee = cfg.to_c(), cfg.ext_c(n_threads)
        A.check(A.lib().bst_train(binned.ctx.h, binned.h, A.ptr(y), len(y), A.ptr(w), A.ptr(vX),
                                  A.ptr(vy), nv, C.byref(cc), C.byref(ee), C.byref(h)))
        m = GBDTModel(h, binned.ctx)
        ms_total, ms_tree, hr = C.c_double(), C.c_double(), C.c_int64()
        A.check(A.lib().bst_last_train_stats(binned.ctx.h, C.byref(ms_total), C.byref(ms_tree), C.byref(hr)))
        m.train_stats = dict(ms_total=ms_total.value, ms_per_tree=ms_tree.value, hist_rows=hr.value)
        buf = np.zeros(max(int(cfg.n_estimators), 1), np.float64)
        nr = C.c_int64()
        A.check(A.lib().bst_last_train_metric(binned.ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
        m.metric_history = buf[: nr.value].copy()
        return m

    @property
    def n_trees(self) -> int:
        return int(self.info.n_trees)

    @property
    def n_features(self) -> int:
        return int(self.info.n_features)

    def export(self):
        arrays = A.alloc_forest(self.info)
        fa = A.forest_struct(arrays)
        A.check(A.lib().bst_model_export(self.h, C.byref(fa)))
        return arrays

    @staticmethod
    def from_arrays(info: A.ModelInfo, arrays, ctx: Optional[Context] = None) -> "GBDTModel":
        ctx = ctx or Context.default()
        fa = A.forest_struct(arrays)
        h = C.c_void_p()
        A.check(A.lib().bst_model_import(ctx.h, C.byref(info), C.byref(fa), C.byref(h)))
        return GBDTModel(h, ctx)

    # -- persistence (persist/mod.rs:88-160; bindings src/model/gbdt.rs:259-327) --
    def to_bytes(self, compression_level: int = 3) -> bytes:
        """The model as .bstr bytes (zstd level 3 by default, 0 = uncompressed)."""
        buf, size = C.POINTER(C.c_uint8)(), C.c_int64()
        A.check(A.lib().bst_model_save_bstr(self.h, int(compression_level), C.byref(buf), C.byref(size)))
        try:
            return C.string_at(buf, size.value)
        finally:
            A.lib().bst_buffer_free(buf)

    def to_json_bytes(self, pretty: bool = False) -> bytes:
        """The model as .bstr.json bytes (UTF-8)."""
        buf, size = C.c_void_p(), C.c_int64()
        A.check(A.lib().bst_model_save_json(self.h, int(pretty), C.byref(buf), C.byref(size)))
        try:
            return C.string_at(buf, size.value)
        finally:
            A.lib().bst_buffer_free(buf)

    @staticmethod
    def from_bytes(data: bytes, ctx: Optional[Context] = None) -> "GBDTModel":
        """Load .bstr bytes; raises BoostersError (code BST_E_FORMAT) on bad input."""
        h = C.c_void_p()
        A.check(A.lib().bst_model_load_bstr(ctx.h if ctx else None, bytes(data), len(data), C.byref(h)))
        return GBDTModel(h, ctx)

    @staticmethod
    def from_json_bytes(data: bytes, ctx: Optional[Context] = None) -> "GBDTModel":
        h = C.c_void_p()
        A.check(A.lib().bst_model_load_json(ctx.h if ctx else None, bytes(data), len(data), C.byref(h)))
        return GBDTModel(h, ctx)

    def save(self, path: str, compression_level: int = 3) -> None:
        A.check(A.lib().bst_model_save_bstr_file(self.h, os.fsencode(path), int(compression_level)))

    @staticmethod
    def load(path: str, ctx: Optional[Context] = None) -> "GBDTModel":
        h = C.c_void_p()
        A.check(A.lib().bst_model_load_bstr_file(ctx.h if ctx else None, os.fsencode(path), C.byref(h)))
        return GBDTModel(h, ctx)

    def __reduce__(self):
        """Pickle through the .bstr bytes (bindings src/model/gbdt.rs:180-188)."""
        return (GBDTModel.from_bytes, (self.to_bytes(),))

    def _predict(self, data, raw: bool, leaf_index: bool = False):
        fm = data.feature_major if isinstance(data, Dataset) else np.ascontiguousarray(data, np.float32)
        F, n = fm.shape
        out = np.zeros((self.info.n_groups, n), np.float32)
        li = np.zeros((self.info.n_trees, n), np.uint32) if leaf_index else None
        if self.ctx is None:  # loaded without a context: bind the default device now
            self.ctx = Context.default()
        A.check(A.lib().bst_predict(self.ctx.h, self.h, A.ptr(fm), n, F, int(raw), A.ptr(out), A.ptr(li)))
        return (out, li) if leaf_index else out

    # -- predict / predict_raw (model.rs:217-258), (n_samples, n_outputs) --
    def predict(self, data: Dataset, n_threads: int = 0) -> np.ndarray:
        return self._predict(data, raw=False).T.copy()

    def predict_raw(self, data: Dataset, n_threads: int = 0) -> np.ndarray:
        return self._predict(data, raw=True).T.copy()

    def predict_feature_major(self, feature_major, raw: bool = False, leaf_index: bool = False):
        """[n_groups, n_samples] output (the Rust core's layout), optional leaf ids."""
        return self._predict(feature_major, raw=raw, leaf_index=leaf_index)

    def __del__(self):
        if getattr(self, "h", None) and A._LIB is not None:
            A._LIB.bst_model_destroy(self.h)
            self.h = None


class Server:
    """Continuous-batching prediction server (bst_server)."""

    def __init__(self, model: GBDTModel, hbm_budget_bytes: int = 1 << 30, max_inflight: int = 4):
        self.model = model
        self.h = C.c_void_p()
        if model.ctx is None:
            model.ctx = Context.default()
        A.check(A.lib().bst_server_create(model.ctx.h, model.h, int(hbm_budget_bytes),
                                          int(max_inflight), C.byref(self.h)))
        self._keep = {}

    def submit(self, feature_major: np.ndarray) -> int:
        fm = np.ascontiguousarray(feature_major, np.float32)
        F, n = fm.shape
        out = np.zeros((self.model.info.n_groups, n), np.float32)
        t = C.c_uint64()
        A.check(A.lib().bst_server_submit(self.h, A.ptr(fm), n, F, A.ptr(out), C.byref(t)))
        self._keep[t.value] = (fm, out)
        return t.value

    def wait(self, ticket: int):
        lat = C.c_int64()
        A.check(A.lib().bst_server_wait(self.h, C.c_uint64(ticket), C.byref(lat)))
        _, out = self._keep.pop(ticket)
        return out, lat.value

    def close(self):
        if getattr(self, "h", None):
            A.check(A.lib().bst_server_destroy(self.h))
            self.h = None

    def __del__(self):
        if getattr(self, "h", None) and A._LIB is not None:
            A._LIB.bst_server_destroy(self.h)
            self.h = None
