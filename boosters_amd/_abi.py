"""ctypes mirror of include/boosters_amd.h (structs, constants, library loader).

The engine library ``libboosters_amd.so`` is built in-tree next to this file.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libboosters_amd.so")

BST_OK = 0
BST_E_CONFIG = -1
BST_E_TARGETS = -2
BST_E_SHAPE = -3
BST_E_OOM = -4
BST_E_DEVICE = -5
BST_E_FORMAT = -6
BST_E_INVALID = -7

OBJ_SQUARED, OBJ_LOGISTIC, OBJ_SOFTMAX, OBJ_LAMBDARANK = 0, 1, 2, 3
OBJ_ABSOLUTE, OBJ_HINGE, OBJ_PINBALL, OBJ_PSEUDO_HUBER, OBJ_POISSON = 4, 5, 6, 7, 8
GROW_DEPTHWISE, GROW_LEAFWISE = 0, 1
METRIC_DEFAULT, METRIC_NONE, METRIC_RMSE, METRIC_LOGLOSS, METRIC_MLOGLOSS = -1, 0, 1, 2, 3
(METRIC_MAE, METRIC_MAPE, METRIC_QUANTILE, METRIC_HUBER, METRIC_POISSON_DEVIANCE, METRIC_ACCURACY,
 METRIC_MARGIN_ACCURACY, METRIC_MULTICLASS_ACCURACY) = 4, 5, 6, 7, 8, 9, 10, 11
TRANSFORM_IDENTITY, TRANSFORM_SIGMOID, TRANSFORM_SOFTMAX = 0, 1, 2
SCHED_PLAIN, SCHED_BATCHED, SCHED_GRAPH = 0, 1, 2


class BinningConfig(C.Structure):
    _fields_ = [
        ("max_bins", C.c_uint32),
        ("sparsity_threshold", C.c_float),
        ("enable_bundling", C.c_int32),
        ("max_categorical_cardinality", C.c_uint32),
        ("sample_cnt", C.c_uint64),
    ]


class GBDTConfigC(C.Structure):
    _fields_ = [
        ("objective", C.c_int32),
        ("n_classes", C.c_uint32),
        ("metric", C.c_int32),
        ("n_trees", C.c_uint32),
        ("learning_rate", C.c_float),
        ("growth", C.c_int32),
        ("max_depth", C.c_uint32),
        ("max_leaves", C.c_uint32),
        ("max_onehot_cats", C.c_uint32),
        ("lambda_", C.c_float),
        ("alpha", C.c_float),
        ("min_child_weight", C.c_float),
        ("min_gain", C.c_float),
        ("min_samples_leaf", C.c_uint32),
        ("subsample", C.c_float),
        ("colsample_bytree", C.c_float),
        ("colsample_bylevel", C.c_float),
        ("binning", BinningConfig),
        ("early_stopping_rounds", C.c_uint32),
        ("cache_size", C.c_uint64),
        ("seed", C.c_uint64),
        ("verbosity", C.c_int32),
        ("huber_delta", C.c_float),
        ("n_alphas", C.c_uint32),
        ("alphas", C.POINTER(C.c_float)),
        ("accuracy_threshold", C.c_float),
        ("metric_huber_delta", C.c_double),
    ]


class ExtConfig(C.Structure):
    _fields_ = [
        ("schedule", C.c_int32),
        ("n_threads", C.c_int32),
        ("goss_top_rate", C.c_float),
        ("goss_other_rate", C.c_float),
        ("query_group_sizes", C.POINTER(C.c_int32)),
        ("n_query_groups", C.c_int64),
        ("autotune", C.c_int32),
        ("hist_fg", C.c_int32),
        ("hist_rows_min", C.c_int32),
        ("hist_blocks_per_cu", C.c_int32),
        ("hist_priv", C.c_int32),
    ]


class ModelInfo(C.Structure):
    _fields_ = [
        ("n_trees", C.c_int64),
        ("n_groups", C.c_int64),
        ("n_features", C.c_int64),
        ("total_nodes", C.c_int64),
        ("total_cat_words", C.c_int64),
        ("output_transform", C.c_int32),
        ("objective", C.c_int32),
    ]


class ForestArrays(C.Structure):
    _fields_ = [
        ("base_score", C.POINTER(C.c_float)),
        ("tree_offset", C.POINTER(C.c_int64)),
        ("tree_group", C.POINTER(C.c_uint32)),
        ("split_index", C.POINTER(C.c_uint32)),
        ("split_threshold", C.POINTER(C.c_float)),
        ("left", C.POINTER(C.c_uint32)),
        ("right", C.POINTER(C.c_uint32)),
        ("default_left", C.POINTER(C.c_uint8)),
        ("is_leaf", C.POINTER(C.c_uint8)),
        ("split_type", C.POINTER(C.c_uint8)),
        ("leaf_value", C.POINTER(C.c_float)),
        ("gain", C.POINTER(C.c_float)),
        ("cover", C.POINTER(C.c_float)),
        ("cat_start", C.POINTER(C.c_uint32)),
        ("cat_size", C.POINTER(C.c_uint32)),
        ("cat_words", C.POINTER(C.c_uint32)),
    ]


class SplitInfoC(C.Structure):
    _fields_ = [
        ("feature", C.c_uint32),
        ("gain", C.c_float),
        ("default_left", C.c_int32),
        ("is_categorical", C.c_int32),
        ("bin", C.c_uint32),
        ("left_cats", C.c_uint32 * 8),
    ]


# Forest exchange arrays as numpy buffers ------------------------------------
FOREST_FIELDS = [
    # name, dtype, size key
    ("base_score", "float32", "n_groups"),
    ("tree_offset", "int64", "n_trees+1"),
    ("tree_group", "uint32", "n_trees"),
    ("split_index", "uint32", "total_nodes"),
    ("split_threshold", "float32", "total_nodes"),
    ("left", "uint32", "total_nodes"),
    ("right", "uint32", "total_nodes"),
    ("default_left", "uint8", "total_nodes"),
    ("is_leaf", "uint8", "total_nodes"),
    ("split_type", "uint8", "total_nodes"),
    ("leaf_value", "float32", "total_nodes"),
    ("gain", "float32", "total_nodes"),
    ("cover", "float32", "total_nodes"),
    ("cat_start", "uint32", "total_nodes"),
    ("cat_size", "uint32", "total_nodes"),
    ("cat_words", "uint32", "total_cat_words"),
]

_CT = {"float32": C.c_float, "int64": C.c_int64, "uint32": C.c_uint32, "uint8": C.c_uint8}


def alloc_forest(info: ModelInfo):
    """Allocate numpy arrays for every field of bst_forest_arrays."""
    import numpy as np

    sizes = {
        "n_groups": info.n_groups,
        "n_trees": info.n_trees,
        "n_trees+1": info.n_trees + 1,
        "total_nodes": info.total_nodes,
        "total_cat_words": max(info.total_cat_words, 1),
    }
    arrays = {name: np.zeros(sizes[key], dtype=dt) for name, dt, key in FOREST_FIELDS}
    return arrays


def forest_struct(arrays) -> ForestArrays:
    fa = ForestArrays()
    for name, dt, _ in FOREST_FIELDS:
        arr = arrays[name]
        assert arr.flags["C_CONTIGUOUS"] and str(arr.dtype) == dt, name
        setattr(fa, name, arr.ctypes.data_as(C.POINTER(_CT[dt])))
    return fa


def ptr(arr, ctype=None):
    """ctypes pointer to a numpy array (or None)."""
    if arr is None:
        return None
    if ctype is None:
        ctype = {
            "float32": C.c_float,
            "float64": C.c_double,
            "uint32": C.c_uint32,
            "int32": C.c_int32,
            "uint16": C.c_uint16,
            "uint8": C.c_uint8,
            "int64": C.c_int64,
        }[str(arr.dtype)]
    return arr.ctypes.data_as(C.POINTER(ctype))


class BoostersError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_LIB = None


def _declare(lib):
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "bst_last_error": (C.c_char_p, []),
        "bst_version": (C.c_char_p, []),
        "bst_gbdt_config_default": (None, [P(GBDTConfigC)]),
        "bst_ext_config_default": (None, [P(ExtConfig)]),
        "bst_gbdt_config_validate": (C.c_int, [P(GBDTConfigC)]),
        "bst_ctx_create": (C.c_int, [C.c_int, P(vp)]),
        "bst_ctx_destroy": (C.c_int, [vp]),
        "bst_ctx_sync": (C.c_int, [vp]),
        "bst_ctx_stream": (vp, [vp]),
        "bst_ctx_set_profiling": (C.c_int, [vp, C.c_int32]),
        "bst_ctx_kernel_stats": (C.c_int, [vp, C.c_int32, P(C.c_int64), P(C.c_double), P(C.c_double)]),
        "bst_kernel_copy": (C.c_int, [vp, vp, vp, C.c_int64]),
        "bst_device_alloc": (C.c_int, [vp, C.c_int64, P(vp)]),
        "bst_device_free": (C.c_int, [vp, vp]),
        "bst_memcpy_h2d": (C.c_int, [vp, vp, vp, C.c_int64]),
        "bst_memcpy_d2h": (C.c_int, [vp, vp, vp, C.c_int64]),
        "bst_dataset_create": (
            C.c_int,
            [vp, P(C.c_float), C.c_int64, C.c_int32, P(C.c_uint8), P(BinningConfig), P(vp)],
        ),
        "bst_dataset_create_device": (
            C.c_int,
            [vp, vp, C.c_int64, C.c_int32, P(C.c_uint8), P(BinningConfig), P(vp)],
        ),
        "bst_dataset_destroy": (C.c_int, [vp]),
        "bst_dataset_feature_info": (
            C.c_int,
            [vp, C.c_int32, P(C.c_uint32), P(C.c_int32), P(C.c_uint32), P(C.c_int32), P(C.c_int32)],
        ),
        "bst_dataset_bin_bounds": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int64]),
        "bst_dataset_get_bins": (C.c_int, [vp, P(C.c_uint16)]),
        "bst_train": (
            C.c_int,
            [
                vp, vp, P(C.c_float), C.c_int64, P(C.c_float), P(C.c_float), P(C.c_float),
                C.c_int64, P(GBDTConfigC), P(ExtConfig), P(vp),
            ],
        ),
        "bst_train_device": (
            C.c_int,
            [vp, vp, vp, C.c_int64, vp, P(GBDTConfigC), P(ExtConfig), P(vp)],
        ),
        "bst_last_train_stats": (C.c_int, [vp, P(C.c_double), P(C.c_double), P(C.c_int64)]),
        "bst_last_train_metric": (C.c_int, [vp, P(C.c_double), C.c_int64, P(C.c_int64)]),
        "bst_ctx_set_keep_predictions": (C.c_int, [vp, C.c_int32]),
        "bst_last_train_predictions": (C.c_int, [vp, P(C.c_float), C.c_int64, P(C.c_int64)]),
        "bst_last_train_tuning": (
            C.c_int, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32),
                      P(C.c_double)]),
        "bst_model_info_get": (C.c_int, [vp, P(ModelInfo)]),
        "bst_model_export": (C.c_int, [vp, P(ForestArrays)]),
        "bst_model_import": (C.c_int, [vp, P(ModelInfo), P(ForestArrays), P(vp)]),
        "bst_model_destroy": (C.c_int, [vp]),
        "bst_model_save_bstr": (C.c_int, [vp, C.c_int32, P(P(C.c_uint8)), P(C.c_int64)]),
        "bst_model_load_bstr": (C.c_int, [vp, C.c_char_p, C.c_int64, P(vp)]),
        "bst_model_save_bstr_file": (C.c_int, [vp, C.c_char_p, C.c_int32]),
        "bst_model_load_bstr_file": (C.c_int, [vp, C.c_char_p, P(vp)]),
        "bst_model_save_json": (C.c_int, [vp, C.c_int32, P(C.c_void_p), P(C.c_int64)]),
        "bst_model_load_json": (C.c_int, [vp, C.c_char_p, C.c_int64, P(vp)]),
        "bst_buffer_free": (None, [vp]),
        "bst_predict": (
            C.c_int,
            [vp, vp, P(C.c_float), C.c_int64, C.c_int32, C.c_int32, P(C.c_float), P(C.c_uint32)],
        ),
        "bst_predict_device": (
            C.c_int,
            [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, vp],
        ),
        "bst_predict_device_noalloc": (
            C.c_int, [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, vp],
        ),
        "bst_server_create": (C.c_int, [vp, vp, C.c_int64, C.c_int32, P(vp)]),
        "bst_server_submit": (
            C.c_int,
            [vp, P(C.c_float), C.c_int64, C.c_int32, P(C.c_float), P(C.c_uint64)],
        ),
        "bst_server_wait": (C.c_int, [vp, C.c_uint64, P(C.c_int64)]),
        "bst_server_destroy": (C.c_int, [vp]),
        "bst_kernel_gradients": (
            C.c_int,
            [vp, C.c_int32, C.c_int32, P(C.c_float), P(C.c_float), P(C.c_float), C.c_int64,
             P(C.c_float), P(C.c_float)],
        ),
        "bst_kernel_gradients_cfg": (
            C.c_int,
            [vp, P(GBDTConfigC), P(C.c_float), P(C.c_float), P(C.c_float), C.c_int64,
             P(C.c_float), P(C.c_float)],
        ),
        "bst_kernel_lambdarank": (
            C.c_int,
            [vp, P(C.c_float), P(C.c_float), P(C.c_float), C.c_int64, P(C.c_int32), C.c_int64,
             P(C.c_float), P(C.c_float)],
        ),
        "bst_kernel_histogram": (
            C.c_int,
            [vp, vp, P(C.c_float), P(C.c_float), P(C.c_uint32), C.c_int64, P(C.c_double), C.c_int64],
        ),
        "bst_dataset_hist_layout": (
            C.c_int,
            [vp, P(C.c_int32), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)],
        ),
        "bst_kernel_partition": (
            C.c_int,
            [vp, vp, P(C.c_uint32), C.c_int64, C.c_int32, C.c_uint32, C.c_int32, P(C.c_uint32),
             C.c_int32, P(C.c_int64)],
        ),
        "bst_kernel_partition_fused": (
            C.c_int,
            [vp, vp, P(C.c_uint32), C.c_int64, C.c_int32, C.c_uint32, C.c_int32, P(C.c_uint32),
             C.c_int32, C.c_int32, P(C.c_int64)],
        ),
        "bst_kernel_find_split": (
            C.c_int,
            [vp, vp, P(C.c_double), C.c_double, C.c_double, C.c_uint32, P(GBDTConfigC), C.c_int32,
             P(SplitInfoC)],
        ),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def exported_symbols():
    """Every entry point the header declares (checked by the ABI test)."""
    return [
        "bst_last_error", "bst_version", "bst_gbdt_config_default", "bst_ext_config_default",
        "bst_gbdt_config_validate", "bst_ctx_create", "bst_ctx_destroy", "bst_ctx_sync",
        "bst_ctx_stream", "bst_ctx_set_profiling", "bst_ctx_kernel_stats", "bst_kernel_copy",
        "bst_device_alloc",
        "bst_device_free", "bst_memcpy_h2d", "bst_memcpy_d2h",
        "bst_dataset_create", "bst_dataset_create_device",
        "bst_dataset_destroy", "bst_dataset_feature_info", "bst_dataset_bin_bounds",
        "bst_dataset_get_bins", "bst_train", "bst_train_device", "bst_last_train_stats",
        "bst_last_train_metric", "bst_ctx_set_keep_predictions", "bst_last_train_predictions",
        "bst_last_train_tuning", "bst_model_info_get", "bst_model_export", "bst_model_import",
        "bst_model_destroy", "bst_model_save_bstr", "bst_model_load_bstr",
        "bst_model_save_bstr_file", "bst_model_load_bstr_file", "bst_model_save_json",
        "bst_model_load_json", "bst_buffer_free", "bst_predict", "bst_predict_device", "bst_predict_device_noalloc", "bst_server_create",
        "bst_server_submit", "bst_server_wait", "bst_server_destroy", "bst_kernel_gradients", "bst_kernel_gradients_cfg",
        "bst_kernel_lambdarank",
        "bst_kernel_histogram", "bst_dataset_hist_layout", "bst_kernel_partition", "bst_kernel_partition_fused",
        "bst_kernel_find_split",
    ]


def lib():
    """Load the in-tree engine library; raise if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"boosters_amd native library not built: {LIB_PATH} is missing "
                "(run __graft_entry__.build())"
            )
        _LIB = _declare(C.CDLL(LIB_PATH))
    return _LIB


def check(code: int):
    if code != BST_OK:
        raise BoostersError(code, lib().bst_last_error().decode(errors="replace"))
