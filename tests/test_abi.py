"""The C ABI library loads and exports every entry point the header declares.

No compute calls: these run on the CPU.  Error paths that need no device are
exercised (null handles, config validation, no-GPU context creation).
"""
import ctypes as C
import os
import re

import pytest

from boosters_amd import _abi as A

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "boosters_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bst_\w+)\s*\(", text)))


def test_header_symbols_exported():
    lib = C.CDLL(A.LIB_PATH)
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_declares_every_symbol():
    assert sorted(set(A.exported_symbols())) == declared()
    A.lib()  # argtypes set for all of them


def test_defaults_mirror_reference():
    # GBDTConfig defaults, model/gbdt/config.rs:111-226.
    c = A.GBDTConfigC()
    A.lib().bst_gbdt_config_default(C.byref(c))
    assert (c.n_trees, round(c.learning_rate, 6), c.max_depth, c.max_leaves) == (100, 0.3, 6, 31)
    assert (c.lambda_, c.alpha, c.min_child_weight, c.min_gain, c.min_samples_leaf) == (1.0, 0.0, 1.0, 0.0, 1)
    assert (c.max_onehot_cats, c.cache_size, c.seed) == (4, 8, 42)
    assert (c.binning.max_bins, c.binning.sample_cnt, c.binning.enable_bundling) == (256, 200000, 1)
    assert c.growth == A.GROW_DEPTHWISE and c.metric == A.METRIC_DEFAULT


@pytest.mark.parametrize("field,value", [
    ("learning_rate", 0.0), ("n_trees", 0), ("subsample", 1.5), ("colsample_bytree", 0.0),
    ("lambda_", -1.0), ("alpha", -0.1), ("min_child_weight", -1.0), ("min_gain", -1.0),
])
def test_config_validation(field, value):
    # GBDTConfig::validate, config.rs:252-310.
    c = A.GBDTConfigC()
    A.lib().bst_gbdt_config_default(C.byref(c))
    assert A.lib().bst_gbdt_config_validate(C.byref(c)) == A.BST_OK
    setattr(c, field, value)
    assert A.lib().bst_gbdt_config_validate(C.byref(c)) == A.BST_E_CONFIG
    assert A.lib().bst_last_error()


def test_null_handles_are_errors_not_crashes():
    L = A.lib()
    assert L.bst_ctx_sync(None) == A.BST_E_INVALID
    assert L.bst_model_info_get(None, None) == A.BST_E_INVALID
    assert L.bst_predict(None, None, None, 0, 0, 0, None, None) == A.BST_E_INVALID
    assert L.bst_server_wait(None, C.c_uint64(1), None) == A.BST_E_INVALID
    assert b"null" in L.bst_last_error()


def test_model_import_validation():
    # A forest with a child id out of range is a format error (persist/error.rs).
    info = A.ModelInfo()
    info.n_trees, info.n_groups, info.n_features, info.total_nodes = 1, 1, 2, 3
    arrays = A.alloc_forest(info)
    arrays["tree_offset"][:] = [0, 3]
    arrays["is_leaf"][:] = [0, 1, 1]
    arrays["left"][0], arrays["right"][0] = 1, 7
    h = C.c_void_p()
    rc = A.lib().bst_model_import(None, C.byref(info), C.byref(A.forest_struct(arrays)), C.byref(h))
    assert rc == A.BST_E_FORMAT
    arrays["right"][0] = 2
    arrays["split_index"][0] = 5  # feature beyond n_features
    rc = A.lib().bst_model_import(None, C.byref(info), C.byref(A.forest_struct(arrays)), C.byref(h))
    assert rc == A.BST_E_FORMAT
    arrays["split_index"][0] = 1
    rc = A.lib().bst_model_import(None, C.byref(info), C.byref(A.forest_struct(arrays)), C.byref(h))
    assert rc == A.BST_OK
    A.lib().bst_model_destroy(h)


def test_no_gpu_context_fails_cleanly():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    rc = A.lib().bst_ctx_create(0, C.byref(h))
    assert rc == A.BST_E_DEVICE
