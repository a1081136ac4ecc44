"""GPU: configuration the reference accepts beyond the benchmark workloads,
trained by the engine and by the oracle on the same data: the forests must be
the oracle's (tie-aware comparator) and the per-round metric within 1e-10.

- colsample_bylevel < 1 (sampling/column.rs:199-224; grower.rs:279-280,443-444),
  alone and on top of colsample_bytree, depth-wise and leaf-wise.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import BinnedDataset, Dataset, GBDTModel
from tests import forest_compare
from tests.conftest import binary_nan_data, make_cfg, make_ext, multiclass_cat_data

pytestmark = pytest.mark.gpu


def _train(ctx, oracle, X, y, cfg, schedule=2, is_cat=None):
    oracle.set_exp_mode(1)
    om = oracle.train(oracle.Binned(X, is_cat), y, cfg, make_ext(0))
    ds = Dataset.from_feature_major(X, categorical_features=None if is_cat is None
                                    else list(np.nonzero(is_cat)[0]))
    b = BinnedDataset(ctx, ds, A.BinningConfig(256, 0.9, 1, 0, 200000))
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(make_ext(schedule)), C.byref(h)))
    gm = GBDTModel(h, ctx)
    buf = np.zeros(4096, np.float64)
    nr = C.c_int64()
    A.check(A.lib().bst_last_train_metric(ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
    return gm, om, buf[: nr.value].copy()


BYLEVEL = {
    "depthwise": dict(colsample_bylevel=0.5, max_depth=6),
    "depthwise_with_bytree": dict(colsample_bylevel=0.6, colsample_bytree=0.7, max_depth=5),
    "leafwise": dict(colsample_bylevel=0.4, growth=1, max_leaves=24),
    "tiny_rate": dict(colsample_bylevel=0.01, max_depth=4),
}


@pytest.mark.parametrize("case", sorted(BYLEVEL))
def test_colsample_bylevel_matches_oracle(ctx, oracle, case):
    X, y = binary_nan_data(20_000, 30, seed=91)
    cfg = make_cfg(objective=A.OBJ_LOGISTIC, n_trees=6, learning_rate=0.3, **BYLEVEL[case])
    gm, om, hist = _train(ctx, oracle, X, y, cfg)
    assert forest_compare.compare(gm.export(), om.arrays)["divergence"] is None
    np.testing.assert_allclose(hist, om.metric_history(), rtol=1e-10)
    # a tree uses features from several level subsets, never more than the tree's
    used = set(gm.export()["split_index"][~gm.export()["is_leaf"].astype(bool)].tolist())
    assert len(used) >= 2


def test_colsample_bylevel_softmax_class_trees(ctx, oracle):
    X, y, is_cat = multiclass_cat_data(12_000, 10, 3, 3, card=8, seed=92)
    cfg = make_cfg(objective=A.OBJ_SOFTMAX, n_classes=3, n_trees=4, max_depth=4,
                   colsample_bylevel=0.5)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, is_cat=is_cat)
    assert forest_compare.compare(gm.export(), om.arrays)["divergence"] is None
    np.testing.assert_allclose(hist, om.metric_history(), rtol=1e-10)
