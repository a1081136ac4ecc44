"""GPU: per-round training and validation metrics and early stopping against
the oracle (evaluate_round, training/eval.rs:218-270; EarlyStopping,
training/callback.rs:67-118; trainer.rs:341-390).

The engine's metric is summed in f64 over device blocks and the oracle's in
one f64 loop, so values agree to summation order: within 1e-10 relative.
The forests themselves are identical (same trees), so early stopping must
pick the same best round and truncate to the same number of trees.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import BinnedDataset, Dataset, GBDTModel
from tests.conftest import binary_nan_data, make_cfg, make_ext, multiclass_cat_data, regression_data

pytestmark = pytest.mark.gpu


def _train(ctx, oracle, X, y, cfg, is_cat=None, schedule=2, weights=None, valid=None):
    oracle.set_exp_mode(1)
    vX, vy = valid if valid is not None else (None, None)
    om = oracle.train(oracle.Binned(X, is_cat), y, cfg, make_ext(0), weights=weights,
                      valid_X=vX, valid_y=vy)
    ds = Dataset.from_feature_major(X, categorical_features=None if is_cat is None
                                    else list(np.nonzero(is_cat)[0]))
    b = BinnedDataset(ctx, ds, A.BinningConfig(256, 0.9, 1, 0, 200000))
    h = C.c_void_p()
    nv = 0 if vy is None else len(vy)
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), A.ptr(weights), A.ptr(vX), A.ptr(vy), nv,
                              C.byref(cfg), C.byref(make_ext(schedule)), C.byref(h)))
    gm = GBDTModel(h, ctx)
    buf = np.zeros(4096, np.float64)
    nr = C.c_int64()
    A.check(A.lib().bst_last_train_metric(ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
    return gm, om, buf[: nr.value].copy()


CASES = {
    "logistic": (lambda: binary_nan_data(20_000, 16, seed=51) + (None,),
                 dict(objective=1, n_trees=12, learning_rate=0.3, max_depth=6)),
    "squared": (lambda: regression_data(15_000, 20, seed=52) + (None,),
                dict(n_trees=12, learning_rate=0.2, max_depth=5)),
    "softmax": (lambda: multiclass_cat_data(12_000, 10, 3, 4, card=12, seed=53),
                dict(objective=2, n_classes=4, n_trees=8, growth=1, max_leaves=24)),
}


@pytest.mark.parametrize("schedule", [1, 2])
@pytest.mark.parametrize("case", sorted(CASES))
def test_training_metric_per_round(ctx, oracle, case, schedule):
    build, kw = CASES[case]
    X, y, is_cat = build()
    cfg = make_cfg(**kw)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, is_cat, schedule)
    want = om.metric_history()
    assert len(hist) == len(want) == cfg.n_trees
    np.testing.assert_allclose(hist, want, rtol=1e-10)
    assert hist[-1] < hist[0]


def test_weighted_training_metric(ctx, oracle):
    """The training metric is weighted by the training weights (eval.rs:218-260)."""
    X, y = binary_nan_data(15_000, 12, seed=54)
    w = (np.random.default_rng(1).random(len(y)) * 2 + 0.1).astype(np.float32)
    cfg = make_cfg(objective=1, n_trees=6, max_depth=5)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, weights=w)
    np.testing.assert_allclose(hist, om.metric_history(), rtol=1e-10)


@pytest.mark.parametrize("case", sorted(CASES))
def test_validation_metric_per_round(ctx, oracle, case):
    build, kw = CASES[case]
    X, y, is_cat = build()
    n = X.shape[1]
    cut = n * 4 // 5
    Xt, yt = np.ascontiguousarray(X[:, :cut]), y[:cut]
    Xv, yv = np.ascontiguousarray(X[:, cut:]), y[cut:]
    cfg = make_cfg(**kw)
    gm, om, hist = _train(ctx, oracle, Xt, yt, cfg, is_cat, valid=(Xv, yv))
    want = om.metric_history()
    assert len(hist) == len(want) == cfg.n_trees
    np.testing.assert_allclose(hist, want, rtol=1e-10)
    # The validation value is the metric of the model's own predictions.
    metric = {"logistic": A.METRIC_LOGLOSS, "squared": A.METRIC_RMSE, "softmax": A.METRIC_MLOGLOSS}[case]
    last = oracle.metric(metric, gm.predict_feature_major(Xv), yv)
    assert hist[-1] == pytest.approx(last, rel=1e-9)


@pytest.mark.parametrize("schedule", [1, 2])
def test_early_stopping_truncates_like_the_oracle(ctx, oracle, schedule):
    """Overfitting on a small noisy set: the validation loss turns up, training
    stops `patience` rounds after the best one and the forest keeps best + 1
    rounds (callback.rs:78-118, trainer.rs:392-400)."""
    X, y = regression_data(1500, 6, seed=55)
    Xv, yv = regression_data(1500, 6, seed=56)
    cfg = make_cfg(n_trees=300, learning_rate=0.5, max_depth=6, early_stopping_rounds=5)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, schedule=schedule, valid=(Xv, yv))
    want = om.metric_history()
    assert len(want) < 300
    assert len(hist) == len(want)
    np.testing.assert_allclose(hist, want, rtol=1e-10)
    best = int(np.argmin(hist))
    assert len(hist) - 1 - best == 5
    assert gm.info.n_trees == om.info.n_trees == best + 1


def test_early_stopping_softmax_counts_rounds(ctx, oracle):
    """With K class trees per round the forest keeps (best + 1) * K trees."""
    X, y, is_cat = multiclass_cat_data(3000, 6, 2, 3, card=8, seed=57)
    Xv, yv, _ = multiclass_cat_data(3000, 6, 2, 3, card=8, seed=58)
    cfg = make_cfg(objective=2, n_classes=3, n_trees=200, learning_rate=0.8, max_depth=6,
                   early_stopping_rounds=3)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, is_cat, valid=(Xv, yv))
    want = om.metric_history()
    assert len(hist) == len(want) < 200
    best = int(np.argmin(hist))
    assert gm.info.n_trees == om.info.n_trees == 3 * (best + 1)
