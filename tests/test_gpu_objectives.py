"""GPU: the remaining objectives and metrics against the oracle
(objectives/regression.rs:68-333, classification.rs:76-123, trainer.rs:265-291,
metrics/regression.rs:55-250, classification.rs:46-192).

- The gradient kernel's (g, h) equal the oracle's bit for bit (same f32
  operations, correctly rounded division and square root, the f64-rounded exp
  both sides use).
- Trained forests are the oracle's (the tie-aware comparator), including the
  renewed leaves of the quantile objectives (absolute, pinball: leaf = lr x
  the alpha-quantile of the leaf's residuals), weighted and unweighted.
- Per-round metric values match within 1e-10 relative (f64 sums in another
  order); early stopping on a higher-is-better metric truncates where the
  oracle does.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import BinnedDataset, Dataset, GBDTModel
from tests import forest_compare
from tests.conftest import binary_nan_data, make_cfg, make_ext, multiclass_cat_data, regression_data

pytestmark = pytest.mark.gpu

f32 = np.float32


def _cfg(**kw):
    alphas = kw.pop("alphas", None)
    c = make_cfg(**kw)
    if alphas is not None:
        arr = np.ascontiguousarray(alphas, np.float32)
        c._alphas_keep = arr
        c.alphas = A.ptr(arr)
        c.n_alphas = len(arr)
    return c


def _train(ctx, oracle, X, y, cfg, schedule=2, weights=None, valid=None, is_cat=None):
    oracle.set_exp_mode(1)
    vX, vy = valid if valid is not None else (None, None)
    om = oracle.train(oracle.Binned(X, is_cat), y, cfg, make_ext(0), weights=weights,
                      valid_X=vX, valid_y=vy)
    ds = Dataset.from_feature_major(X, categorical_features=None if is_cat is None
                                    else list(np.nonzero(is_cat)[0]))
    b = BinnedDataset(ctx, ds, A.BinningConfig(256, 0.9, 1, 0, 200000))
    h = C.c_void_p()
    nv = 0 if vy is None else len(vy)
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), A.ptr(weights), A.ptr(vX), A.ptr(vy),
                              nv, C.byref(cfg), C.byref(make_ext(schedule)), C.byref(h)))
    gm = GBDTModel(h, ctx)
    buf = np.zeros(4096, np.float64)
    nr = C.c_int64()
    A.check(A.lib().bst_last_train_metric(ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
    return gm, om, buf[: nr.value].copy()


def _targets(kind, X, rng):
    n = X.shape[1]
    if kind == A.OBJ_POISSON:
        return rng.poisson(np.exp(0.6 * np.nan_to_num(X[0]))).astype(f32)
    if kind == A.OBJ_HINGE:
        return (np.nan_to_num(X[0]) + 0.4 * rng.normal(0, 1, n) > 0).astype(f32)
    return (2 * np.nan_to_num(X[0]) - np.nan_to_num(X[1]) + rng.standard_t(2.0, n)).astype(f32)


GRAD_CASES = {
    "absolute": dict(objective=A.OBJ_ABSOLUTE),
    "hinge": dict(objective=A.OBJ_HINGE),
    "pinball": dict(objective=A.OBJ_PINBALL, alphas=[0.1, 0.5, 0.9]),
    "huber": dict(objective=A.OBJ_PSEUDO_HUBER, huber_delta=0.7),
    "poisson": dict(objective=A.OBJ_POISSON),
}


@pytest.mark.parametrize("case", sorted(GRAD_CASES))
def test_gradient_kernel_bit_exact(ctx, oracle, case):
    kw = dict(GRAD_CASES[case])
    cfg = _cfg(**kw)
    K = len(kw.get("alphas", [0]))
    rng = np.random.default_rng(71)
    n = 100_003
    p = rng.normal(0, 3, (K, n)).astype(f32)
    p[:, :7] = np.array([0.0, 31.0, -31.0, 1.0, -1.0, 2.5, 1e-30], f32)[None, :]
    y = (rng.poisson(2.0, n) if case == "poisson" else rng.normal(0, 3, n)).astype(f32)
    if case == "hinge":
        y = (rng.random(n) < 0.5).astype(f32)
    y[:2] = p[0, :2]  # zero residuals (signum of +0.0)
    w = rng.uniform(0.1, 3, n).astype(f32)
    oracle.set_exp_mode(1)
    for weights in (None, w):
        og, oh = oracle.gradients_cfg(cfg, p, y, weights)
        g = np.zeros_like(p)
        h = np.zeros_like(p)
        A.check(A.lib().bst_kernel_gradients_cfg(ctx.h, C.byref(cfg), A.ptr(p), A.ptr(y),
                                                 A.ptr(weights), n, A.ptr(g), A.ptr(h)))
        np.testing.assert_array_equal(g, og)
        np.testing.assert_array_equal(h, oh)


TRAIN_CASES = {
    "absolute": (dict(objective=A.OBJ_ABSOLUTE), 2),
    "pinball": (dict(objective=A.OBJ_PINBALL, alphas=[0.25, 0.75]), 2),
    "huber": (dict(objective=A.OBJ_PSEUDO_HUBER, huber_delta=1.5), 2),
    "poisson": (dict(objective=A.OBJ_POISSON), 2),
    "hinge": (dict(objective=A.OBJ_HINGE), 2),
    "absolute_batched": (dict(objective=A.OBJ_ABSOLUTE, growth=1, max_leaves=20), 1),
}


@pytest.mark.parametrize("case", sorted(TRAIN_CASES))
def test_training_matches_oracle(ctx, oracle, case):
    kw, schedule = TRAIN_CASES[case]
    X, _ = regression_data(20_000, 12, seed=72)
    y = _targets(kw["objective"], X, np.random.default_rng(73))
    cfg = _cfg(n_trees=8, learning_rate=0.3, max_depth=5, **kw)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, schedule=schedule)
    assert forest_compare.compare(gm.export(), om.arrays)["divergence"] is None
    np.testing.assert_allclose(hist, om.metric_history(), rtol=1e-10)
    Xt, _ = regression_data(4000, 12, seed=74)
    np.testing.assert_allclose(gm.predict_feature_major(Xt, raw=True),
                               oracle.predict(om.info, om.arrays, Xt, raw=True), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("objective", [A.OBJ_ABSOLUTE, A.OBJ_PINBALL])
def test_weighted_leaf_renewal(ctx, oracle, objective):
    """Weighted quantiles: stable sort by residual, f32 totals in the leaf's
    row order, f32 running sums in sorted order (utils.rs:81-126)."""
    X, _ = regression_data(12_000, 8, seed=75)
    rng = np.random.default_rng(76)
    y = _targets(objective, X, rng)
    w = rng.uniform(0.2, 4.0, len(y)).astype(f32)
    kw = dict(objective=objective)
    if objective == A.OBJ_PINBALL:
        kw["alphas"] = [0.3]
    cfg = _cfg(n_trees=5, learning_rate=0.5, max_depth=4, **kw)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, weights=w)
    assert forest_compare.compare(gm.export(), om.arrays)["divergence"] is None
    np.testing.assert_allclose(hist, om.metric_history(), rtol=1e-10)


METRIC_CASES = {
    "mae": (A.METRIC_MAE, dict(), "reg"),
    "mape": (A.METRIC_MAPE, dict(), "reg"),
    "huber": (A.METRIC_HUBER, dict(metric_huber_delta=0.8), "reg"),
    "poisson_deviance": (A.METRIC_POISSON_DEVIANCE, dict(objective=A.OBJ_POISSON), "pois"),
    "quantile": (A.METRIC_QUANTILE, dict(alphas=[0.3]), "reg"),
    "accuracy": (A.METRIC_ACCURACY, dict(objective=A.OBJ_LOGISTIC, accuracy_threshold=0.6), "bin"),
    "margin_accuracy": (A.METRIC_MARGIN_ACCURACY, dict(objective=A.OBJ_LOGISTIC), "bin"),
    "multiclass_accuracy": (A.METRIC_MULTICLASS_ACCURACY,
                            dict(objective=A.OBJ_SOFTMAX, n_classes=4, growth=1, max_leaves=16), "multi"),
}


@pytest.mark.parametrize("case", sorted(METRIC_CASES))
def test_metric_per_round(ctx, oracle, case):
    metric, kw, data = METRIC_CASES[case]
    is_cat = None
    if data == "bin":
        X, y = binary_nan_data(15_000, 10, seed=77)
    elif data == "multi":
        X, y, is_cat = multiclass_cat_data(10_000, 8, 2, 4, card=10, seed=78)
    else:
        X, _ = regression_data(15_000, 10, seed=79)
        y = _targets(A.OBJ_POISSON if data == "pois" else A.OBJ_SQUARED, X, np.random.default_rng(80))
        if data == "reg":
            y = np.abs(y) + 0.5  # MAPE divides by |y|
    cfg = _cfg(n_trees=6, max_depth=5, metric=metric, **kw)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, is_cat=is_cat)
    want = om.metric_history()
    assert len(hist) == len(want) == 6
    np.testing.assert_allclose(hist, want, rtol=1e-10)


def test_early_stopping_on_a_higher_is_better_metric(ctx, oracle):
    """Hinge's default metric, margin accuracy, is higher-is-better: training
    stops `patience` rounds after its best validation value
    (callback.rs:78-118, metrics/mod.rs:286-301)."""
    X, y = binary_nan_data(3000, 8, seed=81)
    Xv, yv = binary_nan_data(3000, 8, seed=82)
    cfg = _cfg(objective=A.OBJ_HINGE, n_trees=200, learning_rate=0.5, max_depth=6,
               early_stopping_rounds=4)
    gm, om, hist = _train(ctx, oracle, X, y, cfg, valid=(Xv, yv))
    want = om.metric_history()
    assert len(hist) == len(want) < 200
    np.testing.assert_allclose(hist, want, rtol=1e-12)
    best = int(np.argmax(hist))
    assert gm.info.n_trees == om.info.n_trees == best + 1
