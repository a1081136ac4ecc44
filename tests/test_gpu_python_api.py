"""GPU: the Python surface (boosters_amd.model, mirroring the reference
bindings' GBDTModel.train / predict / predict_raw, _boosters_rs.pyi:497,543,557)
against the oracle.

Datasets are [n_samples, n_features] as in the bindings; predictions come back
as (n_samples, n_outputs).  The trained forest must be the oracle's (same
trees, leaf values within 1e-6), predictions within 1e-6, the per-round
metric the oracle's, and a pickled model must predict the same values.
"""
import pickle

import numpy as np
import pytest

from boosters_amd import (Dataset, GBDTConfig, GBDTModel, GrowthStrategy, Metric, Objective)
from tests import forest_compare
from tests.conftest import (binary_nan_data, make_cfg, make_ext, multiclass_cat_data,
                            regression_data)

pytestmark = pytest.mark.gpu


def _oracle_model(oracle, Xfm, y, cfg_c, is_cat=None, valid=None):
    oracle.set_exp_mode(1)
    vX, vy = valid if valid is not None else (None, None)
    return oracle.train(oracle.Binned(Xfm, is_cat), y, cfg_c, make_ext(0), valid_X=vX, valid_y=vy)


def test_regression_train_predict_predict_raw(ctx, oracle):
    Xfm, y = regression_data(8000, 20, seed=81)
    X = np.ascontiguousarray(Xfm.T)
    cfg = GBDTConfig(n_estimators=15, learning_rate=0.1, max_depth=5)
    m = GBDTModel.train(Dataset(X, y), cfg)
    assert m.n_trees == 15 and m.n_features == 20
    om = _oracle_model(oracle, Xfm, y, cfg.to_c())
    r = forest_compare.compare(m.export(), om.arrays)
    assert r["divergence"] is None
    Xt, _ = regression_data(3000, 20, seed=82)
    p = m.predict(Dataset(np.ascontiguousarray(Xt.T)))
    assert p.shape == (3000, 1)
    ref = oracle.predict(om.info, om.arrays, Xt)
    np.testing.assert_allclose(p[:, 0], ref[0], rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(m.predict_raw(Dataset(np.ascontiguousarray(Xt.T))), p)  # identity
    np.testing.assert_allclose(m.metric_history, om.metric_history(), rtol=1e-10)


def test_logistic_with_validation_and_early_stopping(ctx, oracle):
    Xfm, y = binary_nan_data(6000, 12, seed=83)
    cut = 4800
    X = np.ascontiguousarray(Xfm.T)
    train = Dataset(X[:cut], y[:cut])
    val = Dataset(X[cut:], y[cut:])
    cfg = GBDTConfig(n_estimators=300, learning_rate=0.6, max_depth=6, objective=Objective.logistic(),
                     early_stopping_rounds=4)
    m = GBDTModel.train(train, cfg, val_set=val)
    om = _oracle_model(oracle, np.ascontiguousarray(Xfm[:, :cut]), y[:cut], cfg.to_c(),
                       valid=(np.ascontiguousarray(Xfm[:, cut:]), y[cut:]))
    assert m.n_trees == om.info.n_trees < 300
    np.testing.assert_allclose(m.metric_history, om.metric_history(), rtol=1e-10)
    prob = m.predict(val)
    raw = m.predict_raw(val)
    assert prob.shape == raw.shape == (len(y) - cut, 1)
    ref = oracle.predict(om.info, om.arrays, np.ascontiguousarray(Xfm[:, cut:]))
    np.testing.assert_allclose(prob[:, 0], ref[0], rtol=1e-6, atol=1e-7)
    assert np.all((prob > 0) & (prob < 1))
    np.testing.assert_allclose(prob, 1 / (1 + np.exp(-raw.astype(np.float64))), rtol=1e-6, atol=1e-7)


def test_softmax_categorical_and_pickle(ctx, oracle):
    Xfm, y, is_cat = multiclass_cat_data(9000, 10, 3, 4, card=12, seed=84)
    X = np.ascontiguousarray(Xfm.T)
    cats = list(np.nonzero(is_cat)[0])
    cfg = GBDTConfig(n_estimators=6, learning_rate=0.3, objective=Objective.softmax(4),
                     growth_strategy=GrowthStrategy.Leafwise, n_leaves=24,
                     metric=Metric.multi_logloss())
    m = GBDTModel.train(Dataset(X, y, categorical_features=cats), cfg)
    assert m.n_trees == 24
    om = _oracle_model(oracle, Xfm, y, cfg.to_c(), is_cat)
    assert forest_compare.compare(m.export(), om.arrays)["divergence"] is None
    Xt, _, _ = multiclass_cat_data(2000, 10, 3, 4, card=12, seed=85)
    dt = Dataset(np.ascontiguousarray(Xt.T), categorical_features=cats)
    p = m.predict(dt)
    assert p.shape == (2000, 4)
    np.testing.assert_allclose(p.sum(axis=1), 1.0, rtol=1e-5)
    ref = oracle.predict(om.info, om.arrays, Xt)
    np.testing.assert_allclose(p, ref.T, rtol=1e-6, atol=1e-7)
    m2 = pickle.loads(pickle.dumps(m))
    assert m2.n_trees == m.n_trees
    np.testing.assert_array_equal(m2.predict(dt), p)
    np.testing.assert_array_equal(m2.predict_raw(dt), m.predict_raw(dt))


def test_config_defaults_and_validation():
    # GBDTConfig defaults of the bindings (_boosters_rs.pyi:455); invalid
    # values are refused at train time with the reference's ConfigError code.
    c = GBDTConfig().to_c()
    d = make_cfg()
    for f in ("n_trees", "max_depth", "max_leaves", "max_onehot_cats", "min_samples_leaf", "seed"):
        assert getattr(c, f) == getattr(d, f), f
    assert abs(c.learning_rate - 0.3) < 1e-7 and c.lambda_ == 1.0 and c.alpha == 0.0
    Xfm, y = regression_data(500, 4, seed=86)
    from boosters_amd import _abi as A
    with pytest.raises(A.BoostersError) as e:
        GBDTModel.train(Dataset(np.ascontiguousarray(Xfm.T), y), GBDTConfig(learning_rate=0.0))
    assert e.value.code == A.BST_E_CONFIG
