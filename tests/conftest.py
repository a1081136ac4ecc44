import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built engine library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    from boosters_amd import Context

    return Context.default()


def make_cfg(**kw):
    """A bst_gbdt_config with the reference defaults (config.rs:111-226) plus overrides."""
    from boosters_amd import _abi as A

    c = A.GBDTConfigC()
    c.objective = A.OBJ_SQUARED
    c.n_classes = 0
    c.metric = A.METRIC_DEFAULT
    c.n_trees = 100
    c.learning_rate = 0.3
    c.growth = A.GROW_DEPTHWISE
    c.max_depth = 6
    c.max_leaves = 31
    c.max_onehot_cats = 4
    c.lambda_ = 1.0
    c.alpha = 0.0
    c.min_child_weight = 1.0
    c.min_gain = 0.0
    c.min_samples_leaf = 1
    c.subsample = 1.0
    c.colsample_bytree = 1.0
    c.colsample_bylevel = 1.0
    c.binning = A.BinningConfig(256, 0.9, 1, 0, 200000)
    c.early_stopping_rounds = 0
    c.cache_size = 8
    c.seed = 42
    c.huber_delta = 1.0
    c.n_alphas = 0
    c.accuracy_threshold = 0.5
    c.metric_huber_delta = 1.0
    c.verbosity = 0
    for k, v in kw.items():
        setattr(c, "lambda_" if k == "lambda" else k, v)
    return c


def make_ext(schedule=1, n_threads=0):
    from boosters_amd import _abi as A

    e = A.ExtConfig()
    e.schedule = schedule
    e.n_threads = n_threads
    e.goss_top_rate = 0.0
    e.goss_other_rate = 0.0
    e.query_group_sizes = None
    e.n_query_groups = 0
    e.autotune = 1
    return e


def regression_data(n, F, seed=0):
    """Config-1 generator: X ~ N(0,1), y = x0 + 0.5 x1^2 + N(0, 0.1^2)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((F, n), dtype=np.float32)
    y = (X[0] + 0.5 * X[1] ** 2 + rng.normal(0, 0.1, n)).astype(np.float32)
    return X, y


def binary_nan_data(n, F, seed=0, nan_p=0.1):
    """Config-2 generator: X ~ N(0,1), 10% NaN, y ~ Bernoulli(sigmoid(w . x[0:10]))."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((F, n), dtype=np.float32)
    mask = rng.random((F, n)) < nan_p
    k = min(10, F)
    w = rng.standard_normal(k)
    score = w @ X[:k].astype(np.float64)
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-score))).astype(np.float32)
    X[mask] = np.nan
    return X, y


def multiclass_cat_data(n, n_num, n_cat, K, card=32, seed=0):
    """Config-3 generator: numeric N(0,1) + categorical ints, argmax of projections + Gumbel."""
    rng = np.random.default_rng(seed)
    Xn = rng.standard_normal((n_num, n), dtype=np.float32)
    Xc = rng.integers(0, card, size=(n_cat, n)).astype(np.float32)
    W = rng.normal(0, 1.0 / np.sqrt(max(n_num, 1)), size=(K, n_num))
    Cm = rng.normal(0, 1.0 / np.sqrt(max(n_cat, 1)), size=(K, n_cat, card))
    z = W @ Xn.astype(np.float64)
    for j in range(n_cat):
        z += Cm[:, j, Xc[j].astype(np.int64)]
    z += rng.gumbel(size=(K, n))
    y = np.argmax(z, axis=0).astype(np.float32)
    X = np.concatenate([Xn, Xc], axis=0)
    is_cat = np.zeros(n_num + n_cat, np.uint8)
    is_cat[n_num:] = 1
    return X, y, is_cat


def ranking_data(n, F, seed=0, zero_p=0.3, group_lo=5, group_hi=500):
    """Config-4 generator: X ~ U[0,1] with exact zeros (p = 0.3), query sizes
    ~ U{group_lo..group_hi} until the rows are covered (last truncated), score
    s = sum_f a_f [x_f > t_f], label = quintile of s (0..4)."""
    rng = np.random.default_rng(seed)
    X = rng.random((F, n), dtype=np.float32)
    X[rng.random((F, n)) < zero_p] = 0.0
    sizes = []
    tot = 0
    while tot < n:
        sz = int(rng.integers(group_lo, group_hi + 1))
        sz = min(sz, n - tot)
        sizes.append(sz)
        tot += sz
    a = rng.standard_normal(F)
    t = rng.random(F)
    s = (a[:, None] * (X > t[:, None])).sum(axis=0)
    q = np.quantile(s, [0.2, 0.4, 0.6, 0.8])
    y = np.searchsorted(q, s, side="right").astype(np.float32)
    return X, y, np.array(sizes, np.int32)


def sparse_cat_data(n, n_dense, n_sparse, card=6, active=0.05, conflicts=0, seed=0):
    """Dense N(0,1) features followed by sparse categorical ones: each row is
    set (a category 1..card-1; 0 is the default) in at most one sparse
    feature, each feature on about `active` of the rows, plus `conflicts`
    rows set in the first two sparse features at once.  y = x0 + a per
    (feature, category) effect + N(0, 0.1^2).  Returns X, y, is_cat."""
    rng = np.random.default_rng(seed)
    Xd = rng.standard_normal((n_dense, n), dtype=np.float32)
    Xs = np.zeros((n_sparse, n), np.float32)
    owner = rng.integers(0, n_sparse, n)
    on = rng.random(n) < active * n_sparse
    cats = rng.integers(1, card, n).astype(np.float32)
    Xs[owner[on], np.nonzero(on)[0]] = cats[on]
    if conflicts and n_sparse >= 2:
        rows = rng.choice(np.nonzero(~on)[0], conflicts, replace=False)  # rows set nowhere else
        Xs[0, rows] = rng.integers(1, card, conflicts)
        Xs[1, rows] = rng.integers(1, card, conflicts)
    eff = rng.normal(0, 1.5, (n_sparse, card))
    eff[:, 0] = 0.0
    y = (Xd[0] if n_dense else np.zeros(n, np.float32)).astype(np.float64)
    for j in range(n_sparse):
        y += eff[j, Xs[j].astype(np.int64)]
    y = (y + rng.normal(0, 0.1, n)).astype(np.float32)
    X = np.concatenate([Xd, Xs], axis=0)
    is_cat = np.zeros(n_dense + n_sparse, np.uint8)
    is_cat[n_dense:] = 1
    return X, y, is_cat
