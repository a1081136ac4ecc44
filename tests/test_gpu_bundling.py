"""GPU: Exclusive Feature Bundling (data/binned/bundling.rs; dataset.rs:229-236,
460-568,898-952; find.rs:472-573; partition.rs:296-337; grower.rs:424-440,
576-650) against the oracle.

- The effective layout (bundles first, then the other features in feature
  order), the codes decoded per feature and every histogram bin are the
  oracle's, bit for bit, with and without conflicting rows and with more
  rows than the 10000 the conflict graph samples.
- Trained forests are the oracle's on every schedule, depth-wise and
  leaf-wise: splits on a bundle are categorical splits stored on its first
  feature with the codes as categories, code 0 (every feature at its
  default) goes right whatever the set, and the children's totals are those
  of the rows that reach them.  Predictions of the stored trees (raw
  features) match.
- enable_bundling = 0 gives the unbundled layout.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import BinnedDataset, Dataset, GBDTModel
from tests import forest_compare
from tests.conftest import make_cfg, make_ext, sparse_cat_data

pytestmark = pytest.mark.gpu


def _binning(enable=1):
    return A.BinningConfig(256, 0.9, enable, 0, 200000)


def _both(ctx, oracle, X, is_cat, enable=1):
    ds = Dataset.from_feature_major(X, categorical_features=list(np.nonzero(is_cat)[0]))
    return BinnedDataset(ctx, ds, _binning(enable)), oracle.Binned(X, is_cat, _binning(enable))


@pytest.mark.parametrize("n,conflicts", [(6000, 0), (6000, 1), (30000, 2)])
def test_bundled_layout_and_bins(ctx, oracle, n, conflicts):
    X, _, is_cat = sparse_cat_data(n, 3, 6, card=5, conflicts=conflicts, seed=n + conflicts)
    b, o = _both(ctx, oracle, X, is_cat)
    lb, lo = b.layout(), o.layout()
    assert o.bundles(), "the data should bundle"
    for k in ("eff_to_orig", "n_bins", "offset"):
        np.testing.assert_array_equal(lb[k], lo[k], err_msg=k)
    assert lb["total_bins"] == lo["total_bins"]
    np.testing.assert_array_equal(b.bins(), o.bins())
    rng = np.random.default_rng(1)
    grad = rng.standard_normal(n).astype(np.float32)
    hess = (rng.random(n) + 0.1).astype(np.float32)
    rows = np.sort(rng.choice(n, n // 3, replace=False)).astype(np.uint32)
    for r in (None, rows):
        out = np.zeros(lb["total_bins"] * 2, np.float64)
        A.check(A.lib().bst_kernel_histogram(ctx.h, b.h, A.ptr(grad), A.ptr(hess), A.ptr(r),
                                             0 if r is None else len(r), A.ptr(out), len(out)))
        ref = o.histogram(grad, hess, r).reshape(-1)
        sel = np.arange(n) if r is None else r
        assert np.max(np.abs(out - ref)) <= 1e-9 * (np.abs(grad[sel]).sum() + hess[sel].sum())


def test_unbundled_layout_when_disabled(ctx, oracle):
    X, _, is_cat = sparse_cat_data(5000, 2, 4, seed=7)
    b, o = _both(ctx, oracle, X, is_cat, enable=0)
    assert o.bundles() == []
    np.testing.assert_array_equal(b.layout()["eff_to_orig"], np.arange(6))
    np.testing.assert_array_equal(b.bins(), o.bins())


CASES = {
    "graph_depthwise": (dict(objective=A.OBJ_SQUARED, max_depth=5, n_trees=6), 2, 8000),
    "batched_depthwise": (dict(objective=A.OBJ_SQUARED, max_depth=5, n_trees=6), 1, 8000),
    "plain_depthwise": (dict(objective=A.OBJ_SQUARED, max_depth=4, n_trees=4), 0, 8000),
    "graph_leafwise_logistic": (dict(objective=A.OBJ_LOGISTIC, growth=1, max_leaves=24,
                                     n_trees=6), 2, 8000),
    "batched_leafwise": (dict(objective=A.OBJ_SQUARED, growth=1, max_leaves=20, n_trees=5), 1,
                         8000),
    "graph_sampled_conflicts": (dict(objective=A.OBJ_SQUARED, max_depth=6, n_trees=5), 2, 40000),
}


def _bundle_splits(arr, first_features):
    """Splits stored on a bundle's first feature: (count, count with code 0 in the set)."""
    k = np.nonzero((arr["is_leaf"] == 0) & (arr["split_type"] == 1) &
                   np.isin(arr["split_index"], first_features))[0]
    with0 = sum(int(arr["cat_size"][i] > 0 and (int(arr["cat_words"][arr["cat_start"][i]]) & 1))
                for i in k)
    return len(k), with0


@pytest.mark.parametrize("case", sorted(CASES))
def test_bundled_training_matches_oracle(ctx, oracle, case):
    kw, schedule, n = CASES[case]
    X, y, is_cat = sparse_cat_data(n, 3, 6, card=5, conflicts=2 if n > 10000 else 0, seed=11)
    if kw["objective"] == A.OBJ_LOGISTIC:
        y = (y > np.median(y)).astype(np.float32)
    cfg = make_cfg(learning_rate=0.3, **kw)
    b, o = _both(ctx, oracle, X, is_cat)
    bundles = o.bundles()
    assert bundles
    oracle.set_exp_mode(1)
    om = oracle.train(o, y, cfg, make_ext(0))
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(make_ext(schedule)), C.byref(h)))
    gm = GBDTModel(h, ctx)
    ga = gm.export()
    assert forest_compare.compare(ga, om.arrays)["divergence"] is None
    buf = np.zeros(4096, np.float64)
    nr = C.c_int64()
    A.check(A.lib().bst_last_train_metric(ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
    np.testing.assert_allclose(buf[: nr.value], om.metric_history(), rtol=1e-10)
    n_split, with0 = _bundle_splits(ga, [bd[0] for bd in bundles])
    assert n_split > 0, "no split on a bundle"
    Xt, _, _ = sparse_cat_data(3000, 3, 6, card=5, seed=12)
    np.testing.assert_allclose(gm.predict_feature_major(Xt, raw=True),
                               oracle.predict(om.info, om.arrays, Xt, raw=True),
                               rtol=1e-6, atol=1e-6)


def test_bundle_split_with_code_zero_in_the_set(ctx, oracle):
    """Across the cases above, splits whose category set holds code 0 occur,
    so the left sums the children take (code 0 routed right) are exercised:
    train the depth-wise graph case with more trees and count them."""
    X, y, is_cat = sparse_cat_data(8000, 3, 6, card=5, seed=11)
    cfg = make_cfg(objective=A.OBJ_SQUARED, max_depth=5, n_trees=20, learning_rate=0.3)
    b, o = _both(ctx, oracle, X, is_cat)
    oracle.set_exp_mode(1)
    om = oracle.train(o, y, cfg, make_ext(0))
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(make_ext(2)), C.byref(h)))
    ga = GBDTModel(h, ctx).export()
    assert forest_compare.compare(ga, om.arrays)["divergence"] is None
    n_split, with0 = _bundle_splits(om.arrays, [bd[0] for bd in o.bundles()])
    assert with0 > 0 and n_split > with0
