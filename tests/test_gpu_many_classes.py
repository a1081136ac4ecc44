"""GPU: softmax with more classes than one launch holds (classification.rs:
141-187, trainer.rs:252, predictor.rs:197-302, transform.rs:83-103).

- Gradients of K = 20 and 37 classes equal the oracle's bit for bit (the
  wide kernel does the register kernel's f32 operations in class order).
- Training grows the class trees of a round in batches of 16 (the last batch
  short) on the host-driven schedules; the forests are the oracle's and the
  per-round multi-logloss matches.
- Prediction accumulates the outputs in windows of 16 (each launch adds only
  its own trees, in forest order) and applies the softmax over all classes
  after: raw and transformed outputs match the oracle's within 1e-6, leaf
  ids bit-exact, one-phase, two-phase and allocation-free launches agree bit
  for bit, and a model file round-trips.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import BinnedDataset, Dataset, GBDTModel
from tests import forest_compare
from tests.conftest import make_cfg, make_ext, multiclass_cat_data

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [20, 37])
def test_wide_softmax_gradients_bit_exact(ctx, oracle, K):
    rng = np.random.default_rng(K)
    n = 30_001
    p = rng.normal(0, 2, (K, n)).astype(np.float32)
    y = rng.integers(0, K, n).astype(np.float32)
    w = rng.uniform(0.2, 3, n).astype(np.float32)
    oracle.set_exp_mode(1)
    for weights in (None, w):
        g = np.zeros_like(p)
        h = np.zeros_like(p)
        A.check(A.lib().bst_kernel_gradients(ctx.h, A.OBJ_SOFTMAX, K, A.ptr(p), A.ptr(y),
                                             A.ptr(weights), n, A.ptr(g), A.ptr(h)))
        og, oh = oracle.gradients(A.OBJ_SOFTMAX, p, y, weights)
        np.testing.assert_array_equal(g, og)
        np.testing.assert_array_equal(h, oh)


def _train(ctx, oracle, K, schedule, n=12_000, **kw):
    X, y, is_cat = multiclass_cat_data(n, 8, 2, K, card=12, seed=K + schedule)
    cfg = make_cfg(objective=A.OBJ_SOFTMAX, n_classes=K, n_trees=3, learning_rate=0.3, **kw)
    oracle.set_exp_mode(1)
    om = oracle.train(oracle.Binned(X, is_cat), y, cfg, make_ext(0))
    ds = Dataset.from_feature_major(X, categorical_features=list(np.nonzero(is_cat)[0]))
    b = BinnedDataset(ctx, ds, A.BinningConfig(256, 0.9, 1, 0, 200000))
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(make_ext(schedule)), C.byref(h)))
    gm = GBDTModel(h, ctx)
    buf = np.zeros(64, np.float64)
    nr = C.c_int64()
    A.check(A.lib().bst_last_train_metric(ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
    return gm, om, buf[: nr.value].copy()


@pytest.mark.parametrize("K,schedule,kw", [
    (20, 1, dict(max_depth=4)),
    (20, 2, dict(max_depth=4)),          # graph requested: more than 16 trees run batched
    (37, 1, dict(growth=1, max_leaves=12)),
    (18, 0, dict(max_depth=3)),
])
def test_many_class_training_matches_oracle(ctx, oracle, K, schedule, kw):
    gm, om, hist = _train(ctx, oracle, K, schedule, **kw)
    ga = gm.export()
    assert gm.info.n_groups == K and gm.info.n_trees == 3 * K
    np.testing.assert_array_equal(ga["tree_group"], np.tile(np.arange(K), 3))
    assert forest_compare.compare(ga, om.arrays)["divergence"] is None
    np.testing.assert_allclose(hist, om.metric_history(), rtol=1e-10)


class _Dev:
    """A device buffer through the library's own allocation calls."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.p = C.c_void_p()
        A.check(A.lib().bst_device_alloc(ctx.h, nbytes, C.byref(self.p)))

    def put(self, arr):
        A.check(A.lib().bst_memcpy_h2d(self.ctx.h, self.p, A.ptr(arr), arr.nbytes))

    def get(self, arr):
        A.check(A.lib().bst_memcpy_d2h(self.ctx.h, A.ptr(arr), self.p, arr.nbytes))
        return arr

    def __del__(self):
        if self.p:
            A.lib().bst_device_free(self.ctx.h, self.p)


def test_many_class_prediction_windows(ctx, oracle):
    K = 37
    # Depth 6: 111 trees of up to 127 nodes span several LDS chunks, so few
    # rows take the two-phase launch.
    gm, om, _ = _train(ctx, oracle, K, 1, max_depth=6)
    Xt, _, _ = multiclass_cat_data(3000, 8, 2, K, card=12, seed=99)
    F = Xt.shape[0]
    for raw in (True, False):
        got, li = gm.predict_feature_major(Xt, raw=raw, leaf_index=True)
        want, wli = oracle.predict(om.info, om.arrays, Xt, raw=raw, leaf_idx=True)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)
        np.testing.assert_array_equal(li, wli)
    # Few rows (the two-phase launch: per-tree leaf values summed per window)
    # and many rows (one phase), through bst_predict_device and the
    # allocation-free entry point: bit-identical, and the oracle's values.
    T = gm.info.n_trees
    for n in (200, 3000):
        X = np.ascontiguousarray(Xt[:, :n])
        dX, da, db = _Dev(ctx, X.nbytes), _Dev(ctx, K * n * 4), _Dev(ctx, K * n * 4)
        la, lb = _Dev(ctx, T * n * 4), _Dev(ctx, T * n * 4)
        dX.put(X)
        L = A.lib()
        A.check(L.bst_predict_device(ctx.h, gm.h, dX.p, n, F, 0, da.p, la.p))
        A.check(L.bst_predict_device_noalloc(ctx.h, gm.h, dX.p, n, F, 0, db.p, lb.p))
        A.check(L.bst_ctx_sync(ctx.h))
        oa = da.get(np.zeros((K, n), np.float32))
        ob = db.get(np.zeros((K, n), np.float32))
        np.testing.assert_array_equal(oa, ob)
        np.testing.assert_array_equal(la.get(np.zeros((T, n), np.uint32)),
                                      lb.get(np.zeros((T, n), np.uint32)))
        np.testing.assert_allclose(oa, oracle.predict(om.info, om.arrays, X), rtol=1e-6,
                                   atol=1e-6)
    back = GBDTModel.from_bytes(gm.to_bytes(), ctx)
    np.testing.assert_array_equal(back.predict_feature_major(Xt, raw=True),
                                  gm.predict_feature_major(Xt, raw=True))
