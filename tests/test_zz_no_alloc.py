"""GPU: the allocation-free predict entry point (bst_predict_device_noalloc)
against bst_predict_device, through the C ABI.

Both run the same kernels, so outputs and leaf indices must be identical bit
for bit.  Two calls in a row on the same device buffers, the second with
fewer rows, check that nothing carried over in the shared static scratch
changes a result; a forest with many small trees makes the scratch hold fewer
rows than the call (row batches); a large call takes the one-phase launch.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import GBDTModel
from tests.conftest import make_cfg, make_ext, multiclass_cat_data

pytestmark = pytest.mark.gpu

F32 = np.float32


def _random_forest(n_trees, depth, F, seed, K=1):
    rng = np.random.default_rng(seed)
    nn = (1 << (depth + 1)) - 1
    n_inner = (1 << depth) - 1
    info = A.ModelInfo()
    info.n_trees, info.n_groups, info.n_features = n_trees, K, F
    info.total_nodes, info.total_cat_words = n_trees * nn, 0
    info.output_transform = A.TRANSFORM_IDENTITY if K == 1 else A.TRANSFORM_SOFTMAX
    info.objective = A.OBJ_SQUARED if K == 1 else A.OBJ_SOFTMAX
    a = A.alloc_forest(info)
    ids = np.arange(nn)
    for t in range(n_trees):
        s = slice(t * nn, (t + 1) * nn)
        leaf = ids >= n_inner
        a["is_leaf"][s] = leaf
        a["split_index"][s] = np.where(leaf, 0, rng.integers(0, F, nn))
        a["split_threshold"][s] = np.where(leaf, 0, rng.uniform(-2, 2, nn)).astype(F32)
        a["default_left"][s] = rng.integers(0, 2, nn)
        a["left"][s] = np.where(leaf, 0, 2 * ids + 1)
        a["right"][s] = np.where(leaf, 0, 2 * ids + 2)
        a["leaf_value"][s] = np.where(leaf, rng.uniform(-1, 1, nn), 0).astype(F32)
    a["tree_offset"][:] = np.arange(n_trees + 1) * nn
    a["tree_group"][:] = np.arange(n_trees) % K
    a["base_score"][:] = rng.uniform(-1, 1, K).astype(F32)
    return info, a


class DevArray:
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


def _both(ctx, gm, X, F, raw, dX, dout_a, dleaf_a, dout_b, dleaf_b):
    n = X.shape[1]
    K, T = gm.info.n_groups, gm.info.n_trees
    dX.put(np.ascontiguousarray(X))
    L = A.lib()
    A.check(L.bst_predict_device(ctx.h, gm.h, dX.p, n, F, raw, dout_a.p, dleaf_a.p))
    A.check(L.bst_predict_device_noalloc(ctx.h, gm.h, dX.p, n, F, raw, dout_b.p, dleaf_b.p))
    A.check(L.bst_ctx_sync(ctx.h))
    oa = dout_a.get(np.zeros((K, n), F32))
    ob = dout_b.get(np.zeros((K, n), F32))
    la = dleaf_a.get(np.zeros((T, n), np.uint32))
    lb = dleaf_b.get(np.zeros((T, n), np.uint32))
    np.testing.assert_array_equal(ob, oa)
    np.testing.assert_array_equal(lb, la)
    return oa, la


@pytest.mark.parametrize("shape", ["deep_regression", "many_small_trees", "softmax"])
def test_noalloc_equals_predict_device_twice(ctx, oracle, shape):
    if shape == "deep_regression":
        info, arr = _random_forest(300, 7, 40, seed=1)
        gm = GBDTModel.from_arrays(info, arr, ctx)
        F, sizes = 40, (5000, 1300)
    elif shape == "many_small_trees":
        # 3000 trees: the scratch holds 8388608 / 3000 = 2796 rows per batch.
        info, arr = _random_forest(3000, 3, 16, seed=2)
        gm = GBDTModel.from_arrays(info, arr, ctx)
        F, sizes = 16, (7000, 2900)
    else:
        X, y, is_cat = multiclass_cat_data(8000, 10, 3, 4, card=12, seed=71)
        from boosters_amd.model import BinnedDataset, Dataset
        b = BinnedDataset(ctx, Dataset.from_feature_major(X, categorical_features=list(np.nonzero(is_cat)[0])),
                          A.BinningConfig(256, 0.9, 1, 0, 200000))
        h = C.c_void_p()
        cfg = make_cfg(objective=2, n_classes=4, n_trees=40, growth=1, max_leaves=20)
        A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0,
                                  C.byref(cfg), C.byref(make_ext(2)), C.byref(h)))
        gm = GBDTModel(h, ctx)
        info, arr = gm.info, gm.export()
        F, sizes = 13, (6000, 999)
    K, T = gm.info.n_groups, gm.info.n_trees
    rng = np.random.default_rng(3)
    nmax = max(sizes)
    # Device buffers sized for the first (larger) call; the model is uploaded
    # by the first bst_predict_device inside _both.
    dX = DevArray(ctx, F * nmax * 4)
    dout_a, dout_b = DevArray(ctx, K * nmax * 4), DevArray(ctx, K * nmax * 4)
    dleaf_a, dleaf_b = DevArray(ctx, T * nmax * 4), DevArray(ctx, T * nmax * 4)
    for n in sizes:
        if shape == "softmax":
            X, _, _ = multiclass_cat_data(n, 10, 3, 4, card=12, seed=int(n))
        else:
            X = rng.standard_normal((F, n), dtype=F32)
            X[0, ::11] = np.nan
        for raw in (1, 0):
            out, li = _both(ctx, gm, X, F, raw, dX, dout_a, dleaf_a, dout_b, dleaf_b)
        # And both equal the oracle's traversal (leaf ids exact, outputs 1e-6).
        want, wli = oracle.predict(info, arr, X, raw=False, leaf_idx=True)
        np.testing.assert_array_equal(li, wli)
        np.testing.assert_allclose(out, want, rtol=1e-6, atol=1e-7)


def test_noalloc_large_call_and_errors(ctx):
    info, arr = _random_forest(200, 6, 24, seed=4)
    gm = GBDTModel.from_arrays(info, arr, ctx)
    n, F = 200_000, 24
    dX = DevArray(ctx, F * n * 4)
    dout = DevArray(ctx, n * 4)
    L = A.lib()
    # Not resident yet: refused (uploading would allocate).
    code = L.bst_predict_device_noalloc(ctx.h, gm.h, dX.p, n, F, 1, dout.p, None)
    assert code == A.BST_E_INVALID
    assert b"not resident" in L.bst_last_error()
    dout_a, dleaf_a = DevArray(ctx, n * 4), DevArray(ctx, 200 * n * 4)
    dleaf_b = DevArray(ctx, 200 * n * 4)
    X = np.random.default_rng(5).standard_normal((F, n), dtype=F32)
    _both(ctx, gm, X, F, 1, dX, dout_a, dleaf_a, dout, dleaf_b)
    # Fewer features than the model's: a shape error, nothing launched.
    assert L.bst_predict_device_noalloc(ctx.h, gm.h, dX.p, n, F - 1, 1, dout.p, None) == A.BST_E_SHAPE
