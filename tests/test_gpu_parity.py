"""GPU parity: every hot-path kernel and the full trainer against the oracle.

Tolerances (the bar): bins, partitions, tree structure, thresholds and leaf
indices bit-exact; histograms within 1e-9 of the f64 bin mass (the engine's
64-bit fixed-point sums are exact up to one quantum per row, far inside the
1e-5 relative budget); gradients bit-exact against the oracle in f64-exp mode
(the engine's device exp), within 2 ulp against glibc expf; leaf values and
predictions within 1e-6 relative.
"""
import ctypes as C

import numpy as np
import pytest

from boosters_amd import _abi as A
from boosters_amd.model import BinnedDataset, Dataset, GBDTModel
from tests.conftest import (binary_nan_data, make_cfg, make_ext, multiclass_cat_data,
                            ranking_data, regression_data)

pytestmark = pytest.mark.gpu


def _binned(ctx, X, is_cat=None, binning=None):
    ds = Dataset.from_feature_major(X, categorical_features=None if is_cat is None else list(np.nonzero(is_cat)[0]))
    return BinnedDataset(ctx, ds, binning or A.BinningConfig(256, 0.9, 1, 0, 200000))


def _edge_matrix(seed=0):
    rng = np.random.default_rng(seed)
    n = 3000
    cols = [
        rng.standard_normal(n).astype(np.float32),
        np.where(rng.random(n) < 0.2, np.nan, rng.standard_normal(n)).astype(np.float32),
        np.full(n, 3.5, np.float32),                       # constant non-zero: 1 bin
        np.zeros(n, np.float32),                           # trivial
        np.full(n, np.nan, np.float32),                    # trivial (all NaN)
        rng.integers(0, 5, n).astype(np.float32),          # few uniques
        np.where(rng.random(n) < 0.5, 0.0, 1.0).astype(np.float32),  # binary
        rng.standard_normal(n).astype(np.float32) * 1e6,
        rng.integers(0, 7, n).astype(np.float32),          # categorical
    ]
    cols[0][:5] = [np.inf, -np.inf, 0.0, -0.0, 1e-30]
    X = np.stack(cols)
    is_cat = np.zeros(len(cols), np.uint8)
    is_cat[-1] = 1
    return X, is_cat


def test_binning_bit_exact(ctx, oracle):
    X, is_cat = _edge_matrix()
    b = _binned(ctx, X, is_cat)
    o = oracle.Binned(X, is_cat)
    np.testing.assert_array_equal(b.bins(), o.bins())
    for f in range(X.shape[0]):
        assert b.feature_info(f) == o.feature_info(f), f
        np.testing.assert_array_equal(b.bounds(f), o.bounds(f))


def test_binning_config2_shape_sampled_edges(ctx, oracle):
    # More rows than the 200k edge sample: exercises the splitmix64 sample and
    # the 257-bin (NaN) hi plane.
    X, _ = binary_nan_data(210_000, 6, seed=3)
    b = _binned(ctx, X)
    o = oracle.Binned(X)
    np.testing.assert_array_equal(b.bins(), o.bins())
    assert b.feature_info(0)["n_bins"] == 257


@pytest.mark.parametrize("obj,K", [(0, 1), (1, 1), (2, 4)])
def test_gradients(ctx, oracle, obj, K):
    rng = np.random.default_rng(obj)
    n = 5000
    preds = (rng.standard_normal((K, n)) * 3).astype(np.float32)
    y = (rng.integers(0, K, n) if obj == 2 else (rng.random(n) < 0.5)).astype(np.float32)
    if obj == 0:
        y = rng.standard_normal(n).astype(np.float32)
    w = rng.random(n).astype(np.float32) + 0.5
    g = np.zeros((K, n), np.float32)
    h = np.zeros((K, n), np.float32)
    A.check(A.lib().bst_kernel_gradients(ctx.h, obj, K, A.ptr(preds), A.ptr(y), A.ptr(w), n, A.ptr(g), A.ptr(h)))
    oracle.set_exp_mode(1)
    og, oh = oracle.gradients(obj, preds, y, w)
    np.testing.assert_array_equal(g, og)
    np.testing.assert_array_equal(h, oh)
    oracle.set_exp_mode(0)
    rg, rh = oracle.gradients(obj, preds, y, w)
    ulp = np.abs(g.view(np.int32).astype(np.int64) - rg.view(np.int32).astype(np.int64))
    # glibc expf is within 1 ulp of the correctly rounded exp; softmax combines
    # K exps, a sum and a division, so the gap can reach a few ulp.
    assert ulp.max() <= (2 if obj != 2 else 6)


def test_lambdarank_gradients(ctx, oracle):
    rng = np.random.default_rng(21)
    # 512 / 513: the two sort paths either side of their switch; 4096: the
    # largest group; 7: a group whose labels are all equal (maxDCG = 0).
    sizes = np.concatenate([[1, 2, 600, 31, 30, 29, 512, 513, 4096, 7],
                            rng.integers(1, 200, 80)]).astype(np.int32)
    n = int(sizes.sum())
    y = rng.integers(0, 5, n).astype(np.float32)
    z = int(sizes[:9].sum())
    y[z:z + 7] = 2.0
    p = (rng.standard_normal(n) * 2).astype(np.float32)
    p[::7] = p[1::7][: len(p[::7])]  # ties inside groups
    p[5] = -0.0
    p[6] = 0.0
    w = (rng.random(n) + 0.5).astype(np.float32)
    for weights in (None, w):
        g = np.zeros(n, np.float32)
        h = np.zeros(n, np.float32)
        A.check(A.lib().bst_kernel_lambdarank(ctx.h, A.ptr(p), A.ptr(y), A.ptr(weights), n,
                                              A.ptr(sizes), len(sizes), A.ptr(g), A.ptr(h)))
        oracle.set_exp_mode(1)
        og, oh = oracle.lambdarank_gradients(p, y, sizes, weights)
        np.testing.assert_array_equal(g, og)
        np.testing.assert_array_equal(h, oh)


def test_histogram_vs_f64(ctx, oracle):
    X, _ = binary_nan_data(40_000, 40, seed=5)
    b = _binned(ctx, X)
    o = oracle.Binned(X)
    rng = np.random.default_rng(1)
    grad = rng.standard_normal(40_000).astype(np.float32)
    hess = (rng.random(40_000) * 0.25 + 1e-6).astype(np.float32)
    rows = np.sort(rng.choice(40_000, 17_000, replace=False)).astype(np.uint32)
    lay = b.layout()
    for r in (None, rows):
        out = np.zeros(lay["total_bins"] * 2, np.float64)
        A.check(A.lib().bst_kernel_histogram(ctx.h, b.h, A.ptr(grad), A.ptr(hess),
                                             A.ptr(r), 0 if r is None else len(r), A.ptr(out), len(out)))
        ref = o.histogram(grad, hess, r).reshape(-1)
        sel = np.arange(40_000) if r is None else r
        mass = np.abs(grad[sel]).sum() + np.abs(hess[sel]).sum()
        assert np.max(np.abs(out - ref)) <= 1e-9 * mass


def test_known_answer_split(ctx, oracle):
    # find.rs test_find_split_multiple_features: feature 1 is the strongest.
    X = np.stack([np.repeat(np.arange(3, dtype=np.float32), 10)] * 3)
    b = _binned(ctx, X)
    hist = np.array([(1, 5), (0, 5), (-1, 5), (10, 5), (0, 5), (-10, 5), (3, 5), (0, 5), (-3, 5)], np.float64)
    cfg = make_cfg()
    out = A.SplitInfoC()
    A.check(A.lib().bst_kernel_find_split(ctx.h, b.h, A.ptr(hist.reshape(-1)), 0.0, 45.0, 45,
                                          C.byref(cfg), 0, C.byref(out)))
    assert out.feature == 1 and out.gain > 0
    ref = oracle.Binned(X).find_split(hist, 0.0, 45.0, 45, cfg)
    assert (out.feature, out.bin, out.default_left) == (ref.feature, ref.bin, ref.default_left)
    assert out.gain == ref.gain


def test_random_splits_match_oracle(ctx, oracle):
    X, y, is_cat = multiclass_cat_data(6000, 10, 3, 3, card=12, seed=2)
    Xn, _ = binary_nan_data(6000, 6, seed=4)
    X = np.concatenate([X, Xn])
    is_cat = np.concatenate([is_cat, np.zeros(6, np.uint8)])
    b = _binned(ctx, X, is_cat)
    o = oracle.Binned(X, is_cat)
    rng = np.random.default_rng(9)
    cfg = make_cfg(max_onehot_cats=4)
    for trial in range(20):
        grad = rng.standard_normal(6000).astype(np.float32)
        hess = (rng.random(6000) + 0.1).astype(np.float32)
        rows = np.sort(rng.choice(6000, rng.integers(50, 6000), replace=False)).astype(np.uint32)
        hist = o.histogram(grad, hess, rows)
        pg = float(np.sum(grad[rows].astype(np.float64)))
        ph = float(np.sum(hess[rows].astype(np.float64)))
        for nt in (0, 1):
            ref = o.find_split(hist, pg, ph, len(rows), cfg, n_threads=nt)
            out = A.SplitInfoC()
            A.check(A.lib().bst_kernel_find_split(ctx.h, b.h, A.ptr(hist.reshape(-1)), pg, ph, len(rows),
                                                  C.byref(cfg), nt, C.byref(out)))
            assert out.feature == ref.feature, trial
            assert out.is_categorical == ref.is_categorical
            assert out.default_left == ref.default_left
            if out.is_categorical:
                assert list(out.left_cats) == list(ref.left_cats)
            else:
                assert out.bin == ref.bin
            assert abs(out.gain - ref.gain) <= 1e-6 * max(1.0, abs(ref.gain))


def test_partition_bit_exact(ctx, oracle):
    X, y, is_cat = multiclass_cat_data(20_000, 4, 2, 3, card=9, seed=7)
    Xn, _ = binary_nan_data(20_000, 3, seed=8, nan_p=0.3)
    X = np.concatenate([X, Xn])
    is_cat = np.concatenate([is_cat, np.zeros(3, np.uint8)])
    b = _binned(ctx, X, is_cat)
    o = oracle.Binned(X, is_cat)
    lay = b.layout()
    rng = np.random.default_rng(3)
    for trial in range(24):
        e = int(rng.integers(0, lay["n_eff"]))
        rows = np.sort(rng.choice(20_000, int(rng.integers(1, 20_000)), replace=False)).astype(np.uint32)
        sp = A.SplitInfoC()
        sp.feature = e
        sp.gain = 1.0
        sp.default_left = int(rng.integers(0, 2))
        info = b.feature_info(int(lay["eff_to_orig"][e]))
        sp.is_categorical = int(info["categorical"])
        if sp.is_categorical:
            cats = rng.integers(0, 2, 8 * 32).astype(bool)
            for k in range(256):
                if cats[k]:
                    sp.left_cats[k >> 5] |= 1 << (k & 31)
            cw = np.array(list(sp.left_cats), np.uint32)
        else:
            sp.bin = int(rng.integers(0, info["n_bins"]))
            cw = None
        ref_rows, ref_left = o.partition(rows, sp)
        r2 = rows.copy()
        lc = C.c_int64()
        A.check(A.lib().bst_kernel_partition(ctx.h, b.h, A.ptr(r2), len(r2), e, sp.bin, sp.default_left,
                                             A.ptr(cw), 8 if cw is not None else 0, C.byref(lc)))
        assert lc.value == ref_left
        np.testing.assert_array_equal(r2, ref_rows)


def test_fused_partition_stable(ctx, oracle):
    """The graph schedule's one-pass partition (k_part_fused) is the oracle's
    stable partition (partition.rs:73-145, 296-382): equal left count, and
    both sides in row order, with the node's rows given in order or stored
    the way the kernel stores a right child (last position first), and with
    tiles taken by block index or claimed through the ticket counter.  Nodes
    of zero to five 4096-row tiles (the tiles' look-back chain), numeric
    splits with NaN in either direction, categorical splits."""
    X, y, is_cat = multiclass_cat_data(20_000, 4, 2, 3, card=9, seed=17)
    Xn, _ = binary_nan_data(20_000, 3, seed=18, nan_p=0.3)
    X = np.concatenate([X, Xn])
    is_cat = np.concatenate([is_cat, np.zeros(3, np.uint8)])
    b = _binned(ctx, X, is_cat)
    o = oracle.Binned(X, is_cat)
    lay = b.layout()
    rng = np.random.default_rng(19)
    sizes = [0, 1, 5, 4095, 4096, 4097, 12_288, 20_000]
    for trial in range(32):
        e = int(rng.integers(0, lay["n_eff"]))
        n_rows = sizes[trial % len(sizes)] if trial < 2 * len(sizes) else int(rng.integers(1, 20_000))
        rows = rng.choice(20_000, n_rows, replace=False).astype(np.uint32)
        if trial % 3 == 0:
            rows = np.sort(rows)
        sp = A.SplitInfoC()
        sp.feature = e
        sp.gain = 1.0
        sp.default_left = int(rng.integers(0, 2))
        info = b.feature_info(int(lay["eff_to_orig"][e]))
        sp.is_categorical = int(info["categorical"])
        if sp.is_categorical:
            cats = rng.integers(0, 2, 8 * 32).astype(bool)
            for k in range(256):
                if cats[k]:
                    sp.left_cats[k >> 5] |= 1 << (k & 31)
            cw = np.array(list(sp.left_cats), np.uint32)
        else:
            sp.bin = int(rng.integers(0, info["n_bins"]))
            cw = None
        ref_rows, ref_left = o.partition(rows, sp)
        for mode in (0, 1, 2, 3):  # bit 0 reversed storage, bit 1 ticket claims
            r2 = rows.copy()
            lc = C.c_int64()
            A.check(A.lib().bst_kernel_partition_fused(ctx.h, b.h, A.ptr(r2), len(r2), e, sp.bin,
                                                       sp.default_left, A.ptr(cw),
                                                       8 if cw is not None else 0, mode,
                                                       C.byref(lc)))
            assert lc.value == ref_left, (trial, e, n_rows, mode)
            np.testing.assert_array_equal(r2, ref_rows, err_msg=f"trial {trial} mode {mode}")


def _forests_equal(gm, om_info, om_arrays, leaf_rtol=1e-6):
    ga = gm.export()
    assert gm.info.n_trees == om_info.n_trees
    assert gm.info.total_nodes == om_info.total_nodes
    for k in ("tree_offset", "tree_group", "split_index", "left", "right", "default_left", "is_leaf",
              "split_type", "cat_start", "cat_size"):
        np.testing.assert_array_equal(ga[k], om_arrays[k], err_msg=k)
    np.testing.assert_array_equal(ga["split_threshold"], om_arrays["split_threshold"])
    if om_info.total_cat_words:
        np.testing.assert_array_equal(ga["cat_words"][: om_info.total_cat_words],
                                      om_arrays["cat_words"][: om_info.total_cat_words])
    np.testing.assert_allclose(ga["leaf_value"], om_arrays["leaf_value"], rtol=leaf_rtol, atol=1e-12)
    np.testing.assert_array_equal(ga["base_score"], om_arrays["base_score"])
    return ga


def _train_both(ctx, oracle, X, y, cfg, is_cat=None, schedule=1, n_threads=0):
    oracle.set_exp_mode(1)
    o = oracle.Binned(X, is_cat)
    om = oracle.train(o, y, cfg, make_ext(0, n_threads))
    b = _binned(ctx, X, is_cat)
    h = C.c_void_p()
    ext = make_ext(schedule, n_threads)
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(ext), C.byref(h)))
    gm = GBDTModel(h, ctx)
    return gm, om, b


def test_train_config1_identical_trees(ctx, oracle):
    X, y = regression_data(5000, 100, seed=0)
    cfg = make_cfg(n_trees=20, learning_rate=0.1, max_depth=6)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg)
    _forests_equal(gm, om.info, om.arrays)


def test_train_binary_nan_depth8(ctx, oracle):
    X, y = binary_nan_data(30_000, 24, seed=1)
    cfg = make_cfg(objective=1, n_trees=12, learning_rate=0.1, max_depth=8)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg)
    _forests_equal(gm, om.info, om.arrays)


def test_train_softmax_leafwise_categorical(ctx, oracle):
    X, y, is_cat = multiclass_cat_data(12_000, 12, 3, 4, card=16, seed=4)
    cfg = make_cfg(objective=2, n_classes=4, n_trees=5, learning_rate=0.1, growth=1, max_leaves=31)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg, is_cat)
    _forests_equal(gm, om.info, om.arrays)


@pytest.mark.parametrize("n_threads", [0, 1])
def test_schedules_identical(ctx, oracle, n_threads):
    X, y = binary_nan_data(8000, 20, seed=6)
    cfg = make_cfg(objective=1, n_trees=4, max_depth=5)
    gm_p, om, b = _train_both(ctx, oracle, X, y, cfg, schedule=0, n_threads=n_threads)
    ga = _forests_equal(gm_p, om.info, om.arrays)
    h = C.c_void_p()
    ext = make_ext(1, n_threads)
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(ext), C.byref(h)))
    gb = GBDTModel(h, ctx).export()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


GRAPH_CASES = {
    # name: (data builder, config kwargs)
    "squared_depth6": (lambda: regression_data(5000, 100, seed=0) + (None,),
                       dict(n_trees=8, learning_rate=0.1, max_depth=6)),
    "logistic_nan_depth8": (lambda: binary_nan_data(30_000, 24, seed=1) + (None,),
                            dict(objective=1, n_trees=6, learning_rate=0.1, max_depth=8)),
    "squared_leafwise": (lambda: regression_data(20_000, 30, seed=5) + (None,),
                         dict(n_trees=6, growth=1, max_leaves=40, learning_rate=0.2)),
    "softmax_leafwise_cat": (lambda: multiclass_cat_data(12_000, 12, 3, 4, card=16, seed=4),
                             dict(objective=2, n_classes=4, n_trees=4, learning_rate=0.1, growth=1,
                                  max_leaves=31)),
    "softmax_depthwise_cat": (lambda: multiclass_cat_data(9000, 8, 2, 3, card=40, seed=13),
                              dict(objective=2, n_classes=3, n_trees=4, max_depth=5,
                                   max_onehot_cats=4)),
    "min_gain_stops_early": (lambda: regression_data(3000, 10, seed=9) + (None,),
                             dict(n_trees=5, max_depth=7, min_gain=5.0, min_samples_leaf=40)),
}


@pytest.mark.parametrize("case", sorted(GRAPH_CASES))
def test_graph_schedule_matches_oracle(ctx, oracle, case):
    """Device-side growth replayed as a hipGraph grows the oracle's forest."""
    build, kw = GRAPH_CASES[case]
    X, y, is_cat = build()
    cfg = make_cfg(**kw)
    gm, om, b = _train_both(ctx, oracle, X, y, cfg, is_cat, schedule=2)
    ga = _forests_equal(gm, om.info, om.arrays)
    # Same forest as the host-driven batched schedule, leaf values included.
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(make_ext(1)), C.byref(h)))
    gb = GBDTModel(h, ctx).export()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


# (features per block, fewest rows per block, blocks per CU, sub-histograms:
# 1 in LDS, 2 in global memory)
HIST_SHAPES = [(16, 512, 4, 1), (32, 2048, 1, 1), (16, 8192, 1, 1), (32, 512, 8, 1),
               (16, 1024, 2, 2), (32, 512, 4, 2)]


@pytest.mark.parametrize("case", ["logistic_nan_depth8", "softmax_leafwise_cat", "squared_leafwise"])
def test_hist_launch_shapes_identical(ctx, oracle, case):
    """Every histogram launch shape the autotuner can pick (16- or 32-feature
    groups, tile sizes, blocks per CU, LDS or global-memory sub-histograms)
    grows the oracle's forest: the sums are exact, so the shape changes time
    only."""
    build, kw = GRAPH_CASES[case]
    X, y, is_cat = build()
    cfg = make_cfg(**kw)
    gm, om, b = _train_both(ctx, oracle, X, y, cfg, is_cat, schedule=2)
    ga = _forests_equal(gm, om.info, om.arrays)
    for sched in (2, 1):
        for fg, rows_min, bpc, priv in HIST_SHAPES:
            ext = make_ext(sched)
            ext.autotune = 0
            ext.hist_fg, ext.hist_rows_min, ext.hist_blocks_per_cu = fg, rows_min, bpc
            ext.hist_priv = priv
            h = C.c_void_p()
            A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0,
                                      C.byref(cfg), C.byref(ext), C.byref(h)))
            gb = GBDTModel(h, ctx).export()
            for k in ga:
                np.testing.assert_array_equal(ga[k], gb[k],
                                              err_msg=f"{k} {sched} {fg} {rows_min} {bpc} {priv}")
            got = [C.c_int32() for _ in range(5)]
            A.check(A.lib().bst_last_train_tuning(ctx.h, *[C.byref(v) for v in got], None))
            assert [v.value for v in got] == [fg, rows_min, bpc, priv, 0]


@pytest.mark.parametrize("E,nan", [(113, True), (40, False), (40, True), (7, True)])
def test_padded_rows_feature_counts(ctx, oracle, E, nan):
    """Histogram lanes of padding features (past the last real feature of the
    last 16- or 32-feature group) read their own zero bytes of the padded
    row instead of testing their index.  Feature counts whose padding needs
    a longer row than the bins and NaN bits alone (113 with NaN bits: the
    16-feature groups' NaN bytes; 40: the 32-feature groups' bins) grow the
    oracle's forest with either group width."""
    X, y = binary_nan_data(20_000, E, seed=E) if nan else regression_data(20_000, E, seed=E)
    cfg = make_cfg(objective=1 if nan else 0, n_trees=3, max_depth=6)
    gm, om, b = _train_both(ctx, oracle, X, y, cfg, None, schedule=2)
    ga = _forests_equal(gm, om.info, om.arrays)
    for fg in (16, 32):
        ext = make_ext(2)
        ext.autotune = 0
        ext.hist_fg, ext.hist_rows_min, ext.hist_blocks_per_cu = fg, 512, 2
        h = C.c_void_p()
        A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                                  C.byref(ext), C.byref(h)))
        gb = GBDTModel(h, ctx).export()
        for k in ga:
            np.testing.assert_array_equal(ga[k], gb[k], err_msg=f"{k} fg {fg}")


def test_repeated_trains_reuse_buffers(ctx, oracle):
    """Training runs in one context reuse its device and pinned blocks (a pool
    keyed by size): a run after runs of another shape grows the oracle's
    forest, the same forest as the first run, so no state is carried from one
    run to the next through a reused block."""
    X, y = binary_nan_data(20_000, 24, seed=41)
    cfg = make_cfg(objective=1, n_trees=4, max_depth=6)
    gm, om, b = _train_both(ctx, oracle, X, y, cfg, None, schedule=2)
    ga = _forests_equal(gm, om.info, om.arrays)
    # Another shape in between: more rows, other features, leaf-wise.
    X2, y2 = regression_data(30_000, 17, seed=42)
    b2 = _binned(ctx, X2)
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b2.h, A.ptr(y2), len(y2), None, None, None, 0,
                              C.byref(make_cfg(n_trees=3, growth=1, max_leaves=20)),
                              C.byref(make_ext(2)), C.byref(h)))
    GBDTModel(h, ctx)
    for _ in range(2):
        h = C.c_void_p()
        A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                                  C.byref(make_ext(2)), C.byref(h)))
        gb = GBDTModel(h, ctx).export()
        for k in ga:
            np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_autotuner_choice_and_cache(ctx, oracle):
    """With autotune on, the shape is measured once per data shape and cached
    in the context; the forest is the oracle's whatever it picks."""
    X, y = binary_nan_data(40_000, 24, seed=31)
    cfg = make_cfg(objective=1, n_trees=3, max_depth=6)
    gm, om, b = _train_both(ctx, oracle, X, y, cfg, None, schedule=2)
    _forests_equal(gm, om.info, om.arrays)
    fg, rows_min, bpc, priv, tuned = (C.c_int32() for _ in range(5))
    ms = C.c_double()
    A.check(A.lib().bst_last_train_tuning(ctx.h, C.byref(fg), C.byref(rows_min), C.byref(bpc),
                                          C.byref(priv), C.byref(tuned), C.byref(ms)))
    assert tuned.value == 1 and fg.value in (16, 32) and rows_min.value in (512, 1024, 2048)
    assert bpc.value in (1, 2, 4) and priv.value in (1, 2) and ms.value > 0
    first = (fg.value, rows_min.value, bpc.value, priv.value)
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(make_ext(2)), C.byref(h)))
    GBDTModel(h, ctx)
    A.check(A.lib().bst_last_train_tuning(ctx.h, C.byref(fg), C.byref(rows_min), C.byref(bpc),
                                          C.byref(priv), C.byref(tuned), C.byref(ms)))
    assert (fg.value, rows_min.value, bpc.value, priv.value) == first and ms.value == 0.0


@pytest.mark.parametrize("case", ["logistic", "softmax_cat"])
def test_graph_schedule_training_predictions(ctx, oracle, case):
    """The in-graph prediction update (the last level's leaves applied by its
    partition) leaves the same per-round training loss, the last round's
    included; softmax grows three class trees per round."""
    if case == "logistic":
        X, y = binary_nan_data(20_000, 16, seed=23)
        is_cat = None
        cfg = make_cfg(objective=1, n_trees=7, learning_rate=0.3, max_depth=6)
    else:
        X, y, is_cat = multiclass_cat_data(9000, 8, 2, 3, card=40, seed=13)
        cfg = make_cfg(objective=2, n_classes=3, n_trees=7, max_depth=5, max_onehot_cats=4)
    b = _binned(ctx, X, is_cat)
    hist = {}
    for sched in (1, 2):
        h = C.c_void_p()
        A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                                  C.byref(make_ext(sched)), C.byref(h)))
        GBDTModel(h, ctx)
        buf = np.zeros(64, np.float64)
        nr = C.c_int64()
        A.check(A.lib().bst_last_train_metric(ctx.h, A.ptr(buf), len(buf), C.byref(nr)))
        hist[sched] = buf[: nr.value].copy()
    assert len(hist[2]) == 7
    np.testing.assert_array_equal(hist[1], hist[2])
    assert hist[2][-1] < hist[2][0]


@pytest.mark.parametrize("schedule", [1, 2])
def test_train_lambdarank(ctx, oracle, schedule):
    X, y, sizes = ranking_data(6000, 24, seed=3, group_lo=5, group_hi=80)
    cfg = make_cfg(objective=3, n_trees=5, learning_rate=0.3, growth=1, max_leaves=31)
    oracle.set_exp_mode(1)
    o = oracle.Binned(X, None)
    e0 = make_ext(0)
    e0.query_group_sizes = A.ptr(sizes)
    e0.n_query_groups = len(sizes)
    om = oracle.train(o, y, cfg, e0)
    b = _binned(ctx, X, None)
    e = make_ext(schedule)
    e.query_group_sizes = A.ptr(sizes)
    e.n_query_groups = len(sizes)
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(e), C.byref(h)))
    gm = GBDTModel(h, ctx)
    _forests_equal(gm, om.info, om.arrays)


def _train_both_ext(ctx, oracle, X, y, cfg, is_cat, schedule, goss=None, sizes=None):
    oracle.set_exp_mode(1)
    o = oracle.Binned(X, is_cat)
    exts = []
    for sched in (0, schedule):
        e = make_ext(sched)
        if goss:
            e.goss_top_rate, e.goss_other_rate = goss
        if sizes is not None:
            e.query_group_sizes = A.ptr(sizes)
            e.n_query_groups = len(sizes)
        exts.append(e)
    om = oracle.train(o, y, cfg, exts[0])
    b = _binned(ctx, X, is_cat)
    h = C.c_void_p()
    A.check(A.lib().bst_train(ctx.h, b.h, A.ptr(y), len(y), None, None, None, 0, C.byref(cfg),
                              C.byref(exts[1]), C.byref(h)))
    return GBDTModel(h, ctx), om


SAMPLING_CASES = {
    # name: (data builder, config kwargs, goss rates or None, schedules)
    "uniform_depthwise": (lambda: binary_nan_data(20_000, 16, seed=31) + (None,),
                          dict(objective=1, n_trees=5, max_depth=6, subsample=0.6), None, (1, 2)),
    "colsample_leafwise": (lambda: regression_data(15_000, 40, seed=32) + (None,),
                           dict(n_trees=5, growth=1, max_leaves=24, colsample_bytree=0.5),
                           None, (1, 2)),
    "softmax_colsample_subsample": (lambda: multiclass_cat_data(9000, 10, 2, 3, card=12, seed=33),
                                    dict(objective=2, n_classes=3, n_trees=3, max_depth=5,
                                         colsample_bytree=0.6, subsample=0.7), None, (1, 2)),
    "goss_logistic": (lambda: binary_nan_data(30_000, 20, seed=34) + (None,),
                      dict(objective=1, n_trees=6, max_depth=6, learning_rate=0.5), (0.2, 0.1),
                      (2,)),
}


@pytest.mark.parametrize("case", sorted(SAMPLING_CASES))
def test_sampling_matches_oracle(ctx, oracle, case):
    build, kw, goss, schedules = SAMPLING_CASES[case]
    X, y, is_cat = build()
    cfg = make_cfg(**kw)
    for sched in schedules:
        gm, om = _train_both_ext(ctx, oracle, X, y, cfg, is_cat, sched, goss=goss)
        _forests_equal(gm, om.info, om.arrays)


def test_goss_colsample_lambdarank(ctx, oracle):
    """The ranking workload's settings at a small size: GOSS 0.2/0.1, half the
    columns per tree, leaf-wise, LambdaRank gradients."""
    X, y, sizes = ranking_data(12_000, 30, seed=35, group_lo=5, group_hi=120)
    cfg = make_cfg(objective=3, n_trees=6, learning_rate=0.3, growth=1, max_leaves=31,
                   colsample_bytree=0.5)
    gm, om = _train_both_ext(ctx, oracle, X, y, cfg, None, 2, goss=(0.2, 0.1), sizes=sizes)
    _forests_equal(gm, om.info, om.arrays)


def test_predict_and_leaf_indices(ctx, oracle):
    X, y, is_cat = multiclass_cat_data(9000, 8, 2, 3, card=10, seed=11)
    cfg = make_cfg(objective=2, n_classes=3, n_trees=6, growth=1, max_leaves=15)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg, is_cat)
    Xt, _, _ = multiclass_cat_data(4000, 8, 2, 3, card=10, seed=12)
    Xt[0, :50] = np.nan
    for raw in (True, False):
        out, li = gm.predict_feature_major(Xt, raw=raw, leaf_index=True)
        ref, rli = oracle.predict(om.info, om.arrays, Xt, raw=raw, leaf_idx=True)
        np.testing.assert_array_equal(li, rli)
        np.testing.assert_allclose(out, ref, rtol=1e-6, atol=1e-7)


def _mixed_forest(seed):
    """Complete trees of depths 3, 5 and 8, two of depth 13 (16383 nodes: too
    large for an LDS chunk, traversed in the 16-byte layout from global
    memory) and caterpillar trees (one leaf per level, depth 20), so that one
    group of chains holds trees of different depths and leaves at every
    depth."""
    rng = np.random.default_rng(seed)
    F = 12
    trees = []  # (is_leaf, left, right) per tree, original ids
    for kind in [3, 13, 5, "cat", 8, 13] + [8] * 16 + ["cat", 5] * 5:
        if kind == "cat":
            D = 20
            nn = 2 * D + 1
            ids = np.arange(nn)
            leaf = (ids % 2 == 1) | (ids == 2 * D)
            left = np.where(leaf, 0, ids + 1)
            right = np.where(leaf, 0, ids + 2)
        else:
            nn = (1 << (kind + 1)) - 1
            ids = np.arange(nn)
            leaf = ids >= (1 << kind) - 1
            left = np.where(leaf, 0, 2 * ids + 1)
            right = np.where(leaf, 0, 2 * ids + 2)
        trees.append((leaf, left, right))
    info = A.ModelInfo()
    info.n_trees, info.n_groups, info.n_features = len(trees), 1, F
    info.total_nodes = sum(len(t[0]) for t in trees)
    info.total_cat_words = 0
    info.output_transform = A.TRANSFORM_IDENTITY
    info.objective = A.OBJ_SQUARED
    a = A.alloc_forest(info)
    off = 0
    for i, (leaf, left, right) in enumerate(trees):
        nn = len(leaf)
        s = slice(off, off + nn)
        a["is_leaf"][s] = leaf
        a["split_index"][s] = np.where(leaf, 0, rng.integers(0, F, nn))
        # A third of the thresholds exactly 0 (x = -0.0 and +0.0 both go right).
        thr = np.where(rng.random(nn) < 0.3, 0.0, rng.uniform(-2, 2, nn))
        a["split_threshold"][s] = np.where(leaf, 0, thr).astype(np.float32)
        a["default_left"][s] = rng.integers(0, 2, nn)
        a["left"][s] = left
        a["right"][s] = right
        a["leaf_value"][s] = np.where(leaf, rng.uniform(-1, 1, nn), 0).astype(np.float32)
        a["tree_offset"][i] = off
        off += nn
    a["tree_offset"][len(trees)] = off
    a["tree_group"][:] = 0
    a["base_score"][:] = 0.25
    return info, a


def test_predict_mixed_tree_layouts(ctx, oracle):
    """LDS-chunk trees and global-memory trees in one forest, trees of
    different depths in one group of chains: outputs and leaf indices equal
    the oracle's, NaN values included."""
    info, arr = _mixed_forest(7)
    gm = GBDTModel.from_arrays(info, arr, ctx)
    rng = np.random.default_rng(8)
    X = rng.uniform(-2.5, 2.5, (info.n_features, 3000)).astype(np.float32)
    X[rng.random(X.shape) < 0.1] = np.nan
    X[:, :40] = 0.0
    X[:, 40:80] = -0.0
    out, li = gm.predict_feature_major(X, raw=True, leaf_index=True)
    ref, rli = oracle.predict(info, arr, X, raw=True, leaf_idx=True)
    np.testing.assert_array_equal(li, rli)
    np.testing.assert_allclose(out, ref, rtol=1e-6, atol=1e-7)


def test_bstr_round_trip_predictions(ctx, oracle):
    """A trained forest written as .bstr (MessagePack, zstd) and .bstr.json and
    read back (persist/mod.rs:88-143) has the same arrays and predicts the
    same values and leaves on the device."""
    X, y, is_cat = multiclass_cat_data(6000, 8, 2, 3, card=10, seed=13)
    cfg = make_cfg(objective=2, n_classes=3, n_trees=4, growth=1, max_leaves=15)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg, is_cat)
    Xt, _, _ = multiclass_cat_data(3000, 8, 2, 3, card=10, seed=14)
    Xt[1, :40] = np.nan
    want, wli = gm.predict_feature_major(Xt, raw=False, leaf_index=True)
    ga = gm.export()
    assert ga["split_type"].any()
    for m2 in (GBDTModel.from_bytes(gm.to_bytes(0), ctx), GBDTModel.from_bytes(gm.to_bytes(3), ctx),
               GBDTModel.from_json_bytes(gm.to_json_bytes(), ctx)):
        ea = m2.export()
        for k in ("split_index", "split_threshold", "left", "right", "default_left", "is_leaf",
                  "split_type", "leaf_value", "tree_group", "base_score"):
            np.testing.assert_array_equal(ea[k], ga[k], err_msg=k)
        got, li = m2.predict_feature_major(Xt, raw=False, leaf_index=True)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(li, wli)


def test_validation_loss_matches(ctx, oracle):
    X, y = binary_nan_data(20_000, 16, seed=21)
    cfg = make_cfg(objective=1, n_trees=10, learning_rate=0.1, max_depth=6)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg)
    gmh = gm.export()
    Xv, yv = binary_nan_data(5000, 16, seed=22)
    pg = gm.predict_feature_major(Xv)
    po = oracle.predict(om.info, om.arrays, Xv)
    lg = oracle.metric(A.METRIC_LOGLOSS, pg, yv)
    lo = oracle.metric(A.METRIC_LOGLOSS, po, yv)
    assert abs(lg - lo) <= 1e-4
    assert gmh["split_index"].shape[0] > 10


def test_server_alone_equals_mixed(ctx, oracle):
    from boosters_amd.model import Server

    X, y = regression_data(6000, 12, seed=2)
    cfg = make_cfg(n_trees=30, max_depth=6)
    gm, om, _ = _train_both(ctx, oracle, X, y, cfg)
    rng = np.random.default_rng(0)
    reqs = [rng.standard_normal((12, int(rng.integers(1, 3000)))).astype(np.float32) for _ in range(16)]
    alone = [gm.predict_feature_major(r, raw=True) for r in reqs]
    srv = Server(gm, hbm_budget_bytes=1 << 26, max_inflight=3)
    tickets = [srv.submit(r) for r in reqs]
    outs = [srv.wait(t)[0] for t in tickets]
    srv.close()
    for a, b in zip(alone, outs):
        np.testing.assert_array_equal(a, b)
