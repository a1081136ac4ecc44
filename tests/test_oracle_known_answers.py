"""Pin the CPU oracle to the reference's own known-answer tests.

Each test cites the reference test it transcribes (paths relative to
crates/boosters/src/).  The oracle is then used as the checker of the HIP
engine, so a wrong oracle would hide a wrong kernel: these run on the CPU.
"""
import math

import numpy as np
import pytest

from boosters_amd import _abi as A
from tests.conftest import make_cfg, make_ext

F32 = np.float32


def col(*vals):
    return np.array([vals], dtype=F32)


# ---------------------------------------------------------------- binning ---
def test_numeric_value_to_bin(oracle):
    # bin_mapper.rs:366-378 (bounds [1,2,3,4]); data {0.5,1.5,2.5,3.5} gives
    # midpoint bounds [1,2,3,MAX] which agree on every probed value.
    b = oracle.Binned(col(0.5, 1.5, 2.5, 3.5))
    np.testing.assert_array_equal(b.bounds(0)[:3], [1.0, 2.0, 3.0])
    for v, e in [(0.5, 0), (1.0, 0), (1.5, 1), (2.0, 1), (2.5, 2), (3.0, 2), (3.5, 3)]:
        assert b.value_to_bin(0, v) == e, v


def test_nan_bin_is_last(oracle):
    # bin_mapper.rs:406-422: NaN goes to the dedicated last bin.
    b = oracle.Binned(col(0.5, 1.5, 2.5, np.nan))
    info = b.feature_info(0)
    assert info["missing"] == 2 and info["n_bins"] == 4 and info["default_bin"] == 3
    assert b.value_to_bin(0, float("nan")) == 3


def test_categorical_value_to_bin(oracle):
    # bin_mapper.rs:448-459: categories 10,20,30 -> bins 0,1,2; unknown -> 0.
    b = oracle.Binned(col(10, 20, 30, 20), is_cat=np.array([1], np.uint8))
    for v, e in [(10, 0), (20, 1), (30, 2), (99, 0)]:
        assert b.value_to_bin(0, v) == e
    assert b.feature_info(0)["categorical"]


def test_trivial_and_constant_features(oracle):
    # feature_analysis.rs:355-357: all-zero / all-NaN are trivial; a constant
    # non-zero feature is kept with a single bin (dataset.rs:697-699).
    X = np.stack([np.zeros(10, F32), np.full(10, np.nan, F32), np.full(10, 2.0, F32),
                  np.arange(10, dtype=F32)])
    b = oracle.Binned(X)
    assert b.feature_info(0)["trivial"] and b.feature_info(1)["trivial"]
    assert not b.feature_info(2)["trivial"] and b.feature_info(2)["n_bins"] == 1
    lay = b.layout()
    assert list(lay["eff_to_orig"]) == [2, 3]


def test_edges_closed_form(oracle):
    # bin_numeric, dataset.rs:683-750: midpoints of sorted uniques when few,
    # quantile midpoints (strictly increasing) when many.
    rng = np.random.default_rng(0)
    few = rng.integers(0, 40, 1000).astype(F32)
    u = np.unique(few).astype(np.float64)
    b = oracle.Binned(few[None])
    np.testing.assert_array_equal(b.bounds(0)[:-1], (u[:-1] + u[1:]) / 2)
    many = rng.standard_normal(5000).astype(F32)
    u = np.unique(many).astype(np.float64)
    nb = 256
    exp = []
    for i in range(1, nb):
        j = min(int(math.floor(i / nb * (len(u) - 1))), len(u) - 2)
        e = (u[j] + u[j + 1]) / 2
        if not exp or e > exp[-1]:
            exp.append(e)
    b = oracle.Binned(many[None])
    np.testing.assert_array_equal(b.bounds(0)[:-1], exp)
    # value_to_bin == number of bounds strictly below v (first bound >= v).
    bins = b.bins()[0]
    ref = np.minimum(np.searchsorted(np.array(exp), many.astype(np.float64), side="left"), len(exp))
    np.testing.assert_array_equal(bins, ref)


def test_sample_for_edges_deterministic(oracle):
    # dataset.rs:598-629: >200k rows -> edges from a 200k splitmix64 sample;
    # the same data must bin identically run to run.
    rng = np.random.default_rng(1)
    X = rng.standard_normal((1, 200_500)).astype(F32)
    a = oracle.Binned(X).bounds(0)
    b = oracle.Binned(X.copy()).bounds(0)
    np.testing.assert_array_equal(a, b)
    full = oracle.Binned(X[:, :200_000]).bounds(0)
    assert not np.array_equal(a, full)  # the sample is not the first 200k rows


# ------------------------------------------------------------- objectives ---
def test_logistic_gradient_at_zero(oracle):
    # classification.rs:248-266
    g, h = oracle.gradients(1, np.zeros((1, 1), F32), np.ones(1, F32))
    assert g[0, 0] == pytest.approx(-0.5, abs=1e-5) and h[0, 0] == pytest.approx(0.25, abs=1e-5)


def test_logistic_weighted(oracle):
    # classification.rs:281-302
    g, _ = oracle.gradients(1, np.zeros((1, 2), F32), np.array([1, 0], F32), np.array([2.0, 0.5], F32))
    assert g[0, 0] == pytest.approx(-1.0, abs=1e-5) and g[0, 1] == pytest.approx(0.25, abs=1e-5)


def test_base_scores(oracle):
    # classification.rs:268-279 (50% positive -> 0), :350-374 (class priors).
    assert oracle.base_score(1, np.array([0, 0, 1, 1], F32))[0] == pytest.approx(0.0, abs=1e-5)
    s = oracle.base_score(2, np.array([0, 0, 1, 2], F32), n_outputs=3)
    assert s[0] > s[1] and s[1] == pytest.approx(s[2], abs=1e-5)
    assert oracle.base_score(0, np.array([1, 2, 3, 6], F32))[0] == pytest.approx(3.0)


def test_softmax_gradient_signs(oracle):
    # classification.rs:323-348
    preds = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]], F32)
    g, h = oracle.gradients(2, preds, np.array([0, 1], F32))
    assert g[0, 0] < 0 and g[1, 0] > 0 and (h > 0).all()
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-6)


def test_squared_gradient(oracle):
    # regression.rs:20-42: g = w (p - y), h = w.
    g, h = oracle.gradients(0, np.array([[1.0, 2.0]], F32), np.array([0.5, 3.0], F32), np.array([2.0, 1.0], F32))
    np.testing.assert_array_equal(g, [[1.0, -1.0]])
    np.testing.assert_array_equal(h, [[2.0, 1.0]])


# ------------------------------------------------------------- histograms ---
def test_histogram_basic(oracle):
    # ops.rs:418-441: bins [0,1,0,2,1,0] -> hist[0] = (10, 5), hist[1].g = 7, hist[2].g = 4.
    b = oracle.Binned(col(0, 1, 0, 2, 1, 0))
    np.testing.assert_array_equal(b.bins()[0], [0, 1, 0, 2, 1, 0])
    h = b.histogram(np.array([1, 2, 3, 4, 5, 6], F32), np.array([0.5, 1, 1.5, 2, 2.5, 3], F32))
    np.testing.assert_allclose(h[0], [10, 5])
    assert h[1, 0] == 7 and h[2, 0] == 4


def test_histogram_gathered(oracle):
    # ops.rs:443-474: rows [0,2,4] of bins [0,1,2,0,1,2].
    b = oracle.Binned(col(0, 1, 2, 0, 1, 2))
    h = b.histogram(np.arange(1, 7, dtype=F32), np.ones(6, F32), rows=np.array([0, 2, 4], np.uint32))
    np.testing.assert_allclose(h[:, 0], [1, 5, 3])


def test_histogram_brute_force_and_invariants(oracle):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((5, 3000)).astype(F32)
    X[rng.random(X.shape) < 0.1] = np.nan
    b = oracle.Binned(X)
    g = rng.standard_normal(3000).astype(F32)
    hh = rng.random(3000).astype(F32)
    rows = np.sort(rng.choice(3000, 1234, replace=False)).astype(np.uint32)
    hist = b.histogram(g, hh, rows)
    bins = b.bins()
    lay = b.layout()
    for e in range(lay["n_eff"]):
        f, off, nb = lay["eff_to_orig"][e], lay["offset"][e], lay["n_bins"][e]
        ref = np.zeros((nb, 2))
        np.add.at(ref[:, 0], bins[f, rows], g[rows].astype(np.float64))
        np.add.at(ref[:, 1], bins[f, rows], hh[rows].astype(np.float64))
        np.testing.assert_allclose(hist[off:off + nb], ref, rtol=1e-12, atol=1e-12)
        # Every feature's bins sum to the node totals.
        assert hist[off:off + nb, 0].sum() == pytest.approx(g[rows].astype(np.float64).sum(), rel=1e-12)


# ------------------------------------------------------------ split / gain --
def _hist_binned(oracle, n_feats, n_bins):
    X = np.stack([np.repeat(np.arange(n_bins, dtype=F32), 3)] * n_feats)
    return oracle.Binned(X)


def test_gain_symmetric(oracle):
    # gain.rs:181-196: (10,5 | -10,5 | 0,10) -> 16.666
    b = _hist_binned(oracle, 1, 2)
    s = b.find_split(np.array([[10, 5], [-10, 5]], float), 0.0, 10.0, 10, make_cfg())
    assert s.gain == pytest.approx(16.6667, abs=1e-3) and s.bin == 0


def test_leaf_weight(oracle):
    # gain.rs:211-229
    assert oracle.leaf_weight(make_cfg(), -10.0, 5.0) == pytest.approx(1.6667, abs=1e-3)
    c = make_cfg(alpha=2.0)
    assert oracle.leaf_weight(c, -1.0, 5.0) == 0.0
    assert oracle.leaf_weight(c, -10.0, 5.0) == pytest.approx(1.3333, abs=1e-3)


def test_valid_split_constraints(oracle):
    # gain.rs:198-209 via the scan: min_child_weight / min_samples_leaf reject.
    b = _hist_binned(oracle, 1, 2)
    h = np.array([[10, 4], [-10, 5]], float)
    assert b.find_split(h, 0.0, 9.0, 20, make_cfg(min_child_weight=5.0)).gain == -np.inf
    assert b.find_split(h, 0.0, 9.0, 20, make_cfg(min_child_weight=4.0)).gain > 0


def test_find_split_strongest_feature(oracle):
    # find.rs:762-801 and :820-866 (column filtering).
    b = _hist_binned(oracle, 3, 3)
    h = np.array([(1, 5), (0, 5), (-1, 5), (10, 5), (0, 5), (-10, 5), (3, 5), (0, 5), (-3, 5)], float)
    cfg = make_cfg()
    assert b.find_split(h, 0.0, 45.0, 45, cfg).feature == 1
    assert b.find_split(h, 0.0, 45.0, 45, cfg, features=[0, 2]).feature == 2
    assert b.find_split(h, 0.0, 45.0, 45, cfg, features=[0]).feature == 0


def test_single_bin_invalid(oracle):
    # find.rs:803-818: a constant feature (1 bin) gives no split.
    b = oracle.Binned(np.full((1, 6), 2.0, F32))
    s = b.find_split(np.array([[5, 5]], float), 5.0, 5.0, 5, make_cfg())
    assert s.gain == -np.inf


def test_categorical_onehot_and_sorted(oracle):
    # find.rs:712-760: one-hot picks category 1; sorted strategy is categorical.
    X = np.repeat(np.arange(3, dtype=F32), 5)[None]
    b = oracle.Binned(X, is_cat=np.array([1], np.uint8))
    s = b.find_split(np.array([[5, 5], [-10, 5], [5, 5]], float), 0.0, 15.0, 15, make_cfg(max_onehot_cats=4))
    assert s.is_categorical and (s.left_cats[0] >> 1) & 1 and s.left_cats[0] == 0b10
    X5 = np.repeat(np.arange(5, dtype=F32), 3)[None]
    b5 = oracle.Binned(X5, is_cat=np.array([1], np.uint8))
    h5 = np.array([[5, 3], [3, 3], [0, 3], [-3, 3], [-5, 3]], float)
    s5 = b5.find_split(h5, 0.0, 15.0, 15, make_cfg(max_onehot_cats=2))
    assert s5.is_categorical and s5.gain > 0
    # Ascending -G/(H+1e-6) puts cats 0,1,2,3,4 in order; {0,1} and {0,1,2}
    # tie (0.5*(64/7 + 64/10) each) and the first in scan order wins.
    assert s5.left_cats[0] == 0b11


def test_missing_direction_scan(oracle):
    # find.rs:333-404: forward (missing right) then backward (missing left).
    X = np.array([[0, 1, 2, np.nan] * 4], F32)
    b = oracle.Binned(X)
    assert b.feature_info(0)["n_bins"] == 4
    # NaN rows carry large negative gradients: they should go with bin 0 (left).
    h = np.array([[-4, 4], [3, 4], [3, 4], [-8, 4]], float)
    s = b.find_split(h, -6.0, 16.0, 16, make_cfg(min_child_weight=0.0))
    assert s.default_left == 1 and s.bin == 0


def test_split_brute_force(oracle):
    """Brute force of every candidate with the reference rules (ties: first in scan order)."""
    rng = np.random.default_rng(7)
    for trial in range(30):
        with_nan = bool(rng.integers(0, 2))
        nb = int(rng.integers(2, 9))
        data = np.array(list(range(nb)) * 3 + ([np.nan] * 3 if with_nan else []), F32)[None]
        b = oracle.Binned(data)
        nbins = b.feature_info(0)["n_bins"]
        assert nbins == nb + (1 if with_nan else 0)
        hist = np.stack([rng.normal(size=nbins) * 3, rng.random(nbins) * 4 + 0.5], axis=1)
        pg, ph = hist[:, 0].sum(), hist[:, 1].sum()
        cnt = int(rng.integers(nbins, 40))
        cfg = make_cfg(min_child_weight=0.5)
        lam = 1.0

        def gain(gl, hl):
            gr, hr = pg - gl, ph - hl
            return np.float32(0.5 * (gl * gl / (hl + lam) + gr * gr / (hr + lam)) - 0.5 * pg * pg / (ph + lam))

        # Candidates in the reference's scan order: forward bins ascending
        # (missing right, left count = bins scanned), then backward bins
        # descending (missing left, right count = bins scanned).
        cands = []
        for k in range(nbins - 1):
            cl = k + 1
            cands.append((k, 0, hist[: k + 1, 0].sum(), hist[: k + 1, 1].sum(), cl, max(cnt - cl, 0)))
        if with_nan:
            me = nbins - 2  # last real bin; the NaN bin stays left
            for j in range(me, 0, -1):
                gr, hr = hist[j:me + 1, 0].sum(), hist[j:me + 1, 1].sum()
                cr = me + 1 - j
                cands.append((j - 1, 1, pg - gr, ph - hr, max(cnt - cr, 0), cr))
        best = (-np.inf, None)
        for k, dl, gl, hl, cl, cr in cands:
            if hl >= 0.5 and ph - hl >= 0.5 and cl >= 1 and cr >= 1:
                gn = gain(gl, hl)
                if gn > best[0]:
                    best = (gn, (k, dl))
        s = b.find_split(hist, pg, ph, cnt, cfg)
        if best[1] is None:
            assert s.gain == -np.inf
        else:
            assert (s.bin, s.default_left) == best[1], trial
            assert s.gain == pytest.approx(float(best[0]), rel=1e-6, abs=1e-6)


def test_rayon_tie_rule(oracle):
    # find.rs:235-269: >= 16 features in parallel keeps the RIGHTMOST maximum,
    # sequential (n_threads = 1) the first.
    nf = 16
    b = _hist_binned(oracle, nf, 2)
    h = np.tile(np.array([[4.0, 3.0], [-4.0, 3.0]]), (nf, 1))
    cfg = make_cfg()
    assert b.find_split(h, 0.0, 6.0, 6, cfg, n_threads=0).feature == nf - 1
    assert b.find_split(h, 0.0, 6.0, 6, cfg, n_threads=1).feature == 0


# --------------------------------------------------------------- partition --
def _part_binned(oracle):
    X = np.array([[0, 1, 0, 1, 0, 1, 0, 1], [0, 0, 0, 0, 1, 1, 1, 1]], F32)
    return oracle.Binned(X)


def _numeric_split(feature, bin_, dl=0):
    s = A.SplitInfoC()
    s.feature, s.bin, s.default_left, s.gain = feature, bin_, dl, 1.0
    return s


def test_partition_known(oracle):
    # partition.rs:430-505
    b = _part_binned(oracle)
    rows, lc = b.partition(np.arange(8), _numeric_split(1, 0))
    assert lc == 4 and list(rows) == [0, 1, 2, 3, 4, 5, 6, 7]
    leaf0 = rows[:4]
    rows, lc = b.partition(np.arange(8), _numeric_split(0, 0))
    assert lc == 4 and list(rows) == [0, 2, 4, 6, 1, 3, 5, 7]
    # test_multiple_splits: leaf 0 (rows 0-3) split on feature 0 -> {0,2} | {1,3}.
    r2, lc2 = b.partition(leaf0, _numeric_split(0, 0))
    assert lc2 == 2 and list(r2) == [0, 2, 1, 3]


def test_partition_stable_brute_force(oracle):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2, 500)).astype(F32)
    X[0, rng.random(500) < 0.2] = np.nan
    b = oracle.Binned(X)
    bins = b.bins()
    info = b.feature_info(0)
    for _ in range(20):
        rows = np.sort(rng.choice(500, 300, replace=False)).astype(np.uint32)
        k, dl = int(rng.integers(0, info["n_bins"] - 1)), int(rng.integers(0, 2))
        out, lc = b.partition(rows, _numeric_split(0, k, dl))
        bb = bins[0, rows]
        go = np.where(bb == info["default_bin"], bool(dl), bb <= k)
        np.testing.assert_array_equal(out, np.concatenate([rows[go], rows[~go]]))
        assert lc == go.sum()


# ------------------------------------------------------- growth / trees -----
def test_threshold_boundary(oracle):
    # grower.rs:813-842: bin <= 0 trains left; inference uses x < next_up(upper).
    X = np.array([[0.0, 1.0]], F32)
    cfg = make_cfg(n_trees=1, max_depth=1, min_child_weight=0.0, learning_rate=1.0)
    m = oracle.train(oracle.Binned(X), np.array([0.0, 10.0], F32), cfg)
    a = m.arrays
    thr = a["split_threshold"][0]
    assert thr == np.nextafter(F32(0.5), F32(np.inf))
    left, right = a["leaf_value"][a["left"][0]], a["leaf_value"][a["right"][0]]
    out = oracle.predict(m.info, a, np.array([[0.5, thr]], F32), raw=True)[0]
    base = a["base_score"][0]
    assert out[0] == pytest.approx(base + left) and out[1] == pytest.approx(base + right)


def test_categorical_domain_and_swap(oracle):
    # grower.rs:845-876: training-left categories {1,3} become the inference
    # "goes right" bitset 0b1010 with children swapped.
    X = np.array([[0, 1, 2, 3] * 5], F32)
    y = np.array([10, 0, 10, 0] * 5, F32)
    cfg = make_cfg(n_trees=1, max_depth=1, max_onehot_cats=2, learning_rate=1.0)
    m = oracle.train(oracle.Binned(X, is_cat=np.array([1], np.uint8)), y, cfg)
    a = m.arrays
    assert a["split_type"][0] == 1 and a["cat_size"][0] == 1 and a["cat_words"][a["cat_start"][0]] == 0b1010
    p = oracle.predict(m.info, a, np.array([[0, 1, 2, 3]], F32), raw=True)[0]
    assert p[1] == p[3] and p[0] == p[2] and p[1] < p[0]


def test_depthwise_leafwise_and_learning_rate(oracle):
    # expansion.rs:342-493 / grower.rs:1000+: depth cap, leaf cap, lr scaling.
    rng = np.random.default_rng(2)
    X = rng.standard_normal((3, 400)).astype(F32)
    y = (X[0] * 2 + X[1]).astype(F32)
    b = oracle.Binned(X)
    m = oracle.train(b, y, make_cfg(n_trees=1, max_depth=2))
    assert m.arrays["is_leaf"].sum() <= 4
    m = oracle.train(b, y, make_cfg(n_trees=1, growth=1, max_leaves=5))
    assert m.arrays["is_leaf"].sum() == 5
    m1 = oracle.train(b, y, make_cfg(n_trees=1, max_depth=3, learning_rate=1.0))
    m2 = oracle.train(b, y, make_cfg(n_trees=1, max_depth=3, learning_rate=0.5))
    lv = m1.arrays["is_leaf"] == 1
    np.testing.assert_allclose(m2.arrays["leaf_value"][lv], m1.arrays["leaf_value"][lv] * np.float32(0.5), rtol=1e-6)


def test_tree_consistent_with_rows(oracle):
    """Leaf values equal -G/(H+lambda)*lr from the rows that land in each leaf,
    and every split is the brute-force best over the node's own rows."""
    rng = np.random.default_rng(11)
    n = 600
    X = rng.standard_normal((3, n)).astype(F32)
    X[1, rng.random(n) < 0.15] = np.nan
    y = (X[0] + np.nan_to_num(X[1]) ** 2 + rng.normal(0, 0.1, n)).astype(F32)
    b = oracle.Binned(X)
    cfg = make_cfg(n_trees=1, max_depth=3, learning_rate=1.0)
    m = oracle.train(b, y, cfg, make_ext(0, 1))
    a = m.arrays
    base = y.astype(np.float64).mean()
    g = (F32(base) - y).astype(np.float64)
    _, leaf = oracle.predict(m.info, a, X, raw=True, leaf_idx=True)
    for node in np.unique(leaf[0]):
        sel = leaf[0] == node
        w = -g[sel].sum() / (sel.sum() + 1.0)
        assert a["leaf_value"][node] == pytest.approx(w, rel=1e-6, abs=1e-7)
    # Root split: brute force over every (feature, bin, direction).
    bins = b.bins()
    best = -np.inf
    pg, ph = g.sum(), float(n)
    for f in range(3):
        nb = b.feature_info(f)["n_bins"]
        miss = b.feature_info(f)["missing"] == 2
        for k in range(nb - 1):
            for dl in ([0, 1] if miss else [0]):
                go = bins[f] <= k
                if miss:
                    go = np.where(bins[f] == nb - 1, bool(dl), go)
                gl, hl = g[go].sum(), go.sum()
                if hl >= 1 and n - hl >= 1:
                    gn = 0.5 * (gl ** 2 / (hl + 1) + (pg - gl) ** 2 / (ph - hl + 1)) - 0.5 * pg ** 2 / (ph + 1)
                    best = max(best, gn)
    assert a["gain"][0] == pytest.approx(best, rel=1e-5)


# ------------------------------------------------------- prediction, etc. ---
def test_transforms(oracle):
    # transform.rs:140-200: sigmoid(0) = 0.5, NaN propagates, softmax rows sum to 1.
    info = A.ModelInfo()
    info.n_trees, info.n_groups, info.n_features, info.total_nodes = 1, 1, 1, 1
    info.total_cat_words, info.output_transform = 0, A.TRANSFORM_SIGMOID
    arrays = A.alloc_forest(info)
    arrays["is_leaf"][:] = 1
    arrays["tree_offset"][:] = [0, 1]
    out = oracle.predict(info, arrays, np.array([[0.0, np.nan]], F32))
    assert out[0, 0] == 0.5
    arrays["base_score"][0] = np.nan
    assert np.isnan(oracle.predict(info, arrays, np.array([[0.0]], F32))[0, 0])
    info.n_groups, info.output_transform = 3, A.TRANSFORM_SOFTMAX
    info.n_trees = 0
    arrays = A.alloc_forest(info)
    arrays["base_score"][:] = [1.0, 2.0, 3.0]
    out = oracle.predict(info, arrays, np.zeros((1, 4), F32))
    np.testing.assert_allclose(out.sum(axis=0), 1.0, rtol=1e-6)


def test_quality_smoke(oracle):
    # tests/quality_smoke.rs:182-201: loose quality bounds on synthetic data.
    from tests.conftest import binary_nan_data, regression_data

    X, y = regression_data(4000, 30, seed=0)
    m = oracle.train(oracle.Binned(X), y, make_cfg(n_trees=50, learning_rate=0.1))
    p = oracle.predict(m.info, m.arrays, X)
    assert np.sqrt(np.mean((p[0] - y) ** 2)) < 2.0
    X, y = binary_nan_data(4000, 30, seed=1)
    m = oracle.train(oracle.Binned(X), y, make_cfg(objective=1, n_trees=50, learning_rate=0.1))
    p = oracle.predict(m.info, m.arrays, X)
    assert oracle.metric(A.METRIC_LOGLOSS, p, y) < 0.65


def test_early_stopping(oracle):
    # callback.rs:78-118: patience counted in rounds since the best one.
    from tests.conftest import regression_data

    X, y = regression_data(800, 5, seed=3)
    Xv, yv = regression_data(800, 5, seed=4)
    cfg = make_cfg(n_trees=300, learning_rate=0.5, max_depth=6, early_stopping_rounds=5)
    m = oracle.train(oracle.Binned(X), y, cfg, valid_X=Xv, valid_y=yv)
    hist = m.metric_history()
    assert len(hist) < 300
    best = int(np.argmin(hist))
    assert len(hist) - 1 - best == 5
    assert m.info.n_trees == best + 1
