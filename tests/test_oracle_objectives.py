"""CPU: the oracle's remaining objectives, base scores, leaf-renewal quantiles
and metrics, pinned against closed forms, hand-computed answers and brute
force (objectives/regression.rs:68-333, classification.rs:76-123,
utils.rs:17-126, metrics/regression.rs:55-250, classification.rs:46-192,
trainer.rs:265-291).

Each check is written from the formula's definition in numpy (f32 where the
reference computes in f32, f64 where it accumulates in f64), never by calling
the oracle twice, so a dropped term, a swapped sign or an off-by-one quantile
index fails.
"""
import numpy as np
import pytest

from boosters_amd import _abi as A
from tests.conftest import make_cfg

pytestmark = []

f32 = np.float32


def _cfg(**kw):
    alphas = kw.pop("alphas", None)
    c = make_cfg(**kw)
    if alphas is not None:
        arr = np.ascontiguousarray(alphas, np.float32)
        c._alphas_keep = arr  # keep the buffer alive with the struct
        c.alphas = A.ptr(arr)
        c.n_alphas = len(arr)
    return c


# ---- gradients --------------------------------------------------------------

def test_absolute_gradients_sign_with_positive_zero(oracle):
    p = np.array([[1.0, 2.0, 3.0, -1.0]], f32)
    y = np.array([2.0, 2.0, 1.0, -1.0], f32)
    w = np.array([1.0, 2.0, 3.0, 0.5], f32)
    g, h = oracle.gradients_cfg(_cfg(objective=A.OBJ_ABSOLUTE), p, y, w)
    # diff = p - y = [-1, 0, 2, 0]; Rust signum(+0.0) = +1
    np.testing.assert_array_equal(g[0], [-1.0, 2.0, 3.0, 0.5])
    np.testing.assert_array_equal(h[0], w)


def test_pinball_gradients_per_alpha(oracle):
    p = np.array([[0.0, 1.0, 2.0], [0.0, 1.0, 2.0]], f32)
    y = np.array([1.0, 1.0, 1.0], f32)
    g, h = oracle.gradients_cfg(_cfg(objective=A.OBJ_PINBALL, alphas=[0.1, 0.9]), p, y)
    # pred < target -> -alpha, else 1 - alpha (f32)
    np.testing.assert_array_equal(g[0], f32([-0.1, f32(1) - f32(0.1), f32(1) - f32(0.1)]))
    np.testing.assert_array_equal(g[1], f32([-0.9, f32(1) - f32(0.9), f32(1) - f32(0.9)]))
    np.testing.assert_array_equal(h, np.ones((2, 3), f32))


def test_pseudo_huber_gradients_closed_form(oracle):
    rng = np.random.default_rng(3)
    p = rng.normal(0, 4, (1, 500)).astype(f32)
    y = rng.normal(0, 4, 500).astype(f32)
    w = rng.uniform(0.5, 2, 500).astype(f32)
    for delta in (0.5, 1.0, 3.0):
        g, h = oracle.gradients_cfg(_cfg(objective=A.OBJ_PSEUDO_HUBER, huber_delta=delta), p, y, w)
        r = p[0].astype(np.float64) - y.astype(np.float64)
        fac = 1.0 + (r / delta) ** 2
        np.testing.assert_allclose(g[0], w * r / np.sqrt(fac), rtol=3e-6, atol=1e-6)
        np.testing.assert_allclose(h[0], w / fac ** 1.5, rtol=3e-6, atol=1e-7)
    # limits: |r| << delta -> (r, 1); |r| >> delta -> (delta sign r, ~0)
    g, h = oracle.gradients_cfg(_cfg(objective=A.OBJ_PSEUDO_HUBER, huber_delta=1.0),
                                np.array([[1e-3, 1e4, -1e4]], f32), np.zeros(3, f32))
    np.testing.assert_allclose(g[0], [1e-3, 1.0, -1.0], rtol=1e-5)
    assert h[0][0] == pytest.approx(1.0, rel=1e-5) and h[0][1] < 1e-11


def test_poisson_gradients_with_clamp(oracle):
    p = np.array([[0.0, 1.0, 100.0, -100.0]], f32)
    y = np.array([1.0, 0.0, 2.0, 3.0], f32)
    w = np.array([1.0, 2.0, 1.0, 1.0], f32)
    oracle.set_exp_mode(1)
    g, h = oracle.gradients_cfg(_cfg(objective=A.OBJ_POISSON), p, y, w)
    e = np.exp(np.clip(p[0].astype(np.float64), -30, 30))
    np.testing.assert_allclose(g[0], w * (e - y), rtol=2e-7)
    # hessian w e^p floored at 1e-6: e^-30 ~ 9.4e-14 -> 1e-6
    np.testing.assert_allclose(h[0], np.maximum(w * e, 1e-6), rtol=2e-7)
    assert h[0][3] == f32(1e-6)


def test_hinge_gradients_labels_to_pm1(oracle):
    p = np.array([[0.5, 2.0, -0.5, -2.0, 1.0]], f32)
    y = np.array([1.0, 1.0, 0.0, 0.0, 1.0], f32)
    w = np.array([1.0, 1.0, 2.0, 1.0, 1.0], f32)
    g, h = oracle.gradients_cfg(_cfg(objective=A.OBJ_HINGE), p, y, w)
    # y -> [+1, +1, -1, -1, +1]; margin = y p = [0.5, 2, 0.5, 2, 1]; margin < 1 -> -w y
    np.testing.assert_array_equal(g[0], [-1.0, 0.0, 2.0, 0.0, 0.0])
    np.testing.assert_array_equal(h[0], w)


# ---- base scores and quantiles ----------------------------------------------

def test_weighted_quantile_known_answers(oracle):
    assert oracle.weighted_quantile([3, 1, 2], None, 0.5) == 2.0
    # total 12, target 6: running weights 1, 2, 12 -> value 3
    assert oracle.weighted_quantile([1, 2, 3], [1, 1, 10], 0.5) == 3.0
    assert oracle.weighted_quantile([1, 2, 3], [10, 1, 1], 0.5) == 1.0
    # alpha near 1: the last value
    assert oracle.weighted_quantile([5, 4, 9, 1], None, 0.99) == 9.0
    assert oracle.weighted_quantile([], None, 0.5) == 0.0


def test_base_scores(oracle):
    rng = np.random.default_rng(5)
    y = rng.normal(3, 2, 1001).astype(f32)
    med = oracle.base_score_cfg(_cfg(objective=A.OBJ_ABSOLUTE), y)
    assert med[0] == np.sort(y)[500]  # n odd, alpha 0.5: running count 501 >= 500.5
    assert oracle.base_score_cfg(_cfg(objective=A.OBJ_PSEUDO_HUBER), y)[0] == med[0]
    q = oracle.base_score_cfg(_cfg(objective=A.OBJ_PINBALL, alphas=[0.1, 0.9]), y, n_outputs=2)
    s = np.sort(y)
    # unit weights: first index with (i + 1) >= alpha * n (f32 products)
    for j, a in enumerate((0.1, 0.9)):
        tgt = f32(a) * f32(len(y))
        assert q[j] == s[int(np.argmax(np.arange(1, len(y) + 1, dtype=f32) >= tgt))]
    counts = rng.poisson(2.5, 2000).astype(f32)
    b = oracle.base_score_cfg(_cfg(objective=A.OBJ_POISSON), counts)[0]
    assert b == f32(np.log(counts.astype(np.float64).mean()))
    assert oracle.base_score_cfg(_cfg(objective=A.OBJ_POISSON), np.zeros(10, f32))[0] == f32(np.log(1e-7))
    assert oracle.base_score_cfg(_cfg(objective=A.OBJ_HINGE), counts)[0] == 0.0


def test_unweighted_leaf_quantile_index(oracle):
    v = np.array([10.0, 40.0, 20.0, 30.0], f32)
    ind = np.arange(4, dtype=np.uint32)
    # k = ceil(alpha n) - 1: 0.5 -> 1, 0.1 -> 0, 0.9 -> 3, 0.75 -> 2
    assert oracle.quantile_indexed(ind, v, 0.5) == 20.0
    assert oracle.quantile_indexed(ind, v, 0.1) == 10.0
    assert oracle.quantile_indexed(ind, v, 0.9) == 40.0
    assert oracle.quantile_indexed(ind, v, 0.75) == 30.0
    assert oracle.quantile_indexed(ind[:1], v, 0.3) == 10.0
    assert np.isnan(oracle.quantile_indexed(ind[:0], v, 0.5))


def test_leaf_quantiles_match_brute_force(oracle):
    rng = np.random.default_rng(9)
    v = rng.normal(0, 1, 3000).astype(f32)
    w = rng.uniform(0.1, 3, 3000).astype(f32)
    for trial in range(20):
        ind = rng.choice(3000, size=int(rng.integers(2, 400)), replace=False).astype(np.uint32)
        a = float(rng.uniform(0.05, 0.95))
        n = len(ind)
        k = min(max(int(np.ceil(f32(a) * f32(n))) - 1, 0), n - 1)
        assert oracle.quantile_indexed(ind, v, a) == np.sort(v[ind])[k]
        # weighted: total in the leaf's order, running sum in sorted order (f32)
        order = np.argsort(v[ind], kind="stable")
        tot = f32(0)
        for x in w[ind]:
            tot = f32(tot + x)
        thr = f32(tot * f32(a))
        cum = f32(0)
        want = v[ind][order[-1]]
        for j in order:
            cum = f32(cum + w[ind][j])
            if cum >= thr:
                want = v[ind][j]
                break
        assert oracle.quantile_indexed(ind, v, a, weights=w) == want


# ---- metrics ----------------------------------------------------------------

def test_regression_metrics_closed_forms(oracle):
    rng = np.random.default_rng(11)
    n = 4000
    y = rng.gamma(2.0, 2.0, n).astype(f32)
    p = (y + rng.normal(0, 1, n)).astype(f32)
    w = rng.uniform(0.2, 2, n).astype(f32)
    yd, pd_, wd = y.astype(np.float64), p.astype(np.float64), w.astype(np.float64)
    c = _cfg(alphas=[0.25], metric_huber_delta=1.5)
    assert oracle.metric_cfg(c, A.METRIC_MAE, p[None], y, w) == pytest.approx(
        np.sum(wd * np.abs(pd_ - yd)) / wd.sum(), rel=1e-12)
    assert oracle.metric_cfg(c, A.METRIC_MAPE, p[None], y, w) == pytest.approx(
        100 * np.sum(wd * np.abs(pd_ - yd) / np.maximum(np.abs(yd), 1e-15)) / wd.sum(), rel=1e-12)
    r = yd - pd_
    pin = np.where(r >= 0, 0.25 * r, 0.75 * -r)
    assert oracle.metric_cfg(c, A.METRIC_QUANTILE, p[None], y, w) == pytest.approx(
        np.sum(wd * pin) / wd.sum(), rel=1e-12)
    a = np.abs(pd_ - yd)
    hub = np.where(a <= 1.5, 0.5 * a * a, 1.5 * (a - 0.75))
    assert oracle.metric_cfg(c, A.METRIC_HUBER, p[None], y, w) == pytest.approx(
        np.sum(wd * hub) / wd.sum(), rel=1e-12)
    mu = np.maximum(pd_, 1e-9)
    dev = np.where(yd > 1e-9, 2 * (yd * np.log(yd / mu) - (yd - mu)), 2 * mu)
    assert oracle.metric_cfg(c, A.METRIC_POISSON_DEVIANCE, p[None], y, w) == pytest.approx(
        np.sum(wd * dev) / wd.sum(), rel=1e-12)


def test_quantile_metric_pools_outputs(oracle):
    y = np.array([1.0, 2.0, 3.0], f32)
    p = np.array([[0.0, 2.0, 4.0], [1.0, 1.0, 1.0]], f32)
    c = _cfg(alphas=[0.1, 0.9])
    # output 0 residuals y - p = [1, 0, -1] -> a0, 0, 1 - a0 ; output 1: [0, 1, 2] -> 0, a1, 2 a1
    # (alphas are f32 in the config, widened to f64)
    a1 = float(f32(0.9))
    assert oracle.metric_cfg(c, A.METRIC_QUANTILE, p, y) == pytest.approx((1.0 + 3 * a1) / 6, rel=1e-12)


def test_accuracy_metrics(oracle):
    y = np.array([1, 0, 1, 0, 1], f32)
    prob = np.array([[0.9, 0.2, 0.6, 0.7, 0.55]], f32)
    w = np.array([1, 1, 1, 2, 1], f32)
    c = _cfg(accuracy_threshold=0.6)
    # threshold 0.6: classes [1, 0, 1, 1, 0]; correct [1, 1, 1, 0, 0]
    assert oracle.metric_cfg(c, A.METRIC_ACCURACY, prob, y, w) == pytest.approx(3 / 6)
    margin = np.array([[0.0, -0.1, 2.0, 0.3, -1.0]], f32)
    # threshold 0 on raw scores: [1, 0, 1, 1, 0]
    assert oracle.metric_cfg(c, A.METRIC_MARGIN_ACCURACY, margin, y, w) == pytest.approx(3 / 6)
    probs = np.array([[0.2, 0.5, 0.1], [0.5, 0.2, 0.1], [0.3, 0.3, 0.8]], f32)
    lab = np.array([1, 0, 2], f32)
    # argmax with the last of equal maxima: row 0 -> class 1; row 1 -> classes 0/1 tie on 0.5?
    # columns: sample 0 = [0.2, 0.5, 0.3] -> 1; sample 1 = [0.5, 0.2, 0.3] -> 0; sample 2 -> 2
    assert oracle.metric_cfg(c, A.METRIC_MULTICLASS_ACCURACY, probs, lab) == 1.0
    tie = np.array([[0.4, 0.1], [0.4, 0.9]], f32)
    # sample 0 ties 0.4 / 0.4 -> the last maximum, class 1
    assert oracle.metric_cfg(c, A.METRIC_MULTICLASS_ACCURACY, tie, np.array([1, 1], f32)) == 1.0
    assert oracle.metric_cfg(c, A.METRIC_MULTICLASS_ACCURACY, tie, np.array([0, 1], f32)) == 0.5


def test_default_metrics(oracle):
    assert oracle.default_metric(A.OBJ_ABSOLUTE) == A.METRIC_RMSE
    assert oracle.default_metric(A.OBJ_PSEUDO_HUBER) == A.METRIC_RMSE
    assert oracle.default_metric(A.OBJ_POISSON) == A.METRIC_RMSE
    assert oracle.default_metric(A.OBJ_PINBALL) == A.METRIC_QUANTILE
    assert oracle.default_metric(A.OBJ_HINGE) == A.METRIC_MARGIN_ACCURACY


# ---- training with leaf renewal ---------------------------------------------

def _leaf_ids(oracle, model, X):
    _, leaf = oracle.predict(model.info, model.arrays, X, leaf_idx=True)
    return leaf


@pytest.mark.parametrize("objective,alphas", [(A.OBJ_ABSOLUTE, None), (A.OBJ_PINBALL, [0.2, 0.8])])
def test_renewed_leaves_are_residual_quantiles(oracle, objective, alphas):
    """After round 1 every leaf holds lr x the alpha-quantile of (y - base)
    over the rows that reach it (trainer.rs:265-291); NaN-free data, so the
    partition's routing and traversal agree."""
    rng = np.random.default_rng(13)
    n, F = 3000, 6
    X = rng.normal(0, 1, (F, n)).astype(f32)
    y = (X[0] * 2 + rng.standard_t(2, n)).astype(f32)
    cfg = _cfg(objective=objective, alphas=alphas, n_trees=1, max_depth=3, learning_rate=0.5)
    m = oracle.train(oracle.Binned(X), y, cfg, None)
    K = 1 if alphas is None else len(alphas)
    assert m.info.n_trees == K
    base = oracle.base_score_cfg(cfg, y, n_outputs=K)
    leaf = _leaf_ids(oracle, m, X)
    off = m.arrays["tree_offset"]
    for k in range(K):
        a = 0.5 if alphas is None else alphas[k]
        res = (y - base[k]).astype(f32)
        lv = m.arrays["leaf_value"][off[k]:off[k + 1]]
        for node in np.unique(leaf[k]):
            rows = res[leaf[k] == node]
            kk = min(max(int(np.ceil(f32(a) * f32(len(rows)))) - 1, 0), len(rows) - 1)
            assert lv[node] == f32(np.sort(rows)[kk]) * f32(0.5)


def test_quantile_regression_orders_outputs(oracle):
    rng = np.random.default_rng(17)
    n, F = 4000, 5
    X = rng.normal(0, 1, (F, n)).astype(f32)
    y = (X[0] + rng.normal(0, 1, n)).astype(f32)
    cfg = _cfg(objective=A.OBJ_PINBALL, alphas=[0.1, 0.5, 0.9], n_trees=20, max_depth=4)
    m = oracle.train(oracle.Binned(X), y, cfg, None)
    p = oracle.predict(m.info, m.arrays, X)
    cover = [(y <= p[k]).mean() for k in range(3)]
    assert cover[0] < cover[1] < cover[2]
    assert abs(cover[0] - 0.1) < 0.05 and abs(cover[2] - 0.9) < 0.05
    hist = m.metric_history()
    assert hist[-1] < hist[0]


@pytest.mark.parametrize("objective", [A.OBJ_POISSON, A.OBJ_PSEUDO_HUBER, A.OBJ_HINGE])
def test_other_objectives_train(oracle, objective):
    rng = np.random.default_rng(19)
    n, F = 3000, 5
    X = rng.normal(0, 1, (F, n)).astype(f32)
    if objective == A.OBJ_POISSON:
        y = rng.poisson(np.exp(0.5 * X[0])).astype(f32)
    elif objective == A.OBJ_HINGE:
        y = (X[0] + 0.3 * rng.normal(0, 1, n) > 0).astype(f32)
    else:
        y = (X[0] * 3 + rng.standard_t(1.5, n)).astype(f32)
    cfg = _cfg(objective=objective, n_trees=15, max_depth=4)
    oracle.set_exp_mode(1)
    m = oracle.train(oracle.Binned(X), y, cfg, None)
    hist = m.metric_history()
    if objective == A.OBJ_HINGE:
        assert hist[-1] > 0.8 and hist[-1] >= hist[0]  # margin accuracy, higher is better
    elif objective == A.OBJ_POISSON:
        # The default metric (RMSE) reads the raw log-space scores, as the
        # reference's does; the Poisson loss itself must fall.
        p = oracle.predict(m.info, m.arrays, X, raw=True)[0].astype(np.float64)
        base = float(oracle.base_score_cfg(cfg, y)[0])
        nll = lambda s: float(np.mean(np.exp(s) - y * s))  # noqa: E731
        assert nll(p) < nll(np.full(n, base))
    else:
        assert hist[-1] < hist[0]
