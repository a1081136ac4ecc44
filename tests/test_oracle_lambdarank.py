"""Pins for the oracle's LambdaRank gradients (CPU only).

The reference has no ranking objective (objectives/mod.rs:196-219), so there
are no reference vectors: the restatement is pinned by a hand-computed case,
by invariants of the pairwise formulation, and by the derivative identity that
defines it (gradient and hessian of the RankNet surrogate weighted by the
NDCG swap deltas at the current ranking).
"""
import math

import numpy as np
import pytest

T = 30  # pair truncation level


def _disc(r):
    return 1.0 / math.log2(2.0 + r)


def test_two_documents_hand_computed(oracle):
    oracle.set_exp_mode(0)
    p = np.array([0.0, 0.0], np.float32)
    y = np.array([1.0, 0.0], np.float32)
    g, h = oracle.lambdarank_gradients(p, y, [2])
    # maxDCG = (2^1 - 1) / log2(2) = 1; dNDCG = (2 - 1) * (1 - 1/log2(3)); rho = 1/2.
    dn = 1.0 - 1.0 / math.log2(3.0)
    assert g[0] == np.float32(-0.5 * dn)
    assert g[1] == np.float32(0.5 * dn)
    assert h[0] == np.float32(0.25 * dn) and h[1] == np.float32(0.25 * dn)


def test_ordering_and_sign(oracle):
    oracle.set_exp_mode(0)
    # Relevant document ranked second: it is pushed up, the other down.
    p = np.array([2.0, -1.0, 0.5], np.float32)
    y = np.array([0.0, 3.0, 1.0], np.float32)
    g, h = oracle.lambdarank_gradients(p, y, [3])
    assert g[1] < 0 and g[0] > 0
    assert (h > 0).all()
    assert abs(float(g.astype(np.float64).sum())) < 1e-6


def test_invariants_random_groups(oracle):
    oracle.set_exp_mode(0)
    rng = np.random.default_rng(5)
    sizes = rng.integers(1, 90, 60).astype(np.int32)
    n = int(sizes.sum())
    y = rng.integers(0, 5, n).astype(np.float32)
    p = rng.standard_normal(n).astype(np.float32)
    g, h = oracle.lambdarank_gradients(p, y, sizes)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for q in range(len(sizes)):
        gq = g[off[q]:off[q + 1]].astype(np.float64)
        assert abs(gq.sum()) <= 1e-5 * max(1.0, np.abs(gq).sum())
    assert (h >= 0).all()
    # All-equal labels (or a single document) give exactly zero.
    y2 = y.copy()
    y2[off[3]:off[4]] = 2.0
    g2, h2 = oracle.lambdarank_gradients(p, y2, sizes)
    assert not g2[off[3]:off[4]].any() and not h2[off[3]:off[4]].any()
    one = np.where(sizes == 1)[0]
    for q in one:
        assert g[off[q]] == 0 and h[off[q]] == 0
    # Weights scale each document's pair sums.
    w = rng.random(n).astype(np.float32) + 0.5
    gw, hw = oracle.lambdarank_gradients(p, y, sizes, w)
    np.testing.assert_allclose(gw, g.astype(np.float64) * w, rtol=2e-7, atol=1e-12)
    np.testing.assert_allclose(hw, h.astype(np.float64) * w, rtol=2e-7, atol=1e-12)


def test_permutation_equivariant(oracle):
    oracle.set_exp_mode(0)
    rng = np.random.default_rng(8)
    m = 70
    p = rng.standard_normal(m).astype(np.float32)  # distinct scores
    y = rng.integers(0, 4, m).astype(np.float32)
    g, h = oracle.lambdarank_gradients(p, y, [m])
    perm = rng.permutation(m)
    gp, hp = oracle.lambdarank_gradients(p[perm], y[perm], [m])
    np.testing.assert_array_equal(gp, g[perm])
    np.testing.assert_array_equal(hp, h[perm])


def test_truncation(oracle):
    oracle.set_exp_mode(0)
    # 40 documents with scores 40..1, so positions equal indices.  Documents
    # 31 and 32 sit below the top-30 cut: their pair with each other is not
    # counted, each pairs only with the 30 zero-label top documents, and the
    # documents after them (label 0, below the cut) get nothing.
    m = 40
    p = np.arange(m, 0, -1).astype(np.float32)
    y = np.zeros(m, np.float32)
    y[31] = 1.0
    y[32] = 2.0
    g, h = oracle.lambdarank_gradients(p, y, [m])
    assert g[31] < 0 and g[32] < 0
    assert (g[:30] > 0).all()
    assert not g[33:].any() and not h[33:].any()
    assert g[30] == 0.0  # position 30 is below the cut and has label 0
    # Without the truncation the 31-32 pair would add to both; check that
    # raising document 31 to label 2 (ties 32) changes 32 only through maxDCG:
    # its pair terms with the top 30 scale by the same factor.
    y2 = y.copy()
    y2[31] = 2.0
    g2, _ = oracle.lambdarank_gradients(p, y2, [m])
    ls1 = sorted(y.tolist(), reverse=True)
    ls2 = sorted(y2.tolist(), reverse=True)
    md1 = sum((2.0 ** ls1[r] - 1.0) * _disc(r) for r in range(T))
    md2 = sum((2.0 ** ls2[r] - 1.0) * _disc(r) for r in range(T))
    assert float(g2[32]) == pytest.approx(float(g[32]) * md1 / md2, rel=1e-6)


def _surrogate(s, y, order_pos, inv_maxdcg):
    """Sum over counted pairs of dNDCG * log(1 + exp(-(s_hi - s_lo))), dNDCG fixed."""
    s = [float(v) for v in s]
    y = [float(v) for v in y]
    m = len(s)
    tot = 0.0
    for a in range(m):
        for b in range(a + 1, m):
            if a >= T:
                continue
            i, j = order_pos[a], order_pos[b]
            if y[i] == y[j]:
                continue
            hi, lo = (i, j) if y[i] > y[j] else (j, i)
            dn = (2.0 ** y[hi] - 2.0 ** y[lo]) * abs(_disc(a) - _disc(b)) * inv_maxdcg
            tot += dn * math.log1p(math.exp(-(s[hi] - s[lo])))
    return tot


def test_gradient_is_derivative_of_pairwise_surrogate(oracle):
    oracle.set_exp_mode(0)
    rng = np.random.default_rng(11)
    m = 45
    s = rng.standard_normal(m).astype(np.float32)
    y = rng.integers(0, 4, m).astype(np.float32)
    g, h = oracle.lambdarank_gradients(s, y, [m])
    order = sorted(range(m), key=lambda i: (-float(s[i]), i))
    ls = sorted(y.tolist(), reverse=True)
    maxdcg = sum((2.0 ** ls[r] - 1.0) * _disc(r) for r in range(min(m, T)))
    s64 = s.astype(np.float64)
    eps = 1e-4
    for i in range(m):
        sp, sm = s64.copy(), s64.copy()
        sp[i] += eps
        sm[i] -= eps
        # dNDCG fixed at the base ranking (the lambda is the gradient of the
        # surrogate with the swap deltas held constant).
        fp = _surrogate(sp, y, order, 1.0 / maxdcg)
        fm = _surrogate(sm, y, order, 1.0 / maxdcg)
        f0 = _surrogate(s64, y, order, 1.0 / maxdcg)
        d1 = (fp - fm) / (2 * eps)
        d2 = (fp - 2 * f0 + fm) / (eps * eps)
        assert d1 == pytest.approx(float(g[i]), rel=1e-4, abs=1e-6)
        assert d2 == pytest.approx(float(h[i]), rel=2e-3, abs=1e-5)
